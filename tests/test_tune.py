"""Host logic of the skin autotuner (chemhmc_amd/tune.py) on a stand-in engine: the density-scaled skin
family, and the paired protocol (every candidate from the same positions, trajectory counter and seeds,
lanes set per candidate, candidates the list MD rejects skipped, the fastest kept, the state restored),
and the cache of tuned settings."""
import json

import numpy as np
import pytest

from chemhmc_amd import Chain, ChmcError, tune


def box(N, rho):
    return (N / rho) ** (1.0 / 3.0)


CHAINS = [Chain(N=512, L=box(512, 0.3), cutoff=2.5), Chain(N=512, L=box(512, 0.6), cutoff=2.5),
          Chain(N=512, L=box(512, 0.9), cutoff=2.5)]


def test_density_scaled_skins():
    assert np.allclose(tune.density_scaled_skins(CHAINS, 0.45, 0.0), 0.45)
    s = tune.density_scaled_skins(CHAINS, 0.4, 1.0)
    assert np.allclose(s, [0.8, 0.4, 0.4 * 0.6 / 0.9])          # base * rho_ref / rho
    s = tune.density_scaled_skins(CHAINS, 0.4, 1.5)
    assert np.allclose(s, [0.4 * 2.0, 0.4, 0.4 * (2 / 3) ** 1.5])   # 2^1.5 clamped to the factor bound 2
    s = tune.density_scaled_skins(CHAINS, 1.0, 3.0)               # factor clamped to [0.4, 2], skin to 2
    assert np.allclose(s, [2.0, 1.0, 0.4])
    anis = [Chain(N=300, L=(8.0, 6.0, 4.5), cutoff=2.5)]
    assert np.allclose(tune.density_scaled_skins(anis, 0.5, 1.0), 0.5 * 0.6 / (300 / (8.0 * 6.0 * 4.5)))


class FakeEngine:
    """Records what the tuner does; a run's cost is a function of the skins and lanes in force."""

    def __init__(self, chains, cost, reject_max=None):
        self.chains, self.nchains = chains, len(chains)
        self.cost, self.reject_max = cost, reject_max
        self.x = [np.full((c.N, 3), float(k)) for k, c in enumerate(chains)]
        self.traj, self.skins, self.lanes, self.check = 7, None, 8, 2
        self.log = []

    def positions(self, k):
        return self.x[k].copy()

    def set_positions(self, x, k):
        self.x[k] = np.array(x)

    def trajectory(self):
        return self.traj

    def set_trajectory(self, t):
        self.traj = t

    def set_chain_skins(self, skins):
        if self.reject_max is not None and max(skins) > self.reject_max:
            self.skins = None
            raise ChmcError("list cells too large")
        self.skins = np.array(skins)

    def set_list_lanes(self, lanes):
        self.lanes = lanes if lanes else (4 if self.skins.max() < 0.5 else 8)

    def list_lanes(self):
        return self.lanes

    def set_list_check(self, m):
        self.check = m

    def list_check(self):
        return self.check

    def speculation(self):
        return 0

    def list_entry_bits(self):
        return 32


def test_tune_skins_paired_protocol(monkeypatch):
    clock = [0.0]
    monkeypatch.setattr(tune.time, "perf_counter", lambda: clock[0])

    def cost(skins, lanes):                       # ms per trajectory: smallest at max skin 0.45, 8 lanes
        return 1.0 + abs(skins.max() - 0.45) + (0.5 if lanes != 8 else 0.0)

    e = FakeEngine(CHAINS, cost, reject_max=1.5)
    x0 = [e.positions(k) for k in range(3)]

    def run(n, seed):
        # every candidate starts from the saved state with the same seeds
        assert e.traj == 7 and all(np.array_equal(e.x[k], x0[k]) for k in range(3))
        e.log.append((seed, n, tuple(e.skins), e.lanes))
        clock[0] += n * cost(e.skins, e.lanes) * 1e-3        # seconds
        e.traj += n
        e.x = [x + 1.0 for x in e.x]

    skins, table = tune.tune_skins(e, run, bases=(0.3, 0.45, 0.6, 0.9), exponents=(0.0, 1.0), reps=5, lanes=0)
    # (0.9, e=1) gives 1.8 at rho = 0.3: rejected by the list MD and skipped; every other candidate timed
    assert set(table) == {"0.3/0.0", "0.45/0.0", "0.6/0.0", "0.9/0.0", "0.3/1.0", "0.45/1.0", "0.6/1.0"}
    assert {s for s, _, _, _ in e.log} == {1001} and {n for _, n, _, _ in e.log} == {5}
    # lanes tuned per candidate: 4 below skin 0.5 (the stand-in's rule), so 0.3/0 runs at 4 lanes
    assert table["0.3/0.0"] == pytest.approx(1.0 + 0.15 + 0.5)
    assert table["0.45/0.0"] == pytest.approx(1.0 + 0.5)          # max skin 0.45 -> 4 lanes
    best = min(table, key=table.get)
    assert np.allclose(skins, tune.density_scaled_skins(CHAINS, float(best.split("/")[0]), float(best.split("/")[1])))
    assert np.allclose(e.skins, skins) and e.lanes == (4 if skins.max() < 0.5 else 8)
    assert e.traj == 7 and all(np.array_equal(e.x[k], x0[k]) for k in range(3))   # state restored


def test_tune_skins_warm_and_no_candidate():
    e = FakeEngine(CHAINS, None, reject_max=0.1)
    with pytest.raises(ChmcError):
        tune.tune_skins(e, lambda n, s: None, bases=(0.3,), exponents=(0.0,))
    e = FakeEngine(CHAINS, None)
    calls = []
    tune.tune_skins(e, lambda n, s: calls.append((n, s)), bases=(0.3, 0.4), exponents=(0.0,), warm=3, reps=4, seed=50)
    assert calls == [(3, 50), (4, 51), (3, 50), (4, 51)]


def test_cached_tune_skins(tmp_path, monkeypatch):
    """A miss runs the search and stores its outcome; a later engine of the same problem gets the stored
    skins, check interval and lanes without a single run; another workload or candidate set misses; a
    damaged cache file reads as empty."""
    clock = [0.0]
    monkeypatch.setattr(tune.time, "perf_counter", lambda: clock[0])
    runs = []

    def make_run(e):
        def run(n, seed):                        # ms per trajectory: smallest at skin 0.4, check 1
            runs.append((n, seed, tuple(e.skins), e.check))
            clock[0] += n * (1.0 + abs(e.skins.max() - 0.4) + 0.1 * e.check) * 1e-3
        return run

    cache = str(tmp_path / "sub" / "tune.json")
    kw = dict(bases=(0.3, 0.4, 0.5), exponents=(0.0,), reps=2, lanes=0, checks=(1, 2))
    e = FakeEngine(CHAINS, None)
    skins, table, hit = tune.cached_tune_skins(e, make_run(e), cache, {"workload": "w"}, **kw)
    assert not hit and len(runs) == 6 and len(table) == 6
    assert np.allclose(skins, 0.4) and e.check == 1 and e.lanes == 4
    e2 = FakeEngine(CHAINS, None)
    e2.check, e2.lanes = 3, 16
    runs.clear()
    skins2, table2, hit2 = tune.cached_tune_skins(e2, make_run(e2), cache, {"workload": "w"}, **kw)
    assert hit2 and not runs and table2 == table
    assert np.allclose(skins2, skins) and np.allclose(e2.skins, skins) and e2.check == 1 and e2.lanes == 4
    tune.cached_tune_skins(e2, make_run(e2), cache, {"workload": "v"}, **kw)
    assert len(runs) == 6
    runs.clear()
    tune.cached_tune_skins(e2, make_run(e2), cache, {"workload": "w"}, **dict(kw, reps=3))
    assert len(runs) == 6
    with open(cache) as f:
        assert len(json.load(f)) == 3
    e3 = FakeEngine(CHAINS[:2], None)                # other chains: a miss
    runs.clear()
    assert not tune.cached_tune_skins(e3, make_run(e3), cache, {"workload": "w"}, **kw)[2] and len(runs) == 6
    with open(cache, "w") as f:
        f.write("{not json")
    runs.clear()
    assert not tune.cached_tune_skins(e2, make_run(e2), cache, {"workload": "w"}, **kw)[2] and len(runs) == 6
