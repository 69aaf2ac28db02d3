"""The learned surrogate on the GPU against the FP64 numpy oracle (oracle/surrogate.py), which is
pinned to torch autograd and finite differences in test_surrogate_oracle.py.

System: N=512, rho*=0.8, T*=1.5, r_c=2.5 (the plain-HMC CPU case), decorrelated by LJ HMC first.
Tolerances (FP32 pair sums with __expf/__cosf against FP64; bfloat16 MFMA layers against the oracle's
bfloat16 emulation, which differ only where an FP32 rounding difference flips a bfloat16 rounding):
  descriptors        |G_gpu - G| <= 2e-4 + 1e-4 |G|
  energy             |E_gpu - E| <= 1e-3 sum_i |e_i|                 (oracle fed the GPU's G)
  dE/dG              max |g_gpu - g| <= 1e-2 max |g|, mean <= 1e-4 max |g|
  forces             rms |f_gpu - f| <= 2e-3 rms |f|                 (oracle from the positions)
  LJ targets         |u_gpu - u| <= 5e-5 except <= 4 atoms with a pair at r_c (<= 0.02)
  loss gradient      per tensor max |d_gpu - d| <= 2e-3 max |d|
  Adam step 1        |dp_gpu - dp| <= 1e-6 + 1e-3 lr
Sampling: HMC driven by surrogate forces keeps the exact FP64 LJ Hamiltonian in the accept test, so
<U/N> of a surrogate-driven chain must agree with an LJ-driven one (|z| < 4, tau_int-corrected).
"""
import time

import numpy as np
import pytest

from chemhmc_amd import stats

pytestmark = pytest.mark.gpu

N, RHO, T, RC = 512, 0.8, 1.5, 2.5
L = (N / RHO) ** (1 / 3)
DT, STEPS = 0.005, 20


@pytest.fixture(scope="module")
def osur():
    from oracle import surrogate
    return surrogate


def _engine(chmc, seed=5, burn=30):
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T))
    e.init_lattice(seed, 0.05)
    if burn:   # accept-all burn-in away from the lattice (thermalization_steps)
        e.hmc(burn, burn + 1, chmc.LEAPFROG, DT, STEPS, seed=seed + 11)
    return e


def _params(osur, seed, G=None):
    p = osur.init_params(seed)
    rng = np.random.default_rng(seed + 100)
    p["b1"] = rng.normal(0, 0.1, 64)
    p["b2"] = rng.normal(0, 0.1, 64)
    p["b3"] = np.array([0.3])
    if G is not None:
        p["shift"] = G.mean(0)
        p["scale"] = 1.0 / (G.std(0) + 1e-3)
    return p


def _as_f32(p):
    """The parameters as the device holds them (FP32 masters)."""
    return {k: np.asarray(v, np.float32).astype(np.float64) for k, v in p.items()}


@pytest.fixture(scope="module")
def system(chmc, osur):
    e = _engine(chmc)
    x = e.positions()
    e.surrogate_set_params(_params(osur, 1))
    e.surrogate_eval()
    G, _ = e.surrogate_features()
    p = _as_f32(_params(osur, 1, G.astype(np.float64)))
    e.surrogate_set_params(p)
    yield e, x, p
    e.close()


def test_params_roundtrip(chmc, system):
    e, _, p = system
    assert np.array_equal(e.surrogate_params(), chmc.pack_surrogate(p))


def test_descriptors(system, osur):
    e, x, p = system
    e.surrogate_eval()
    G, _ = e.surrogate_features()
    Go = osur.descriptors(x, L, RC)
    err = np.abs(G - Go)
    print(f"descriptor max abs err {err.max():.3e}, max |G| {np.abs(Go).max():.3f}")
    assert np.all(err <= 2e-4 + 1e-4 * np.abs(Go))


def test_energy_and_input_gradient(system, osur):
    e, x, p = system
    E = e.surrogate_eval()[0]
    G, g = e.surrogate_features()
    eo, cache = osur.mlp(G.astype(np.float64), p, bf=True)
    go = osur.input_grad(G.astype(np.float64), p, bf=True, cache=cache)
    print(f"E gpu {E:.6f} oracle {eo.sum():.6f} (sum|e| {np.abs(eo).sum():.3f})")
    assert abs(E - eo.sum()) <= 1e-3 * np.abs(eo).sum()
    d = np.abs(g - go)
    print(f"dE/dG max err {d.max():.3e} mean {d.mean():.3e} max |g| {np.abs(go).max():.3e}")
    assert d.max() <= 1e-2 * np.abs(go).max()
    assert d.mean() <= 1e-4 * np.abs(go).max()


def test_forces(system, osur):
    e, x, p = system
    e.surrogate_eval(store_force=True)
    f = e.force()
    Eo, fo, _ = osur.energy_grad(x, L, RC, p, bf=True)
    rms = np.sqrt((fo ** 2).sum(1).mean())
    err = np.sqrt(((f - fo) ** 2).sum(1).mean())
    print(f"force rms err {err:.3e} of rms {rms:.3e}; net force {np.abs(f.sum(0)).max():.3e}")
    assert err <= 2e-3 * rms
    assert np.abs(f.sum(0)).max() <= 1e-3 * rms * np.sqrt(N)     # translation invariance of E


def test_targets_and_loss_gradient(chmc, system, osur):
    e, x, p = system
    e.surrogate_reserve(2)
    rows = e.surrogate_collect()
    assert rows == N
    G, u = e.surrogate_training_data(0, N)
    perm = e.bins()[0]                          # training rows are in binned (slot) order
    uo = osur.lj_atom_energies(x, L, RC)[perm]
    du = np.abs(u - uo)
    # FP64 sums over FP32 pair distances (typical 1e-6); a pair within ~1e-6 of r_c can fall on the
    # other side of the cutoff in FP32 and moves u_i by |2 sr6 (sr6 - 1)| ~ 8e-3 (Q3)
    far = du > 5e-5
    print(f"targets: max err {du.max():.3e}, {far.sum()} atoms beyond 5e-5 ({np.sort(du)[-3:]})")
    assert far.sum() <= 4 and du.max() < 0.02
    assert abs(u.astype(np.float64).sum() - e.compute_potential()[0]) <= 1e-4 * abs(uo).sum()
    grad = e.surrogate_loss_grad(0, N)
    lo, go = osur.loss_grad(p, G.astype(np.float64), u.astype(np.float64), bf=True)
    gd = chmc.unpack_surrogate(grad, trained_only=True)
    for k, v in go.items():
        err = np.abs(gd[k] - v).max()
        print(f"d loss / d {k}: max err {err:.3e} of max {np.abs(v).max():.3e}")
        assert err <= 2e-3 * np.abs(v).max()
    assert abs(grad[-1] / N - lo) <= 1e-3 * lo


def test_adam_first_step(chmc, chmc_fresh_surrogate, osur):
    e, p = chmc_fresh_surrogate
    grad = e.surrogate_loss_grad(0, N)
    lr = 1e-3
    e.surrogate_train(1, N, lr=lr, seed=3)      # one batch of all N rows: offset 0
    new = chmc.unpack_surrogate(e.surrogate_params())
    gd = chmc.unpack_surrogate(grad, trained_only=True)
    q = {k: v.copy() for k, v in p.items()}
    osur.adam_step(q, gd, {}, lr=lr)
    for k in gd:
        err = np.abs(new[k] - q[k]).max()
        assert err <= 1e-6 + 1e-3 * lr, (k, err)
    assert np.array_equal(new["shift"], p["shift"]) and np.array_equal(new["scale"], p["scale"])


@pytest.fixture
def chmc_fresh_surrogate(chmc, osur):
    e = _engine(chmc, seed=9)
    e.surrogate_set_params(_params(osur, 2))
    e.surrogate_eval()
    G, _ = e.surrogate_features()
    p = _as_f32(_params(osur, 2, G.astype(np.float64)))
    e.surrogate_set_params(p)
    e.surrogate_reserve(1)
    e.surrogate_collect()
    yield e, p
    e.close()


def _train(chmc, e, osur, nconf=40, nsteps=3000, seed=21):
    """Training set from an LJ-driven chain, normalised descriptors, Adam on the GPU."""
    e.surrogate_set_params(_params(osur, seed))
    e.surrogate_reserve(nconf)
    for k in range(nconf):
        e.hmc(5, 0, chmc.LEAPFROG, DT, STEPS, seed=seed * 1000 + k)
        e.surrogate_collect()
    G, _ = e.surrogate_training_data(0, nconf * N)
    p = _params(osur, seed, G.astype(np.float64))
    e.surrogate_set_params(p)
    return e.surrogate_train(nsteps, 2048, lr=2e-3, seed=seed)


@pytest.fixture(scope="module")
def trained(chmc, osur):
    e = _engine(chmc, seed=13, burn=100)
    hist = _train(chmc, e, osur)
    yield e, hist
    e.close()


def test_training_reduces_loss(trained):
    _, hist = trained
    first, last = hist[:20].mean(), hist[-50:].mean()
    print(f"training MSE per atom: first {first:.4e} last {last:.4e}")
    assert np.all(np.isfinite(hist))
    assert last < 0.2 * first


def test_surrogate_hmc_graph_matches_direct(chmc, trained):
    """Bit-identical records from graph replays and direct launches.  Both runs start from the same
    positions set through the API, so both rebuild their neighbour lists there (the list contents, and
    with them the FP32 summation order, depend on where the lists were last built)."""
    t, _ = trained
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T), atom_id_order=True)
    e.init_lattice(13, 0.05)
    e.hmc(30, 31, chmc.LEAPFROG, DT, STEPS, seed=24)
    e.surrogate_set_params(t.surrogate_params())
    x0, t0 = e.positions(), e.trajectory()
    e.set_positions(x0)
    r1, _ = e.hmc(6, 0, chmc.LEAPFROG, DT, STEPS, seed=77, use_graph=True, surrogate=True)
    e.set_positions(x0)
    e.set_trajectory(t0)                        # same heat-bath counters
    r2, _ = e.hmc(6, 0, chmc.LEAPFROG, DT, STEPS, seed=77, use_graph=False, surrogate=True)
    e.close()
    assert np.all(r1["surrogate"] == 1)
    assert np.array_equal(r1, r2)


def test_time_surrogate_leaves_the_state(chmc, trained):
    """chmc_time_surrogate times the descriptor + MLP launch and the surrogate force sweep at dt = 0:
    positive times, positions and momenta unchanged bitwise, and a surrogate run after it equal to one
    from the same positions and counters set through the API (both on lists built at those positions);
    without surrogate parameters it is an error."""
    t, _ = trained
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T), atom_id_order=True)
    e.init_lattice(13, 0.05)
    with pytest.raises(chmc.ChmcError):
        e.time_surrogate(2)
    e.hmc(30, 31, chmc.LEAPFROG, DT, STEPS, seed=24)
    e.surrogate_set_params(t.surrogate_params())
    x0, p0, t0 = e.positions(), e.momenta(), e.trajectory()
    ms_desc, ms_sweep = e.time_surrogate(5)
    assert ms_desc > 0 and ms_sweep > 0
    assert np.array_equal(e.positions(), x0) and np.array_equal(e.momenta(), p0) and e.trajectory() == t0
    r1, _ = e.hmc(4, 0, chmc.LEAPFROG, DT, STEPS, seed=79, surrogate=True)
    e.set_positions(x0)
    e.set_trajectory(t0)
    r2, _ = e.hmc(4, 0, chmc.LEAPFROG, DT, STEPS, seed=79, surrogate=True)
    e.close()
    assert np.all(r1["surrogate"] == 1)
    assert np.array_equal(r1, r2)


def test_fallback_switches_to_lj(chmc, trained):
    e, _ = trained
    r, _ = e.hmc(25, 0, chmc.LEAPFROG, DT, STEPS, seed=78, surrogate=True, fallback_window=10, fallback_acc=1.01)
    assert np.all(r["surrogate"][:10] == 1) and np.all(r["surrogate"][10:] == 0)
    r, _ = e.hmc(25, 0, chmc.LEAPFROG, DT, STEPS, seed=79, surrogate=True, fallback_window=10, fallback_acc=0.0)
    assert np.all(r["surrogate"] == 1)


@pytest.mark.parametrize("fallback_acc,bits", [(0.0, 0), (1.01, 0), (0.0, 16)])
def test_speculative_surrogate_equals_sequential(chmc, trained, fallback_acc, bits):
    """Surrogate-driven trajectories run as 4 speculative copies of one chain give the sequential chain's
    records, positions and momenta bit for bit, also across the fallback to LJ forces after the first
    10-trajectory window (fallback_acc 1.01: every window's acceptance is below it), and with 16-bit
    list entries (bits 16) under the surrogate's descriptor and force kernels."""
    t, _ = trained
    ch = chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T)
    seq, spec = chmc.Engine(ch), chmc.Engine(ch, speculate=4)
    x0, p = t.positions(), t.surrogate_params()
    for e in (seq, spec):
        e.set_list_entry_bits(bits)
        e.set_list_lanes(8)
        e.surrogate_set_params(p)
        e.set_positions(x0)
    kw = dict(seed=83, surrogate=True, fallback_window=10, fallback_acc=fallback_acc)
    r1, n1 = seq.hmc(30, 0, chmc.LEAPFROG, 2 * DT, STEPS, **kw)
    r2, n2 = spec.hmc(30, 0, chmc.LEAPFROG, 2 * DT, STEPS, **kw)
    print("accepted", n1[0], "of 30; surrogate", r1["surrogate"][:, 0])
    assert 1 <= n1[0] <= 29, n1
    assert np.all(r1["surrogate"][:10] == 1)
    assert np.all(r1["surrogate"][10:] == (0 if fallback_acc > 1 else 1))
    assert r2.shape == (30, 1)
    assert np.array_equal(r1[:, 0], r2[:, 0]) and n1[0] == n2[0]
    assert np.array_equal(seq.positions(), spec.positions())
    assert np.array_equal(seq.momenta(), spec.momenta())
    seq.close()
    spec.close()


@pytest.mark.slow
def test_surrogate_chain_samples_lj_and_reports_ess(chmc, trained):
    """Same distribution as the LJ-driven chain; prints acceptance, time and ESS/s of U for both."""
    e, _ = trained
    x0 = e.positions()
    out = {}
    for name, surr in (("lj", False), ("surrogate", True)):
        e.set_positions(x0)
        e.hmc(100, 0, chmc.LEAPFROG, DT, STEPS, seed=500, surrogate=surr)
        chmc_sync = e.kinetic          # any call that synchronises
        chmc_sync()
        t0 = time.perf_counter()
        r, _ = e.hmc(2000, 0, chmc.LEAPFROG, DT, STEPS, seed=501, surrogate=surr)
        dt = time.perf_counter() - t0
        u = r["u"][:, 0] / N
        m, err, tau = stats.mean_err(u)
        ess = stats.ess(u)
        out[name] = (m, err)
        print(f"{name}: acceptance {r['accepted'].mean():.3f}  <U/N> {m:.5f} +- {err:.5f}  tau_int {tau:.2f}  "
              f"{dt * 1e3 / 2000:.3f} ms/traj  ESS/s {ess / dt:.1f}")
    z = stats.zscore(out["lj"], out["surrogate"])
    print(f"z(<U/N>) = {z:.2f}")
    assert abs(z) < 4


def test_force_matching_gradient(chmc, system, osur):
    """Energy + force matching loss gradient of the training kernels (residual forces, dL/dg, forward
    tangent and reverse pass over whole configurations) against the oracle's (pinned to torch double
    backward), for the force weight alone and mixed with the energy term."""
    e, x, p = system
    for w_force in (1.0, 0.3):
        grad, le, lf = e.surrogate_fm_loss_grad(w_force)
        me, mf, go = osur.force_loss_grad(p, [x], L, RC, w_force, bf=True)
        print(f"w_force {w_force}: loss_e gpu {le:.5e} oracle {me:.5e}; loss_f gpu {lf:.5e} oracle {mf:.5e}")
        assert abs(le - me) <= 2e-3 * me and abs(lf - mf) <= 2e-3 * mf
        gd = chmc.unpack_surrogate(grad, trained_only=True)
        for k, v in go.items():
            err = np.abs(gd[k] - v).max()
            print(f"  d loss / d {k}: max err {err:.3e} of max {np.abs(v).max():.3e}")
            assert err <= 3e-3 * np.abs(v).max()


def test_force_matching_training_reduces_both_losses(chmc, osur):
    e = _engine(chmc, seed=17, burn=60)
    e.surrogate_set_params(_params(osur, 3))
    e.surrogate_eval()
    G, _ = e.surrogate_features()
    e.surrogate_set_params(_params(osur, 3, G.astype(np.float64)))
    e.surrogate_reserve_configs(30)
    for k in range(30):
        e.hmc(3, 3, chmc.LEAPFROG, DT, STEPS, seed=700 + k)
        assert e.surrogate_collect_config() == k + 1
    le, lf = e.surrogate_train_fm(1500, 1.0, lr=2e-3, seed=8)
    e.close()
    print(f"energy MSE {le[:20].mean():.4e} -> {le[-50:].mean():.4e}; force MSE {lf[:20].mean():.4e} -> {lf[-50:].mean():.4e}")
    assert np.all(np.isfinite(le)) and np.all(np.isfinite(lf))
    assert le[-50:].mean() < 0.3 * le[:20].mean() and lf[-50:].mean() < 0.5 * lf[:20].mean()


def test_training_checkpoint_resumes_bitwise(chmc, osur):
    """Parameters + Adam moments + step count are a complete checkpoint: 40 Adam steps in one run equal
    20 steps, a checkpoint into a fresh engine, and 20 more (same rows: one batch of all rows, offset 0)."""
    def eng():
        e = _engine(chmc, seed=19, burn=20)
        e.surrogate_set_params(_params(osur, 5))
        e.surrogate_reserve(1)
        e.surrogate_collect()
        return e
    a = eng()
    a.surrogate_train(40, N, lr=1e-3, seed=3)
    b = eng()
    b.surrogate_train(20, N, lr=1e-3, seed=3)
    p, (m, v, st) = b.surrogate_params(), b.surrogate_adam_state()
    assert st == 20
    c = eng()
    c.surrogate_set_params(p)
    c.surrogate_set_adam_state(m, v, st)
    c.surrogate_train(20, N, lr=1e-3, seed=3)
    assert np.array_equal(c.surrogate_params(), a.surrogate_params())
    for e in (a, b, c):
        e.close()
