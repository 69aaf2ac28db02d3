"""The surrogate workload at full size: N = 32768 (rho* = 0.8442, T* = 0.722, r_c = 2.5, OMF2 lambda =
0.1932, 10 x 0.01), a 32-64-64-1 network on 32 radial descriptors trained on the GPU with energy +
force matching on snapshots of this system's own plain-HMC chain, then surrogate-driven HMC whose
Metropolis step uses the exact FP64 LJ Hamiltonian.

  training        both losses fall
  forces          the surrogate force of the MD kernels (neighbour-list descriptors, MLP, list sweep)
                  against the FP64 oracle of the same network: rms error <= 2e-3 of the rms force
  exact accept    on accept-all trajectories driven by the surrogate, every record's dH equals
                  beta (K_f + V_f - K_i - V_i) from the oracle's FP64 LJ energy of the initial and final
                  positions and FP64 kinetic energies of the heat-bath and final momenta (|diff| < 1e-6)
  decisions       on Metropolis trajectories driven by the surrogate, accepted iff u < exp(-dH) with the
                  chain's mt19937_64 uniforms, and a rejected trajectory leaves the positions unchanged
  graph = direct  records of graph-replayed and directly launched surrogate trajectories are bitwise equal
  fallback        with a window of 2 and an unreachable acceptance threshold the run switches to LJ forces
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, RHO, T, RC = 32768, 0.8442, 0.722, 2.5
L = (N / RHO) ** (1 / 3)
DT, STEPS, LAM = 0.01, 10, 0.1932
NCONF = 24


@pytest.fixture(scope="module")
def osur():
    from oracle import surrogate
    return surrogate


@pytest.fixture(scope="module")
def c4(chmc, osur):
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T))
    e.init_lattice(31, 0.05)
    e.hmc(60, 61, chmc.OMF2, DT, STEPS, lam=LAM, seed=32)                # accept-all burn-in
    p = osur.init_params(4)
    e.surrogate_set_params(p)
    e.surrogate_eval()
    G, _ = e.surrogate_features()
    p["shift"] = G.mean(0).astype(np.float64)
    p["scale"] = 1.0 / (G.std(0).astype(np.float64) + 1e-3)
    e.surrogate_set_params(p)
    e.surrogate_reserve_configs(NCONF)
    for k in range(NCONF):
        e.hmc(2, 2, chmc.OMF2, DT, STEPS, lam=LAM, seed=400 + k)
        e.surrogate_collect_config()
    le, lf = e.surrogate_train_fm(400, 1.0, lr=1e-3, seed=9)
    e.hmc(5, 6, chmc.OMF2, DT, STEPS, lam=LAM, seed=33)                  # off the last snapshot
    yield e, le, lf
    e.close()


def test_training_reduces_both_losses(c4):
    _, le, lf = c4
    print(f"energy MSE {le[:10].mean():.4e} -> {le[-20:].mean():.4e}; force MSE {lf[:10].mean():.4e} -> {lf[-20:].mean():.4e}")
    assert np.all(np.isfinite(le)) and np.all(np.isfinite(lf))
    assert le[-20:].mean() < 0.5 * le[:10].mean() and lf[-20:].mean() < 0.7 * lf[:10].mean()


def test_md_forces_vs_oracle(chmc, c4, osur):
    e, _, _ = c4
    x = e.positions()
    p = chmc.unpack_surrogate(e.surrogate_params())
    E = e.surrogate_eval(store_force=True)[0]
    f = e.force()
    Eo, fo, _ = osur.energy_grad(x, L, RC, p, bf=True)
    rms = np.sqrt((fo ** 2).sum(1).mean())
    err = np.sqrt(((f - fo) ** 2).sum(1).mean())
    print(f"E gpu {E:.4f} oracle {Eo:.4f}; force rms err {err:.3e} of rms {rms:.3e}")
    assert err <= 2e-3 * rms
    assert abs(E - Eo) <= 1e-3 * abs(Eo) + 1.0


def test_exact_accept_dh(chmc, orc, c4):
    """dH of surrogate-driven trajectories is the exact FP64 LJ Hamiltonian difference."""
    e, _, _ = c4
    for k in range(3):
        x0 = e.positions()
        t0 = e.trajectory()
        seed = 810 + k
        r, _ = e.hmc(1, 10 ** 9, chmc.OMF2, DT, STEPS, lam=LAM, seed=seed, surrogate=True)   # thermalisation: kept
        assert r["surrogate"][0, 0] == 1 and r["accepted"][0, 0] == 1
        xf, pf = e.positions(), e.momenta()
        pi = orc.heatbath(N, 1.0, 1 / T, seed, 0, t0 + 1)
        Vi = orc.potential_binning(x0, [L] * 3, RC)
        Vf = orc.potential_binning(xf, [L] * 3, RC)
        dh = (orc.kinetic(pf, 1.0) + Vf - orc.kinetic(pi, 1.0) - Vi) / T
        print(f"dH record {r['dh'][0, 0]:.9f}  FP64 recomputation {dh:.9f}  V_f {Vf:.4f}")
        assert abs(r["dh"][0, 0] - dh) < 1e-6
        assert abs(r["u"][0, 0] - Vf) < 1e-8 * abs(Vf)


def test_metropolis_decisions(chmc, orc, c4):
    e, _, _ = c4
    e.hmc(1, 0, chmc.OMF2, DT, STEPS, lam=LAM, seed=1)        # a Metropolis run: the counter is past thermalisation
    nt = 12
    x0 = e.positions()
    r, nacc = e.hmc(nt, 0, chmc.OMF2, DT, STEPS, lam=LAM, seed=900, surrogate=True)
    g = orc.MT19937_64(900)
    u = np.array([g.uniform() for _ in range(nt)])
    dh = r["dh"][:, 0]
    acc = r["accepted"][:, 0].astype(bool)
    print(f"surrogate trajectories: dH {np.array2string(dh, precision=2)}; accepted {int(acc.sum())} of {nt}")
    assert np.array_equal(acc, u < np.exp(-dh))
    assert nacc[0] == acc.sum()
    if not acc.any():
        assert np.array_equal(e.positions(), x0)


def test_graph_equals_direct(chmc, c4):
    e, _, _ = c4
    x0, t0 = e.positions(), e.trajectory()
    e.set_positions(x0)                         # both runs start from lists rebuilt at x0
    r1, _ = e.hmc(4, 0, chmc.OMF2, DT, STEPS, lam=LAM, seed=77, use_graph=True, surrogate=True)
    e.set_positions(x0)
    e.set_trajectory(t0)
    r2, _ = e.hmc(4, 0, chmc.OMF2, DT, STEPS, lam=LAM, seed=77, use_graph=False, surrogate=True)
    assert np.all(r1["surrogate"] == 1)
    assert np.array_equal(r1, r2)


def test_fallback_switches_to_lj(chmc, c4):
    e, _, _ = c4
    r, _ = e.hmc(6, 0, chmc.OMF2, DT, STEPS, lam=LAM, seed=78, surrogate=True, fallback_window=2, fallback_acc=1.01)
    assert np.all(r["surrogate"][:2] == 1) and np.all(r["surrogate"][2:] == 0)
