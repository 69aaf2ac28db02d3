"""The remaining integrators of the drop-in boundary on the GPU against the oracle:

  lp_leapfrog / lp_omf4 (code-carsten/integrator.hh:51-79, 156-197): every update rounded to the grid
      10^-n_prec.  FP64 forces: every state component on the grid and equal to the oracle's except where
      the two force sums round an increment to neighbouring grid points (at most 1e-3 of the components,
      each by one grid step); bit-exact time reversal of the GPU run itself.  FP32 list MD: on the grid and
      within the FP32-force bar of the oracle.  Whole lp_leapfrog HMC trajectories (graph-captured) against
      the oracle chain: dH within 1e-8, identical decisions.
  randomized trajectory length (integrator.cpp:37-49, HMC.cpp:82-85): binomial step counts drawn from the
      same mt19937_64 stream as the oracle, records equal (dH within 1e-8, identical decisions).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RC = 2.5


def _blob_state(n_prec, seed=5):
    """64 atoms in the middle of a 20-sided box (no atom reaches a face), x and p on the grid."""
    g = 10.0 ** n_prec
    rng = np.random.default_rng(seed)
    site = np.stack(np.meshgrid(*[np.arange(4)] * 3, indexing="ij"), -1).reshape(64, 3)
    x = np.round((8.35 + 1.1 * site + rng.uniform(-0.05, 0.05, (64, 3))) * g) / g
    return x, g


def _on_grid(v, g):
    return bool(np.all(np.round(v * g) / g == v))


@pytest.mark.parametrize("scheme", [1, 3])
def test_lp_integrators_fp64_vs_oracle(chmc, orc, scheme):
    n = 6
    x0, g = _blob_state(n)
    L = 20.0
    e = chmc.Engine(chmc.Chain(N=64, L=L, cutoff=RC, beta=1 / 1.5), fp64_forces=True)
    e.set_positions(x0)
    p0 = np.round(orc.heatbath(64, 1.0, 1 / 1.5, 5, 0, 1) * g) / g
    e.set_momenta(p0)
    e.set_lp_exponent(n)
    assert e.lp_exponent() == n
    s = orc.Sim(64, [L] * 3, cutoff=RC, beta=1 / 1.5)
    s.set_lp_exponent(n)
    s.x[:] = x0
    s.p[:] = p0
    e.integrate(scheme, 0.004, 12, 0.1932)
    s.integrate(scheme, 0.004, 12, 0.1932)
    x, p = e.positions(), e.momenta()
    assert _on_grid(x, g) and _on_grid(p, g)
    for a, b in ((x, s.x), (p, s.p)):
        d = np.abs(a - b)
        print("components off the oracle:", int(np.count_nonzero(d)), "max", d.max())
        assert np.count_nonzero(d) <= 1e-3 * d.size + 1 and d.max() <= 1.5 / g
    # bit-exact time reversal of the GPU run
    e.set_momenta(-p)
    e.integrate(scheme, 0.004, 12, 0.1932)
    assert np.array_equal(e.positions(), x0) and np.array_equal(-e.momenta(), p0)
    e.close()


def test_lp_leapfrog_fp32_lists_vs_oracle(chmc, orc):
    N, rho, T = 4000, 0.8442, 0.722
    L = (N / rho) ** (1 / 3)
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T))
    e.init_lattice(3, 0.05)
    e.heatbath(3, 1)
    e.set_lp_exponent(8)
    s = orc.Sim(N, [L] * 3, cutoff=RC, beta=1 / T)
    s.set_lp_exponent(8)
    s.x[:] = e.positions()
    s.p[:] = e.momenta()
    e.integrate(chmc.LP_LEAPFROG, 0.005, 10)
    s.integrate(orc.LP_LEAPFROG, 0.005, 10)
    x = e.positions()
    assert _on_grid(x, 1e8) and _on_grid(e.momenta(), 1e8)
    d = x - s.x
    d -= L * np.round(d / L)
    print("rms", np.sqrt(np.mean(d ** 2)), "max", np.abs(d).max())
    assert np.sqrt(np.mean(d ** 2)) < 1e-6 and np.abs(d).max() < 2e-4
    e.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_lp_leapfrog_hmc_vs_oracle(chmc, orc, use_graph):
    N, rho, T = 512, 0.8, 1.5
    L = (N / rho) ** (1 / 3)
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T), fp64_forces=True)
    e.init_lattice(7, 0.05)
    e.set_lp_exponent(10)
    s = orc.Sim(N, [L] * 3, cutoff=RC, beta=1 / T)
    s.set_lp_exponent(10)
    s.x[:] = e.positions()
    r, nacc = e.hmc(15, 3, chmc.LP_LEAPFROG, 0.005, 20, seed=4, use_graph=use_graph)
    ro = s.hmc(15, 3, orc.LP_LEAPFROG, 0.005, 20, seed=4)
    r = r[:, 0]
    print("dh gpu   ", np.array2string(r["dh"], precision=4))
    print("dh oracle", np.array2string(ro["dh"], precision=4))
    assert np.max(np.abs(r["dh"] - ro["dh"])) < 1e-8
    assert np.array_equal(r["accepted"], ro["acc"]) and nacc[0] == ro["nacc"]
    assert _on_grid(e.positions(), 1e10)
    e.close()


def test_randomized_trajectory_length_vs_oracle(chmc, orc):
    """randomize_trajectory: each trajectory's steps ~ Binomial(2 * steps, 1/2) from the chain's
    mt19937_64, drawn before its Metropolis uniform (integrator.cpp:37-49, HMC.cpp:82-85, :97)."""
    N, rho, T = 512, 0.8, 1.5
    L = (N / rho) ** (1 / 3)
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T), fp64_forces=True)
    e.init_lattice(7, 0.05)
    s = orc.Sim(N, [L] * 3, cutoff=RC, beta=1 / T)
    s.x[:] = e.positions()
    r, nacc = e.hmc(20, 3, chmc.LEAPFROG, 0.005, 10, seed=21, randomize=True)
    ro = s.hmc(20, 3, orc.LEAPFROG, 0.005, 10, seed=21, randomize=True)
    r = r[:, 0]
    # the step counts themselves, from the same generator
    g = orc.MT19937_64(21)
    steps = []
    for t in range(20):
        steps.append(g.binomial_steps(10))
        if t + 1 >= 3:
            g.uniform()
    print("steps", steps)
    assert len(set(steps)) > 3
    assert np.max(np.abs(r["dh"] - ro["dh"])) < 1e-8
    assert np.array_equal(r["accepted"], ro["acc"]) and nacc[0] == ro["nacc"]
    e.close()
