"""Statistical parity of plain GPU HMC (FP32 pair forces, FP64 x/p/dH) with the FP64 oracle chain at
config 1 (N=512, rho*=0.8, T*=1.5, rc=2.5, leapfrog 20 x 0.005): acceptance, <U/N>, virial pressure
and g(r) (200 bins to L/2).  Errors are tau_int-corrected (Sokal windowing); the gate is 2 statistical
sigma per scalar observable and a mean z^2 over the populated g(r) bins below 2 (z-scores are printed).  Both chains start from the same seeded lattice
and diverge (FP32 vs FP64 forces), so they are independent samples of the same distribution.
Burn-in accepts every trajectory (the reference's thermalization_steps): from the lattice the first
trajectory carries dH ~ 7 (pairs of the sqrt(5) a = 2.41 shell leaving the unshifted cutoff, Q3), so a
Metropolis chain started there can stay rejected for hundreds of trajectories.
"""
import numpy as np
import pytest

from chemhmc_amd import stats

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N, RHO, T, RC = 512, 0.8, 1.5, 2.5
L = (N / RHO) ** (1 / 3)
NB = 200
DT, STEPS = 0.005, 20
BURN, SEG, NSEG = 300, 50, 50


def _run_gpu(chmc):
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T))
    e.init_lattice(1, 0.05)
    e.hmc(BURN, BURN + 1, chmc.LEAPFROG, DT, STEPS, seed=100, use_graph=True)    # accept-all burn-in
    recs, g = [], []
    for k in range(NSEG):
        r, _ = e.hmc(SEG, 0, chmc.LEAPFROG, DT, STEPS, seed=1000 + 7919 * k, use_graph=True)
        recs.append(r[:, 0])
        g.append(e.rdf(NB, L / (2 * NB))[1])
    r = np.concatenate(recs)
    return r["u"] / N, r["w"], r["accepted"].astype(float), np.array(g)


def _run_oracle(orc):
    s = orc.Sim(N, [L] * 3, cutoff=RC, beta=1 / T)
    s.x[:] = orc.lattice_init(N, [L] * 3, 1, 0, 0.05)
    s.hmc(BURN, BURN + 1, orc.LEAPFROG, DT, STEPS, seed=200)
    u, w, a, g = [], [], [], []
    for k in range(NSEG):
        r = s.hmc(SEG, 0, orc.LEAPFROG, DT, STEPS, seed=2000 + 7919 * k)
        u.append(r["u"] / N)
        w.append(r["w"])
        a.append(r["acc"].astype(float))
        g.append(orc.rdf(s.x, [L] * 3, NB, L / (2 * NB))[1])
    return np.concatenate(u), np.concatenate(w), np.concatenate(a), np.array(g)


def test_plain_hmc_matches_oracle_statistically(chmc, orc):
    ug, wg, ag, gg = _run_gpu(chmc)
    uo, wo, ao, go = _run_oracle(orc)
    V = L ** 3
    pg = stats.pressure(wg, N, V, 1 / T)
    po = stats.pressure(wo, N, V, 1 / T)
    z = {}
    for name, a, b in [("acceptance", ag, ao), ("U/N", ug, uo), ("P", pg, po)]:
        ma, ea, _ = stats.mean_err(a)
        mb, eb, _ = stats.mean_err(b)
        z[name] = stats.zscore((ma, ea), (mb, eb))
        print(f"{name}: gpu {ma:.5f} +- {ea:.5f}  oracle {mb:.5f} +- {eb:.5f}  z = {z[name]:.2f}")
    # g(r): per-bin mean over the configuration samples, error from the sample scatter
    mask = (gg.mean(0) > 0.05) & (go.mean(0) > 0.05)
    eg = gg.std(0, ddof=1) / np.sqrt(gg.shape[0])
    eo = go.std(0, ddof=1) / np.sqrt(go.shape[0])
    zg = (gg.mean(0) - go.mean(0))[mask] / np.hypot(eg, eo)[mask]
    print(f"g(r): {mask.sum()} bins, mean z^2 = {np.mean(zg ** 2):.2f}, max |z| = {np.max(np.abs(zg)):.2f}")
    assert all(v < 2.0 for v in z.values()), z
    assert np.mean(zg ** 2) < 2.0
