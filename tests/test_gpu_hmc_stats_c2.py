"""Parity of plain GPU HMC with the FP64 oracle at N = 32768 (rho* = 0.8442, T* = 0.722, r_c = 2.5,
OMF2 lambda = 0.1932, 10 x 0.01; FP32 pair forces on the neighbour lists, FP64 x/p/V/K/dH).

1. Acceptance, trajectory by trajectory: from the same positions, heat-bath stream and Metropolis uniform
   the GPU trajectory and the oracle's give dH within DH_TOL and the same decision wherever the uniform is
   not within that band of exp(-dH).  Both restart from the oracle's state before every trajectory.
2. <U/N>, virial pressure, acceptance and g(r) (500 bins to L/2) of independent chains, each within 2
   statistical sigma: 8 oracle chains (tests/golden/oracle_chains_N32768.npz, written by
   tests/golden/make_oracle_chain_c2.py, about 40 minutes of host time) against 32 GPU chains run by one
   batched context with the same protocol (200 accept-all trajectories off the lattice, then 1000
   Metropolis trajectories, g(r) after every 50).  About 2 % of the trajectories are accepted here, so a
   single chain decorrelates over hundreds of trajectories; the errors are those of the chain means
   (sample scatter over chains), not tau_int estimates of one chain.  z-scores are printed.
"""
import os

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "oracle_chains_N32768.npz")
N, RHO, T, RC = 32768, 0.8442, 0.722, 2.5
L = (N / RHO) ** (1 / 3)
DT, STEPS, LAM = 0.01, 10, 0.1932
DH_TOL = 0.15            # pairs within ~1e-6 of r_c at the last step can sit on either side (0.016 each)
NGPU = 32


def test_trajectory_dh_and_decisions_vs_oracle_N32768(chmc, orc):
    with chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T)) as e:
        e.init_lattice(11, 0.05)
        e.hmc(60, 61, chmc.OMF2, DT, STEPS, lam=LAM, seed=12)            # accept-all burn-in
        s = orc.Sim(N, [L] * 3, cutoff=RC, beta=1 / T)
        s.x[:] = e.positions()
        dg, do, same, banded = [], [], 0, 0
        for k in range(24):
            seed = 500 + k
            e.set_positions(s.x)
            e.set_trajectory(0)
            r, _ = e.hmc(1, 0, chmc.OMF2, DT, STEPS, lam=LAM, seed=seed, use_graph=False)
            ro = s.hmc(1, 0, orc.OMF2, DT, STEPS, lam=LAM, seed=seed)
            u = orc.MT19937_64(seed).uniform()
            a, b = float(r["dh"][0, 0]), float(ro["dh"][0])
            dg.append(a)
            do.append(b)
            assert abs(a - b) < DH_TOL, (k, a, b)
            if abs(np.log(u) + b) > DH_TOL:
                assert int(r["accepted"][0, 0]) == int(ro["acc"][0]), (k, a, b, u)
                same += 1
            else:
                banded += 1
            if ro["acc"][0]:
                assert np.max(np.abs(e.positions() - s.x)) < 1e-4
        print("dH gpu   ", np.array2string(np.array(dg), precision=3))
        print("dH oracle", np.array2string(np.array(do), precision=3))
        print(f"max |ddH| {np.max(np.abs(np.array(dg) - np.array(do))):.3e}; decisions compared {same}, in the band {banded}")


def _load():
    z = np.load(FIXTURE, allow_pickle=False)
    n, nch, burn, seg, nseg, nb = (int(v) for v in z["protocol"])
    assert n == N and tuple(z["params"][:6]) == (RHO, T, RC, DT, STEPS, LAM)
    return dict(nch=nch, burn=burn, seg=seg, nseg=nseg, nb=nb, u=z["u_per_n"], w=z["w"],
                acc=z["accepted"].astype(float), g=z["g"].astype(np.float64))


def test_independent_chains_N32768_within_2_sigma(chmc):
    from chemhmc_amd import stats
    o = _load()
    nb = o["nb"]
    with chmc.Engine([chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T)] * NGPU) as e:
        e.init_lattice(200, 0.05)
        e.hmc(o["burn"], o["burn"] + 1, chmc.OMF2, DT, STEPS, lam=LAM, seed=600, use_graph=True)
        recs, g = [], []
        for k in range(o["nseg"]):
            r, _ = e.hmc(o["seg"], 0, chmc.OMF2, DT, STEPS, lam=LAM, seed=7000 + 7919 * k, use_graph=True)
            recs.append(r)
            g.append([e.rdf(nb, L / (2 * nb), chain=c)[1] for c in range(NGPU)])
    r = np.concatenate(recs)                              # [trajectories, chains]
    gg = np.array(g).transpose(1, 0, 2)                   # [chains, segments, bins]
    V = L ** 3
    per_chain = {
        "acceptance": (r["accepted"].astype(float).mean(0), o["acc"].mean(1)),
        "U/N": ((r["u"] / N).mean(0), o["u"].mean(1)),
        "P": (stats.pressure(r["w"], N, V, 1 / T).mean(0), stats.pressure(o["w"], N, V, 1 / T).mean(1)),
    }
    z = {}
    for name, (a, b) in per_chain.items():
        ma, ea = a.mean(), a.std(ddof=1) / np.sqrt(a.size)
        mb, eb = b.mean(), b.std(ddof=1) / np.sqrt(b.size)
        z[name] = stats.zscore((ma, ea), (mb, eb))
        print(f"{name}: gpu {ma:.5f} +- {ea:.5f} ({a.size} chains)  oracle {mb:.5f} +- {eb:.5f} ({b.size} chains)  "
              f"z = {z[name]:.2f}")
    ga, gb = gg.mean(1), o["g"].mean(1)                   # per-chain mean g(r)
    mask = (ga.mean(0) > 0.05) & (gb.mean(0) > 0.05)
    ea = ga.std(0, ddof=1) / np.sqrt(ga.shape[0])
    eb = gb.std(0, ddof=1) / np.sqrt(gb.shape[0])
    zg = (ga.mean(0) - gb.mean(0))[mask] / np.hypot(ea, eb)[mask]
    print(f"g(r): {mask.sum()} bins, mean z^2 = {np.mean(zg ** 2):.2f}, max |z| = {np.max(np.abs(zg)):.2f}")
    assert all(np.isfinite(v) and v < 2.0 for v in z.values()), z
    assert np.mean(zg ** 2) < 2.0
