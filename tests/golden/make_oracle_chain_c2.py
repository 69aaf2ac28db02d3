"""Writes tests/golden/oracle_chains_N32768.npz: independent FP64 oracle plain-HMC chains at N = 32768
(32^3 lattice + jitter 0.05, rho* = 0.8442, T* = 0.722, r_c = 2.5, OMF2 lambda = 0.1932, 10 x 0.01),
the samples the GPU chains of tests/test_gpu_hmc_stats_c2.py are compared with.

At this state point about 2 % of the trajectories are accepted, so one chain decorrelates over several
hundred trajectories and a single chain's tau_int-corrected error is not reliable; the comparison is
made between independent chains (error of the mean from their scatter).  Protocol per chain c (shared
with the test): lattice jitter seed 100 + c, BURN accept-all trajectories (the reference's
thermalization_steps, HMC.cpp:98-100), then NSEG segments of SEG Metropolis trajectories; per
trajectory U/N, the virial W and the decision; after each segment g(r) in 500 bins of L/1000.
About 40 minutes on 8 host cores.  Run from the repository root: python tests/golden/make_oracle_chain_c2.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

N, RHO, T, RC = 32768, 0.8442, 0.722, 2.5
L = (N / RHO) ** (1 / 3)
NB = 500
DT, STEPS, LAM = 0.01, 10, 0.1932
NCHAIN, BURN, SEG, NSEG = 8, 200, 50, 20


def main():
    oracle.build()
    t0 = time.perf_counter()
    u, w, a, dh, g = [], [], [], [], []
    for c in range(NCHAIN):
        s = oracle.Sim(N, [L] * 3, cutoff=RC, beta=1 / T)
        s.x[:] = oracle.lattice_init(N, [L] * 3, 100 + c, 0, 0.05)
        s.hmc(BURN, BURN + 1, oracle.OMF2, DT, STEPS, lam=LAM, seed=300 + c)
        cu, cw, ca, cd, cg = [], [], [], [], []
        for k in range(NSEG):
            r = s.hmc(SEG, 0, oracle.OMF2, DT, STEPS, lam=LAM, seed=3000 + 100 * c + 7919 * k)
            cu.append(r["u"] / N)
            cw.append(r["w"])
            ca.append(r["acc"])
            cd.append(r["dh"])
            cg.append(oracle.rdf(s.x, [L] * 3, NB, L / (2 * NB))[1])
        u.append(np.concatenate(cu))
        w.append(np.concatenate(cw))
        a.append(np.concatenate(ca))
        dh.append(np.concatenate(cd))
        g.append(np.array(cg))
        print(f"chain {c + 1}/{NCHAIN}: accepted {int(a[-1].sum())}, <U/N> {u[-1].mean():.5f}, "
              f"{time.perf_counter() - t0:.0f} s", flush=True)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_chains_N32768.npz")
    np.savez_compressed(out, u_per_n=np.array(u), w=np.array(w), accepted=np.array(a).astype(np.int8),
                        dh=np.array(dh), g=np.array(g).astype(np.float32),
                        protocol=np.array([N, NCHAIN, BURN, SEG, NSEG, NB], dtype=np.int64),
                        params=np.array([RHO, T, RC, DT, STEPS, LAM, L]))


if __name__ == "__main__":
    main()
