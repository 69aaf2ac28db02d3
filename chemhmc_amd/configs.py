"""The five workloads of the benchmark definition, as chain lists and trajectory parameters.

c1  plain HMC, N=512, rho*=0.8, T*=1.5, leapfrog 20 x 0.005, FP64 forces (the reference's CPU case)
c2  plain HMC, N=32768 (32^3), rho*=0.8442, T*=0.722, OMF2 (lambda=0.1932) 10 x 0.01, FP32 forces
c3  256 independent chains: N log-uniform in [256, 8192] rounded to a cube, rho* ~ U(0.3, 0.9),
    T* ~ U(0.7, 2.0), leapfrog 15 x 0.005, one launch per stage for the whole batch
c4  as c2 with the surrogate-driven trajectory (a trained per-atom MLP drives the MD, the exact FP64 LJ
    Hamiltonian judges each Metropolis step; the benchmark's surrogate workload builds it)
c5  N=1,048,576 (first N sites of a 102^3 lattice), rho*=0.8442, T*=0.722, OMF2 10 x 0.01

The box side follows from N / rho (L = 33.8585 at c2; the benchmark text's 33.68 corresponds to
rho*=0.8577).  Every chain starts from the seeded lattice + jitter of +-0.05 sigma.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import Chain, LEAPFROG, OMF2


@dataclass
class Workload:
    name: str
    chains: list
    scheme: int
    dt: float
    steps: int
    lam: float = 0.1932
    fp64_forces: bool = False
    rdf_bins: int = 0
    ntraj: int = 0
    notes: dict = field(default_factory=dict)

    @property
    def natoms(self):
        return sum(c.N for c in self.chains)


def box(N, rho):
    return (N / rho) ** (1.0 / 3.0)


def c1():
    return Workload("c1", [Chain(N=512, L=box(512, 0.8), cutoff=2.5, beta=1 / 1.5)], LEAPFROG, 0.005, 20,
                    fp64_forces=True, rdf_bins=200, ntraj=1000)


def c2():
    return Workload("c2", [Chain(N=32768, L=box(32768, 0.8442), cutoff=2.5, beta=1 / 0.722)], OMF2, 0.01, 10,
                    lam=0.1932, rdf_bins=500, ntraj=2000)


def c3(seed=3, nchains=256):
    rng = np.random.default_rng(seed)
    chains = []
    for _ in range(nchains):
        n = float(np.exp(rng.uniform(np.log(256.0), np.log(8192.0))))
        side = max(int(round(n ** (1.0 / 3.0))), 2)
        N = side ** 3
        rho = float(rng.uniform(0.3, 0.9))
        T = float(rng.uniform(0.7, 2.0))
        chains.append(Chain(N=N, L=box(N, rho), cutoff=2.5, beta=1 / T))
    return Workload("c3", chains, LEAPFROG, 0.005, 15, ntraj=1000)


def c5():
    return Workload("c5", [Chain(N=1048576, L=box(1048576, 0.8442), cutoff=2.5, beta=1 / 0.722)], OMF2, 0.01, 10,
                    lam=0.1932, rdf_bins=1000, ntraj=100)


WORKLOADS = {"c1": c1, "c2": c2, "c3": c3, "c5": c5}
