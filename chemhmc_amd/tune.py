"""Autotuner of the neighbour-list skins of a batch of chains.

A larger skin makes every list sweep read more entries and every rebuild test more pairs, but the lists
are rebuilt less often.  How the two balance depends on the chain: at low density an atom has few
neighbours, so a wider shell costs little.  The tuner searches a two-parameter family of per-chain
skins, skin_k = base * clamp((rho_ref / rho_k) ** exponent, lo, hi), by timing a few trajectories of
the batch with each candidate, and keeps the fastest.  exponent = 0 is one skin for every chain.
Candidates the list MD cannot hold (a list cell of more than 256 atoms) are skipped.
"""
from __future__ import annotations

import time

import numpy as np

from . import ChmcError


def chain_density(chain):
    L = np.broadcast_to(np.asarray(chain.L, float), (3,))
    return chain.N / float(np.prod(L))


def density_scaled_skins(chains, base, exponent, rho_ref=0.6, lo=0.4, hi=2.0):
    """skin_k = base * clamp((rho_ref / rho_k) ** exponent, lo, hi), capped at 2 (the largest skin the
    list MD takes)."""
    rho = np.array([chain_density(c) for c in chains])
    return np.minimum(base * np.clip((rho_ref / rho) ** exponent, lo, hi), 2.0)


def snapshot(engine):
    """Positions of every chain and the trajectory counter (the heat bath's stream position)."""
    return [engine.positions(k) for k in range(engine.nchains)], engine.trajectory()


def restore(engine, snap):
    xs, traj = snap
    for k, x in enumerate(xs):
        engine.set_positions(x, k)
    engine.set_trajectory(traj)


def tune_skins(engine, run, bases=(0.35, 0.4, 0.45, 0.5, 0.6), exponents=(0.0, 1.0, 1.5), warm=0, reps=24,
               seed=1000, lanes=0, checks=None):
    """Time `run(ntraj, seed)` (trajectories of the engine's workload) for every (base, exponent) as a
    paired comparison: each candidate starts from the same positions and trajectory counter and runs the
    same random streams, warm untimed trajectories and then reps timed ones in one call, as a timed run
    that follows the tuner starts (fresh lists, then a rebuild whenever an atom has moved skin / 2; a
    window that skips the first rebuild interval misranks the skins).  The lanes per atom of the list
    sweep are set for every candidate (`lanes`, 0: autotuned at that candidate's lists), since the best
    lanes change with the list length and a skin timed at the wrong lanes is misjudged.  `checks`
    (optional) also searches the rebuild-check interval of list MD (chmc_set_list_check) with every
    skin: a longer interval saves launches but lowers the rebuild trigger, which pays where rebuilds are
    cheap.  The engine is returned to the starting state with the fastest skins, their lanes and check
    interval.  Returns (skins, {"base/exponent": ms per trajectory}, with "/m" appended when checks are
    searched)."""
    snap = snapshot(engine)
    table, best = {}, None
    for ex in exponents:
        for b in bases:
            skins = density_scaled_skins(engine.chains, b, ex)
            try:
                engine.set_chain_skins(skins)
            except ChmcError:
                continue
            for m in (checks if checks else (None,)):
                if m is not None:
                    engine.set_list_check(m)
                restore(engine, snap)
                engine.set_list_lanes(lanes)
                lp = engine.list_lanes()
                if warm > 0:
                    run(warm, seed)
                t0 = time.perf_counter()
                run(reps, seed + 1)
                ms = (time.perf_counter() - t0) / reps * 1e3
                table[f"{b}/{ex}" if m is None else f"{b}/{ex}/{m}"] = ms
                if best is None or ms < best[0]:
                    best = (ms, skins, lp, m)
    if best is None:
        raise ChmcError("tune_skins: no candidate skin fits the list MD")
    engine.set_chain_skins(best[1])
    if best[3] is not None:
        engine.set_list_check(best[3])
    restore(engine, snap)
    engine.set_list_lanes(best[2])
    return best[1], table
