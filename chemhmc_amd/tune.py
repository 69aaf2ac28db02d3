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


def tune_skins(engine, run, bases=(0.35, 0.4, 0.45, 0.5, 0.6), exponents=(0.0, 1.0, 1.5), warm=8, reps=12,
               seed=1000):
    """Time `run(ntraj, seed)` (trajectories of the engine's workload) for every (base, exponent): warm
    trajectories after the skins change (list set-up, graph capture, and the first rebuild interval:
    lists are fresh after a change, so a short window would miss the rebuilds of a large skin and
    favour it), then reps timed ones.  The
    candidates are timed twice, in one order and then in the reverse order, and the two times summed,
    so that a steady drift of the state (a batch still relaxing from its start) does not favour the
    candidates timed last.  Set the list lanes first: with more lanes per atom than a sparse chain
    needs, list entries cost little and larger skins look cheaper than they are at the right lanes.
    The engine keeps the fastest skins.  Returns (skins, {"base/exponent": ms
    per trajectory})."""
    cands = []
    for ex in exponents:
        for b in bases:
            skins = density_scaled_skins(engine.chains, b, ex)
            try:
                engine.set_chain_skins(skins)
            except ChmcError:
                continue
            cands.append((f"{b}/{ex}", skins))
    if not cands:
        raise ChmcError("tune_skins: no candidate skin fits the list MD")
    table = {name: 0.0 for name, _ in cands}
    k = 0
    for order in (cands, cands[::-1]):
        for name, skins in order:
            engine.set_chain_skins(skins)
            run(warm, seed + 2 * k)
            t0 = time.perf_counter()
            run(reps, seed + 2 * k + 1)
            table[name] += (time.perf_counter() - t0) / reps * 1e3 / 2
            k += 1
    name, skins = min(cands, key=lambda c: table[c[0]])
    engine.set_chain_skins(skins)
    return skins, table
