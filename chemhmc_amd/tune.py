"""Autotuner of the neighbour-list skins of a batch of chains.

A larger skin makes every list sweep read more entries and every rebuild test more pairs, but the lists
are rebuilt less often.  How the two balance depends on the chain: at low density an atom has few
neighbours, so a wider shell costs little.  The tuner searches a two-parameter family of per-chain
skins, skin_k = base * clamp((rho_ref / rho_k) ** exponent, lo, hi), by timing a few trajectories of
the batch with each candidate, and keeps the fastest.  exponent = 0 is one skin for every chain.
Candidates the list MD cannot hold (a list cell of more than 256 atoms) are skipped.  cached_tune_skins
keeps the outcome in a JSON file, keyed by the chains, the workload, the candidate set and the library
build, so that a later run of the same problem sets the tuned skins, check interval and lanes without
searching again.
"""
from __future__ import annotations

import hashlib
import json
import os
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


def cache_key(engine, workload, candidates):
    """Identity of a tuning problem: the chains (atoms, box, cut-off, LJ parameters, mass, beta), the
    engine's speculative copies and list entry format, the caller's workload description, the candidate
    settings of the search and the library build (size and modification time of libchmc.so), so that
    another system, search or build is tuned afresh."""
    from . import LIB_PATH
    try:
        st = os.stat(LIB_PATH)
        build = [st.st_size, st.st_mtime_ns]
    except OSError:
        build = None
    desc = {"chains": [[c.N, list(c.box()), c.cutoff, c.sigma, c.eps, c.mass, c.beta] for c in engine.chains],
            "speculation": engine.speculation(), "entry_bits": engine.list_entry_bits(), "library": build,
            "workload": workload, "candidates": candidates}
    return hashlib.sha256(json.dumps(desc, sort_keys=True, default=list).encode()).hexdigest()


def _read_cache(path):
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    return data if isinstance(data, dict) else {}


def cached_tune_skins(engine, run, cache, workload, **kw):
    """tune_skins(engine, run, **kw) with its outcome kept in the JSON file `cache` under
    cache_key(engine, workload, kw).  A hit sets the stored skins, rebuild-check interval (when the search
    covered it) and lanes, and runs nothing; a miss runs the search and adds its outcome to the file
    (written whole to a temporary file and renamed over it).  Returns (skins, table, hit)."""
    key = cache_key(engine, workload, kw)
    entry = _read_cache(cache).get(key)
    if isinstance(entry, dict):
        skins = np.asarray(entry["skins"], dtype=np.float64)
        engine.set_chain_skins(skins)
        if entry.get("check") is not None:
            engine.set_list_check(int(entry["check"]))
        engine.set_list_lanes(int(entry["lanes"]))
        return skins, dict(entry["table"]), True
    skins, table = tune_skins(engine, run, **kw)
    data = _read_cache(cache)
    data[key] = {"skins": [float(v) for v in skins], "check": engine.list_check() if kw.get("checks") else None,
                 "lanes": int(engine.list_lanes()), "table": table}
    d = os.path.dirname(os.path.abspath(cache))
    os.makedirs(d, exist_ok=True)
    tmp = os.path.join(d, f".{os.path.basename(cache)}.{os.getpid()}.tmp")
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, cache)
    return skins, table, False
