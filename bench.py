"""Headline benchmark: plain HMC of the Lennard-Jones liquid at N = 32768, rho* = 0.8442, T* = 0.722,
rc = 2.5, 2nd-order Omelyan (lambda = 0.1932) with 10 steps of dt = 0.01, FP32 pair forces with FP64
positions / momenta / dH, on one MI355X.

A "step" is one HMC trajectory (heat bath, 10 OMF2 steps = 21 force evaluations, FP64 V_f, K_f,
Metropolis, restore-on-reject), its kernels launched directly on the context stream (--graph replays
captured hipGraphs instead: no faster per trajectory here, and each call pays its captures).  The chain
runs as speculative copies (--spec; by default 8, or for runs shorter than 160 trajectories the count
that splits them into whole batches): a batch runs the chain's next trajectories at once, and the
records are the sequential chain's, bit for bit.  The timed region starts with the state resident in
HBM and ends after the device has finished the K-th trajectory.

  value      atom-steps/s = N * 10 MD steps * K trajectories / seconds
  thermalisation  a fixed burn-in of accept-all trajectories (the reference's thermalization_steps)
             takes the chain off its lattice start before anything is tuned or timed, whatever --warmup
             says; --warmup adds ordinary (Metropolis) untimed trajectories after it
  ess        ESS/s of U/N (Sokal-windowed tau_int) from a run of at least ESS_TRAJ trajectories (the timed
             run itself when it is that long, else a separate run after it, timed on its own); null when
             U/N never changed in that run (no accepted trajectory: no information about tau_int)
  roofline   the dominant kernel, the FP32 neighbour-list sweep (force + kick + drift), against the FP32
             VALU peak with the algorithmic work of the reference's 27-cell sweep it replaces:
             8 * N_cand + 13 * N_int flops per atom-evaluation (N_cand, N_int counted on the device for the
             current configuration) / its average duration measured with HIP events on the context's
             stream (50 launches back to back); its HBM bytes (algorithmic and rocprof-measured) beside it
  cpu_baseline  the FP64 C oracle (a port of the reference algorithm, cpu-opt binning) timed on this
             host's cores for a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, RHO, T, RC = 32768, 0.8442, 0.722, 2.5
DT, NSTEPS, LAM = 0.01, 10, 0.1932
THERM = 200                     # accept-all burn-in trajectories off the lattice start
ESS_TRAJ = 4000                 # trajectories of the run the ESS/s figure is measured on
PEAK_FP32_TFLOPS = 157.3        # MI355X vector FP32 (vendor figure; packed FP32)
PEAK_HBM_GBS = 8000.0
# list sweep bytes per atom besides its list row: own FP32 coordinates 16, entry count 4, p 24 + 24 and
# x 24 + 24 (read and written, slot order), rebuild reference 16, next FP32 coordinates 16
SWEEP_BYTES_PER_ATOM = 148.0
# HBM traffic of one list sweep (force + kick + drift) from rocprofv3 --pmc FETCH_SIZE and --pmc
# WRITE_SIZE (medians over the profiled dispatches, KiB).  Calibration (--workload calib, cold 1 GiB
# device copies): FETCH_SIZE reads 524,300 KiB for 1,048,576 KiB read, WRITE_SIZE 1,048,576 KiB for as
# many written, so FETCH_SIZE is scaled by FETCH_SCALE and WRITE_SIZE taken as it is
# (profiles/hbm_counter_calibration.md).  Keyed by (skin, lanes, speculative copies in the launch) at
# N=32768 (profiles/lj_list_sweep_N32768.md, profiles/lj_list_sweep_spec8_N32768.md): reported as the
# roofline's traffic when the run's tuned skin, lanes and copies are those, null otherwise
FETCH_SCALE = 2.0
SWEEP_TRAFFIC_PROFILED = {(0.6, 8, 1): (FETCH_SCALE * 9976.0 + 2048.0) * 1024.0,
                          (0.7, 8, 1): (FETCH_SCALE * 9970.0 + 2048.0) * 1024.0,
                          (0.7, 816, 1): (FETCH_SCALE * 10086.8 + 2048.0) * 1024.0,
                          (0.4, 812, 8): (FETCH_SCALE * 60488.0 + 16384.2) * 1024.0,
                          (0.4, 811, 8): (FETCH_SCALE * 60490.2 + 16388.0) * 1024.0,
                          (0.4, 811, 10): (FETCH_SCALE * 75927.1 + 20483.4) * 1024.0,
                          (0.4, 812, 10): (FETCH_SCALE * 75931.4 + 20480.3) * 1024.0,
                          (0.5, 811, 10): (FETCH_SCALE * 78359.7 + 20488.6) * 1024.0,
                          (0.5, 812, 10): (FETCH_SCALE * 78363.0 + 20482.5) * 1024.0}
# the same at N=1,048,576 (workload c5), keyed by (skin, lanes, list entry bits): 32-bit entries at skin
# 0.5 with 8 lanes and skin 0.7 with the 816 shape, 16-bit entries (the default at this size) at skins
# 0.5 and 0.6 with 8 lanes and skin 0.7 with the 816 shape (profiles/lj_list_sweep_N1048576.md)
SWEEP_TRAFFIC_PROFILED_C5 = {(0.5, 8, 32): (FETCH_SCALE * 322982.0 + 65536.0) * 1024.0,
                             (0.7, 816, 32): (FETCH_SCALE * 323121.7 + 65536.3) * 1024.0,
                             (0.5, 8, 16): (FETCH_SCALE * 229185.4 + 65536.0) * 1024.0,
                             (0.6, 8, 16): (FETCH_SCALE * 229220.5 + 65536.2) * 1024.0,
                             (0.7, 816, 16): (FETCH_SCALE * 229060.2 + 65536.0) * 1024.0}
# the same for the 256-chain batch (workload c3) at the per-chain skin families base 0.4, exponent 1.5
# and base 0.45, exponent 1.0, 4 lanes per atom (profiles/lj_list_sweep_c3_batch.md), keyed by the skin
# family the run uses
SWEEP_TRAFFIC_PROFILED_C3 = {("0.4/1.5", 4): (FETCH_SCALE * 115006.0 + 38233.0) * 1024.0,
                             ("0.45/1.0", 4): (FETCH_SCALE * 116471.5 + 38304.6) * 1024.0}


def _tune(args, e, run, workload, **kw):
    """The skin search (tune.tune_skins), or with --tune-cache its stored outcome for the same problem
    (tune.cached_tune_skins).  Returns (skins, table, how the settings were found)."""
    from chemhmc_amd import tune
    if args.tune_cache:
        skins, table, hit = tune.cached_tune_skins(e, run, args.tune_cache, workload, **kw)
        return skins, table, "cache" if hit else "search (stored)"
    skins, table = tune.tune_skins(e, run, **kw)
    return skins, table, "search"


def _omp_threads(n):
    """Set the OpenMP team size of the oracle library (libgomp) for the next parallel regions."""
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
        return True
    except OSError:
        return False


def _cpu_model():
    """The host CPU's model name (lscpu's 'Model name', read from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(seconds_budget=20.0):
    """FP64 oracle (OpenMP) on the host, on the N=32768 workload: OMF2 trajectories with the O(N)
    counting-sort binning on all cores until ~budget (the reported value), one trajectory on one core,
    and one trajectory with the reference's O(N * cells) count/set binning (its cost model) on all
    cores."""
    import oracle
    L = (N / RHO) ** (1 / 3)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    s = oracle.Sim(N, [L] * 3, cutoff=RC, beta=1 / T)
    s.x[:] = oracle.lattice_init(N, [L] * 3, 1, 0, 0.05)
    ntraj = 0
    t0 = time.perf_counter()
    while True:
        s.hmc(1, 0, oracle.OMF2, DT, NSTEPS, lam=LAM, seed=1 + ntraj)
        ntraj += 1
        el = time.perf_counter() - t0
        if el > seconds_budget or ntraj >= 2000:
            break
    out = {"value": N * NSTEPS * ntraj / el, "unit": "atom-steps/s", "cores": threads, "kind": "port",
           "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
           "sample": f"{ntraj} OMF2 trajectories of the N=32768 workload (FP64 oracle, O(N) binning), {el:.1f} s"}
    if _omp_threads(1):
        t0 = time.perf_counter()
        s.hmc(1, 0, oracle.OMF2, DT, NSTEPS, lam=LAM, seed=999)
        el1 = time.perf_counter() - t0
        out["one_core"] = {"value": N * NSTEPS / el1, "sample": f"1 trajectory on 1 core, {el1:.1f} s"}
        _omp_threads(threads)
    r = oracle.Sim(N, [L] * 3, cutoff=RC, beta=1 / T, refalgo=True)
    r.x[:] = s.x
    t0 = time.perf_counter()
    r.hmc(1, 0, oracle.OMF2, DT, NSTEPS, lam=LAM, seed=998)
    elr = time.perf_counter() - t0
    out["ref_algo"] = {"value": N * NSTEPS / elr, "cores": threads,
                       "sample": f"1 trajectory with the reference's O(N*cells) binning (22 binnings), {elr:.1f} s"}
    return out


def other_workload(args):
    """--workload c3 (256 batched chains, LEAP 15 x 0.005) or c5 (N = 1,048,576, OMF2 10 x 0.01): the
    same measurement on those configurations, one JSON line."""
    import chemhmc_amd as chmc
    from chemhmc_amd import configs, stats, tune

    w = configs.WORKLOADS[args.workload]()
    e = chmc.Engine(w.chains)
    e.set_list_entry_bits(args.entry_bits)
    e.init_lattice(1, 0.05)
    e.set_list_check(args.list_check)
    nat = w.natoms
    steps = args.steps if args.steps_given else (200 if args.workload == "c3" else 50)
    # both workloads relax from their lattice starts over the first ~100 trajectories (more rebuilds
    # early): 150 accept-all trajectories (--therm for counter runs) whatever --warmup says, then --warmup
    # ordinary ones; the tuners and the timed run start after that
    therm = args.therm if args.therm_given else 150
    warm = args.warmup if args.warmup_given else 0
    t0 = time.perf_counter()
    e.hmc(therm, therm, w.scheme, w.dt, w.steps, lam=w.lam, seed=10, use_graph=args.graph)
    if warm > 0:
        e.hmc(warm, 0, w.scheme, w.dt, w.steps, lam=w.lam, seed=11, use_graph=args.graph)
    t_warm = time.perf_counter() - t0
    # skin autotuner (wall time of a few trajectories per candidate, paired from the warmed-up state,
    # which it restores), unless --skin fixes one skin for every chain or --skin-family BASE/EXP fixes
    # the per-chain skins (profiling runs): a batch of chains searches per-chain skins scaled with the
    # chain's density, a single chain one skin; lanes per atom tuned unless --lanes fixes them
    skins = {}
    family = None
    tuner = "fixed (--skin-family)" if args.skin_family is not None else "fixed (--skin)"
    if args.skin_family is not None:
        family = args.skin_family
        fb, fe = (float(t) for t in family.split("/"))
        e.set_chain_skins(tune.density_scaled_skins(w.chains, fb, fe))
        e.set_list_lanes(args.lanes)
    elif args.skin is None:
        def run(n, sd):
            e.hmc(n, 0, w.scheme, w.dt, w.steps, lam=w.lam, seed=sd, use_graph=args.graph)
        checks = None if args.list_check_given else (1, 2, 3)
        desc = {"workload": args.workload, "scheme": w.scheme, "dt": w.dt, "steps": w.steps, "lam": w.lam,
                "graph": bool(args.graph), "list_check": args.list_check if args.list_check_given else None}
        if len(w.chains) > 1:
            _, skins, tuner = _tune(args, e, run, desc, lanes=args.lanes, checks=checks)
        else:
            _, skins, tuner = _tune(args, e, run, desc, bases=(0.3, 0.4, 0.5, 0.6, 0.7, 0.8), exponents=(0.0,),
                                    lanes=args.lanes, checks=checks)
        family = "/".join(min(skins, key=skins.get).split("/")[:2])
    else:
        e.set_skin(args.skin)
        e.set_list_lanes(args.lanes)
    chain_skins = np.array([e.chain_skin(k) for k in range(len(w.chains))])
    t0 = time.perf_counter()
    rec, nacc = e.hmc(steps, 0, w.scheme, w.dt, w.steps, lam=w.lam, seed=12, use_graph=args.graph)
    el = time.perf_counter() - t0
    value = nat * w.steps * steps / el
    ent = nint = 0
    for k in range(len(w.chains)):
        a, b, _ = e.list_stats(k)
        ent += a
        nint += b
    ms_sweep, ms_rebuild = e.time_list(reps=20)
    bits = e.list_entry_bits()
    flops = 8.0 * ent + 13.0 * nint
    nbytes = bits / 8.0 * ent + SWEEP_BYTES_PER_ATOM * nat
    achieved = nbytes / (ms_sweep * 1e-3) / 1e9
    # the FP32 work of the 27-cell sweep the lists replace (candidates and interacting pairs of the cell
    # path, counted on the device at the final positions)
    ncell = sum(e.pair_counts(k)[0] for k in range(len(w.chains)))
    flops_cell = 8.0 * ncell + 13.0 * nint
    tflops_cell = flops_cell / (ms_sweep * 1e-3) / 1e12
    u = rec["u"] / np.array([c.N for c in w.chains])[None, :]
    # chains whose U/N never changed in the timed run carry no tau_int and contribute no ESS
    taus = [stats.tau_int(u[:, k])[0] for k in range(len(w.chains))]
    moving = [t for t in taus if np.isfinite(t)]
    ess = sum(steps / (2 * t) for t in moving)
    out = {
        "metric": "atom-steps/s (LJ force+integrator) and ESS/s of U/N", "value": value, "unit": "atom-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": warm, "ms_per_step": el / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 pair forces, f64 x/p/dH",
        "data": "synthetic (seeded lattice + jitter, Philox momenta)",
        "config": {"workload": f"{args.workload}: {len(w.chains)} chain(s), {nat} atoms, "
                               f"{'LEAP' if w.scheme == chmc.LEAPFROG else 'OMF2'} {w.steps}x{w.dt}",
                   "atoms": nat, "chains": len(w.chains), "list_lanes": e.list_lanes(), "list_check": e.list_check(),
                   "list_entry_bits": bits, "skin": e.skin(),
                   "chain_skin_range": [float(chain_skins.min()), float(chain_skins.max())], "skin_family": family,
                   "ms_per_traj_by_skin": skins, "tuner": tuner, "warmup_seconds": t_warm},
        "acceptance": float(nacc.sum()) / (steps * len(w.chains)),
        "ess_per_s_sum_over_chains": ess / el if moving else None, "chains_with_tau_int": len(moving),
        "roofline": {"bound": "hbm", "kernel": "k_lsweep<force|kick|drift> (FP32 list sweep)", "achieved": achieved,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                     "traffic": (SWEEP_TRAFFIC_PROFILED_C5.get((round(float(e.skin()), 2), int(e.list_lanes()), bits))
                                 if args.workload == "c5" else
                                 SWEEP_TRAFFIC_PROFILED_C3.get((family, int(e.list_lanes())))
                                 if args.workload == "c3" else None),
                     "bytes_per_launch": nbytes, "bytes": f"{bits // 8} B per list entry + {SWEEP_BYTES_PER_ATOM:.0f} B per atom",
                     "flops_per_launch": flops, "tflops": flops / (ms_sweep * 1e-3) / 1e12,
                     "fp32_27cell": {"work": "8 * N_cand + 13 * N_int per atom-evaluation, N_cand of the 27-cell stencil",
                                     "flops_per_launch": flops_cell, "tflops": tflops_cell,
                                     "frac": tflops_cell / PEAK_FP32_TFLOPS,
                                     "cell_candidates_per_atom": ncell / nat},
                     "ms_per_launch": ms_sweep, "ms_rebuild": ms_rebuild,
                     "list_entries_per_atom": ent / nat, "n_int_per_atom": nint / nat},
    }
    print(json.dumps(out))


def surrogate_c5_workload(args):
    """--workload c5s: the c4 surrogate fine-tuned with energy + force matching on 200 snapshots of the
    N = 1,048,576 system (rho*=0.8442, T*=0.722; 5 GB of positions held on the device), then
    surrogate-driven HMC at that size against plain HMC from the same state (OMF2 10 x 0.01): acceptance
    and ms per trajectory of both, and the fine-tuning losses."""
    import chemhmc_amd as chmc
    from chemhmc_amd import configs

    params, _, _, _, _ = train_c4_surrogate(args, chmc)
    w = configs.WORKLOADS["c5"]()
    nat = w.natoms
    e = chmc.Engine(w.chains)
    e.init_lattice(1, 0.05)
    warm = 10                   # accept-all trajectories off the lattice start
    e.hmc(warm, warm, w.scheme, w.dt, w.steps, lam=w.lam, seed=11, use_graph=args.graph)
    if args.warmup_given and args.warmup > 0:
        e.hmc(args.warmup, 0, w.scheme, w.dt, w.steps, lam=w.lam, seed=9, use_graph=args.graph)
        warm += args.warmup
    e.set_skin(0.5)
    e.surrogate_set_params(params)
    nconf = 200
    e.surrogate_reserve_configs(nconf)
    t0 = time.perf_counter()
    for k in range(nconf):
        e.hmc(1, 10 ** 9, w.scheme, w.dt, w.steps, lam=w.lam, seed=200 + k, use_graph=False)
        e.surrogate_collect_config()
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    le, lf = e.surrogate_train_fm(args.finetune_steps, C4_W_FORCE, lr=3e-4, seed=7)
    t_train = time.perf_counter() - t0
    e.hmc(2, 10 ** 9, w.scheme, w.dt, w.steps, lam=w.lam, seed=299, use_graph=False)     # off the last snapshot
    x0 = e.positions()
    steps = args.steps if args.steps_given else 20
    out = {}
    for name, surr in (("plain", False), ("surrogate", True)):
        e.set_positions(x0)
        e.set_trajectory(warm + nconf + 2)
        e.kinetic()
        t0 = time.perf_counter()
        rec, nacc = e.hmc(steps, 0, w.scheme, w.dt, w.steps, lam=w.lam, seed=12, use_graph=args.graph,
                          surrogate=surr, fallback_window=0)
        el = time.perf_counter() - t0
        out[name] = {"acceptance": float(nacc[0]) / steps, "ms_per_traj": el / steps * 1e3,
                     "u_per_n": float(rec["u"][:, 0].mean() / nat), "value": nat * w.steps * steps / el}
    res = {
        "metric": "atom-steps/s (LJ force+integrator) and ESS/s of U/N", "value": out["surrogate"]["value"],
        "unit": "atom-steps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": out["surrogate"]["ms_per_traj"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 MFMA surrogate MD, f64 LJ accept", "data": "synthetic (seeded lattice + jitter, Philox momenta)",
        "config": {"workload": "c5s: surrogate HMC N=1048576 rho=0.8442 T=0.722 OMF2(0.1932) 10x0.01; the c4 surrogate "
                               "fine-tuned on 200 configurations of this system",
                   "base_train_configs": args.train_configs, "finetune_configs": nconf, "finetune_steps": args.finetune_steps,
                   "w_force": C4_W_FORCE, "finetune_lr": 3e-4, "collect_seconds": t_collect, "finetune_seconds": t_train,
                   "energy_mse_first": float(le[:20].mean()), "energy_mse_last": float(le[-100:].mean()),
                   "force_mse_first": float(lf[:20].mean()), "force_mse_last": float(lf[-100:].mean())},
        "plain": out["plain"], "surrogate": out["surrogate"], "force_evaluation_kernels": kernels,
        "ms_per_traj_ratio_surrogate_over_plain": out["surrogate"]["ms_per_traj"] / out["plain"]["ms_per_traj"],
    }
    print(json.dumps(res))


C4_NCONF = 5000                 # training snapshots of the N=32768 plain-HMC chain
C4_W_FORCE = 1.0                # weight of the force-matching term in the training loss


def train_c4_surrogate(args, chmc):
    """The surrogate of the c4 benchmark: 32-64-64-1 on 32 radial descriptors, trained on the GPU with
    energy + force matching on C4_NCONF snapshots of the plain-HMC chain of the N=32768 system itself
    (rho*=0.8442, T*=0.722, OMF2 10 x 0.01): one snapshot after each accept-all trajectory (the
    thermalisation phase of the reference's loop) following the burn-in off the lattice.  Returns the
    parameters, the collect and train seconds and the two loss histories."""
    L = (N / RHO) ** (1 / 3)
    rng = np.random.default_rng(4)
    tr = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T))
    tr.init_lattice(3, 0.05)
    rec, _ = tr.hmc(THERM, THERM + 1, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=5)          # accept-all burn-in
    p0 = {"W1": rng.normal(0, 1 / np.sqrt(32), (64, 32)), "b1": np.zeros(64),
          "W2": rng.normal(0, 1 / np.sqrt(64), (64, 64)), "b2": np.zeros(64),
          "w3": rng.normal(0, 1 / np.sqrt(64), 64), "b3": np.zeros(1), "shift": np.zeros(32), "scale": np.ones(32)}
    tr.surrogate_set_params(p0)
    tr.surrogate_eval()
    G, _ = tr.surrogate_features()
    p0["shift"] = G.mean(0).astype(np.float64)
    p0["scale"] = 1.0 / (G.std(0).astype(np.float64) + 1e-3)
    p0["b3"] = np.array([float(rec["u"][-1, 0]) / N])
    tr.surrogate_set_params(p0)
    nconf = args.train_configs
    tr.surrogate_reserve_configs(nconf)
    t0 = time.perf_counter()
    for k in range(nconf):
        tr.hmc(1, 10 ** 9, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=1000 + k, use_graph=False)
        tr.surrogate_collect_config()
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    le, lf = tr.surrogate_train_fm(args.train_steps, C4_W_FORCE, lr=1e-3, seed=6)
    t_train = time.perf_counter() - t0
    params = tr.surrogate_params()
    tr.close()
    return params, t_collect, t_train, le, lf


def surrogate_workload(args):
    """--workload c4: the learned surrogate (trained by train_c4_surrogate) drives the MD of HMC at
    N=32768, rho*=0.8442, T*=0.722 (OMF2 10 x 0.01) with the exact FP64 LJ accept step, against plain
    HMC from the same state and random streams: acceptance, ms per trajectory and ESS/s of U/N of both."""
    import chemhmc_amd as chmc
    from chemhmc_amd import stats

    params, t_collect, t_train, le, lf = train_c4_surrogate(args, chmc)
    L = (N / RHO) ** (1 / 3)
    # both legs run the chain as args.spec speculative copies (chmc_set_speculation), as the c2 line does
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T), speculate=args.spec)
    e.init_lattice(1, 0.05)
    e.set_list_check(args.list_check)
    e.set_list_start_rebuild(bool(args.start_rebuild))
    e.hmc(THERM, THERM, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=10)
    if args.warmup > 0:
        e.hmc(args.warmup, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=11)
    e.surrogate_set_params(params)
    x0 = e.positions()
    t_start = e.trajectory()
    steps = args.steps if args.steps_given else 2000
    out = {}
    for name, surr in (("plain", False), ("surrogate", True)):
        e.set_positions(x0)
        e.set_trajectory(t_start)
        e.hmc(2, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=12, use_graph=args.graph, surrogate=surr)   # graph warm-up
        e.set_positions(x0)
        e.set_trajectory(t_start)
        e.kinetic()
        t0 = time.perf_counter()
        rec, nacc = e.hmc(steps, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=12, use_graph=args.graph,
                          surrogate=surr, fallback_window=0)
        el = time.perf_counter() - t0
        uu = rec["u"][:, 0] / N
        tau, _ = stats.tau_int(uu)
        ok = bool(np.isfinite(tau))
        out[name] = {"acceptance": float(nacc[0]) / steps, "ms_per_traj": el / steps * 1e3,
                     "mean_dh": float(rec["dh"][:, 0].mean()), "median_abs_dh": float(np.median(np.abs(rec["dh"][:, 0]))),
                     "u_per_n": float(uu.mean()), "tau_int_u": tau if ok else None,
                     "ess_per_s": steps / (2 * tau) / el if ok else None, "reliable": stats.reliable(uu),
                     "value": N * NSTEPS * steps / el}
    # per-launch device times of one force evaluation over every speculative copy (HIP events, 50 launches
    # each at the current positions): the surrogate's descriptor + MLP and force sweep, the LJ list sweep
    ms_desc, ms_ssweep = e.time_surrogate(reps=50)
    ms_lj, _ = e.time_list(reps=50)
    kernels = {"atoms_per_launch": N * e.speculation(), "ms_desc_mlp": ms_desc, "ms_surrogate_sweep": ms_ssweep,
               "ms_lj_list_sweep": ms_lj, "ratio_per_force_evaluation": (ms_desc + ms_ssweep) / ms_lj}
    res = {
        "metric": "atom-steps/s (LJ force+integrator) and ESS/s of U/N", "value": out["surrogate"]["value"],
        "unit": "atom-steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": out["surrogate"]["ms_per_traj"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 MFMA surrogate MD, f64 LJ accept", "data": "synthetic (seeded lattice + jitter, Philox momenta)",
        "config": {"workload": "c4: surrogate HMC N=32768 rho=0.8442 T=0.722 OMF2(0.1932) 10x0.01; surrogate 32-64-64-1 "
                               "trained with energy + force matching on snapshots of the same system's plain-HMC chain",
                   "speculative_copies": e.speculation(), "start_rebuild": e.list_start_rebuild(),
                   "train_configs": args.train_configs, "train_steps": args.train_steps, "w_force": C4_W_FORCE, "lr": 1e-3,
                   "collect_seconds": t_collect, "train_seconds": t_train,
                   "energy_mse_first": float(le[:20].mean()), "energy_mse_last": float(le[-100:].mean()),
                   "force_mse_first": float(lf[:20].mean()), "force_mse_last": float(lf[-100:].mean())},
        "plain": out["plain"], "surrogate": out["surrogate"], "force_evaluation_kernels": kernels,
        "ms_per_traj_ratio_surrogate_over_plain": out["surrogate"]["ms_per_traj"] / out["plain"]["ms_per_traj"],
        "ess_per_s_ratio_surrogate_over_plain": (out["surrogate"]["ess_per_s"] / out["plain"]["ess_per_s"]
                                                  if out["surrogate"]["ess_per_s"] and out["plain"]["ess_per_s"]
                                                  else None),
    }
    print(json.dumps(res))


def cell_md_run(args, e, chmc, stats):
    """The same c2 measurement through the 27-cell sweep with a re-binning per force evaluation (the
    reference's cost model, CHMC_CELL_MD) instead of the neighbour lists."""
    e.hmc(THERM, THERM, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=10, use_graph=args.graph)
    if args.warmup > 0:
        e.hmc(args.warmup, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=11, use_graph=args.graph)
    e.set_force_tiling(0)
    t0 = time.perf_counter()
    rec, nacc = e.hmc(args.steps, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=12, use_graph=args.graph)
    el = time.perf_counter() - t0
    u = rec["u"][:, 0] / N
    tau, _ = stats.tau_int(u)
    tau = tau if np.isfinite(tau) else None
    ncand, nint = e.pair_counts(0)
    flops = 8.0 * ncand + 13.0 * nint
    ms_bin, ms_force = e.time_force(reps=50)
    achieved = flops / (ms_force * 1e-3) / 1e12
    out = {"metric": "atom-steps/s (LJ force+integrator) and ESS/s of U/N", "value": N * NSTEPS * args.steps / el,
           "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32 pair forces, f64 x/p/dH", "data": "synthetic (seeded lattice + jitter, Philox momenta)",
           "config": {"workload": "plain HMC N=32768 rho=0.8442 T=0.722 rc=2.5 OMF2(lambda=0.1932) 10x0.01",
                      "md_path": "27-cell sweep, re-binning per evaluation", "force_tiling": e.force_tiling()},
           "acceptance": float(nacc[0]) / args.steps, "tau_int_u": tau,
           "ess_per_s": args.steps / (2 * tau) / el if tau else None,
           "roofline": {"bound": "valu", "kernel": "k_lj<float> force+kick sweep", "achieved": achieved,
                        "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                        "traffic": None, "flops_per_launch": flops, "ms_per_launch": ms_force, "ms_binning": ms_bin,
                        "n_cand_per_atom": ncand / N, "n_int_per_atom": nint / N}}
    print(json.dumps(out))


def kernels_workload(args):
    """--workload kernels: each kernel of the hot path exercised on the N=32768 workload so that a
    kernel trace (rocprofv3) times them all in one run: the O(N) binning and the 27-cell FP32 sweep
    (chmc_time_force), the RDF in 500 bins, graph-replayed HMC trajectories (heat bath, list sweeps and
    rebuilds, FP64 energy, Metropolis, accept/restore) and the surrogate's descriptor, MLP and force
    kernels.  Prints one JSON line with the host-visible timings and the algorithmic units per launch
    (atoms, pairs) that the per-kernel rates divide by."""
    import chemhmc_amd as chmc

    L = (N / RHO) ** (1 / 3)
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T))
    e.init_lattice(1, 0.05)
    e.hmc(args.therm, args.therm, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=10)
    out = {"workload": "kernels at N=32768 rho=0.8442 T=0.722", "N": N}
    ms_bin, ms_force = e.time_force(reps=args.reps)
    ncand, nint = e.pair_counts(0)
    out["binning_ms"], out["cell_sweep_ms"] = ms_bin, ms_force
    out["cell_sweep_flops"] = 8.0 * ncand + 13.0 * nint
    nb = 500
    t0 = time.perf_counter()
    for _ in range(args.reps):
        e.rdf(nb, L / (2 * nb))
    out["rdf_ms_host"] = (time.perf_counter() - t0) / args.reps * 1e3
    out["rdf_pairs"] = N * (N - 1) / 2
    t0 = time.perf_counter()
    e.hmc(args.reps, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=11)
    out["hmc_ms_per_traj_host"] = (time.perf_counter() - t0) / args.reps * 1e3
    rng = np.random.default_rng(1)
    p = {"W1": rng.normal(0, 1 / np.sqrt(32), (64, 32)), "b1": np.zeros(64),
         "W2": rng.normal(0, 1 / np.sqrt(64), (64, 64)), "b2": np.zeros(64),
         "w3": rng.normal(0, 1 / np.sqrt(64), 64), "b3": np.zeros(1), "shift": np.zeros(32), "scale": np.ones(32)}
    e.surrogate_set_params(p)
    t0 = time.perf_counter()
    for _ in range(args.reps):
        e.surrogate_eval(store_force=True)
    out["surrogate_eval_ms_host"] = (time.perf_counter() - t0) / args.reps * 1e3
    out["surrogate_mlp_flops_fwd"] = 2.0 * (32 * 64 + 64 * 64 + 64) * N
    # device-to-device copies of a known size (the positions, 3 N doubles each): the calibration of the
    # FETCH_SIZE / WRITE_SIZE counters against a plain copy
    e.surrogate_reserve_configs(4)
    for _ in range(4):
        e.surrogate_collect_config()
    out["calibration_copy_bytes"] = 3 * N * 8
    e.close()
    print(json.dumps(out))


def calib_workload(args):
    """Calibration of the HBM counters: four device-to-device copies of 1 GiB, each source written long
    before (the other buffers evict it from the L2 and the MALL), so every byte is read from HBM once and
    written once; profiled with rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE."""
    import torch
    n = 1 << 28                                    # FP32 elements: 1 GiB per buffer
    bufs = [torch.full((n,), float(i + 1), device="cuda") for i in range(4)]
    torch.cuda.synchronize()
    for src, dst in ((0, 1), (2, 3), (1, 0), (3, 2)):
        bufs[dst].copy_(bufs[src])
        torch.cuda.synchronize()
    print(json.dumps({"workload": "calib", "copies": 4, "bytes_per_copy": 4 * n}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed HMC trajectories")
    ap.add_argument("--warmup", type=int, default=100, help="untimed trajectories (also thermalisation)")
    ap.add_argument("--tune", type=int, default=200,
                    help="trajectories per skin in the skin autotuner (long enough that the graph capture of "
                         "each timed call is a small part of it)")
    ap.add_argument("--tune-cache", default=os.environ.get("CHMC_TUNE_CACHE"),
                    help="JSON file of tuned settings (default $CHMC_TUNE_CACHE, unset: search every run): a run "
                         "of a problem tuned before (same chains, workload, candidates and library build) takes "
                         "the stored skins, check interval and lanes without searching; otherwise the search's "
                         "outcome is added to it")
    ap.add_argument("--skin", type=float, default=None, help="fixed list skin (no skin autotuning)")
    ap.add_argument("--skin-family", default=None,
                    help="fixed per-chain skins BASE/EXP of the density-scaled family (c3, c5; no skin autotuning)")
    ap.add_argument("--lanes", type=int, default=0, choices=[0, 4, 8, 16, 810, 811, 812, 816],
                    help="fixed list-sweep lanes per atom (0: autotune)")
    ap.add_argument("--entry-bits", type=int, default=0, choices=[0, 16, 32],
                    help="list entry format (chmc_set_list_entry_bits): 0 chooses by the lists' size")
    ap.add_argument("--list-check", type=int, default=2, choices=[1, 2, 3, 4],
                    help="list MD: rebuild checks before every LIST_CHECK-th force evaluation (chmc_set_list_check); "
                         "tuned with the skin (1, 2 or 3) unless given; 2 where the skin is fixed")
    ap.add_argument("--spec", type=int, default=0, choices=[0, 1, 2, 4, 8, 10, 16],
                    help="c2, c4: speculative copies of the chain (chmc_set_speculation; 1: the sequential loop; 0: "
                         "8, or for runs shorter than 160 trajectories the count that splits them into whole "
                         "batches of at most 10).  The chain, its records and its acceptance are the sequential "
                         "run's, bit for bit")
    ap.add_argument("--start-rebuild", type=int, default=1, choices=[0, 1],
                    help="c2, c4: rebuild the lists at the start of the trajectory after an accepted one "
                         "(chmc_set_list_start_rebuild; pays at low acceptance)")
    ap.add_argument("--therm", type=int, default=THERM,
                    help="accept-all burn-in trajectories off the lattice (c2, kernels; c3 and c5 default to 150; "
                         "shorter for counter runs)")
    ap.add_argument("--no-ess", action="store_true",
                    help="c2: ESS from the timed run itself, no separate long run (counter and trace runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--graph", action="store_true",
                    help="replay captured hipGraphs of the trajectories instead of launching them directly (each "
                         "hmc call captures its own: ~3 ms with eight speculative copies, no faster per trajectory)")
    ap.add_argument("--no-graph", action="store_true", help="direct launches (the default; overrides --graph)")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5", "c5s", "kernels", "calib"],
                    help="c2 (default, the headline line), c3 batched chains, c4 surrogate vs plain, c5 N=1M, "
                         "c5s the surrogate fine-tuned and run at N=1M, kernels: every hot-path kernel a few "
                         "times at N=32768 (for kernel traces), calib: 1 GiB device copies (HBM counter calibration)")
    ap.add_argument("--reps", type=int, default=20, help="kernels: repetitions of each kernel group")
    ap.add_argument("--train-steps", type=int, default=4000, help="c4, c5s: Adam steps of the surrogate")
    ap.add_argument("--train-configs", type=int, default=C4_NCONF, help="c4: training snapshots of the N=32768 chain")
    ap.add_argument("--finetune-steps", type=int, default=1000, help="c5s: Adam steps of the fine-tuning")
    ap.add_argument("--cell-md", action="store_true",
                    help="c2 through the 27-cell sweep with a re-binning per evaluation (CHMC_CELL_MD)")
    args = ap.parse_args()
    args.steps_given = "--steps" in sys.argv
    args.warmup_given = "--warmup" in sys.argv
    args.list_check_given = "--list-check" in sys.argv
    args.therm_given = "--therm" in sys.argv
    args.graph = args.graph and not args.no_graph
    if args.spec == 0:
        # a batch runs all its copies whether the run needs them or not: a short run takes the copy count
        # that divides it into whole batches (20 trajectories: 2 batches of 10 instead of 8 + 8 + 4)
        nb = -(-args.steps // 10)
        args.spec = 8 if args.steps >= 160 else -(-args.steps // nb)
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload == "kernels":
        kernels_workload(args)
        return
    if args.workload == "calib":
        calib_workload(args)
        return
    if args.workload == "c4":
        surrogate_workload(args)
        return
    if args.workload == "c5s":
        surrogate_c5_workload(args)
        return
    if args.workload != "c2":
        other_workload(args)
        return

    import chemhmc_amd as chmc
    from chemhmc_amd import stats, tune

    L = (N / RHO) ** (1 / 3)
    copies = 1 if args.cell_md else args.spec
    e = chmc.Engine(chmc.Chain(N=N, L=L, cutoff=RC, beta=1 / T), cell_md=args.cell_md, speculate=copies)
    if not args.cell_md:
        e.set_list_entry_bits(args.entry_bits)
    e.init_lattice(1, 0.05)
    e.set_list_check(args.list_check)
    e.set_list_start_rebuild(bool(args.start_rebuild))
    if args.cell_md:
        cell_md_run(args, e, chmc, stats)
        return
    # accept-all thermalisation away from the lattice (thermalization_steps), then the untimed warm-up
    e.hmc(args.therm, args.therm, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=10, use_graph=args.graph)
    if args.warmup > 0:
        e.hmc(args.warmup, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=11, use_graph=args.graph)
    # skin autotuner: wall time of a short run per skin from the equilibrated state (paired: every skin
    # starts from that state with the same random streams, and the state is restored afterwards, so the
    # timed run continues from it whatever the tuner ran), with the lanes per atom of the list sweep
    # (0: autotuned) chosen for every skin
    skins = {}
    tuner = "fixed (--skin)"
    if args.skin is None:
        def run(n, sd):
            e.hmc(n, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=sd, use_graph=args.graph)
        _, table, tuner = _tune(args, e, run, {"workload": "c2", "scheme": "OMF2", "dt": DT, "steps": NSTEPS, "lam": LAM,
                                               "graph": bool(args.graph), "start_rebuild": bool(args.start_rebuild),
                                               "list_check": args.list_check if args.list_check_given else None},
                                bases=(0.3, 0.4, 0.5, 0.6, 0.7, 0.8), exponents=(0.0,), reps=args.tune,
                                lanes=args.lanes, checks=None if args.list_check_given else (1, 2, 3))
        skins = {k.split("/")[0] + ("/" + k.split("/")[2] if k.count("/") == 2 else ""): v for k, v in table.items()}
    else:
        e.set_skin(args.skin)
        e.set_list_lanes(args.lanes)
    skin = e.skin()
    lanes = e.list_lanes()
    _, _, rebuilds0 = e.list_stats()

    t0 = time.perf_counter()
    rec, nacc = e.hmc(args.steps, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=12, use_graph=args.graph)
    el = time.perf_counter() - t0
    ms = el / args.steps * 1e3
    value = N * NSTEPS * args.steps / el
    entries, nint, rebuilds1 = e.list_stats()
    # list_stats itself rebuilds every copy once; with speculative copies every copy's rebuilds count
    rebuilds_per_traj = (rebuilds1 - rebuilds0 - copies) / args.steps

    # ESS/s of U/N on a run long enough for tau_int (the timed run itself when it is)
    if args.no_ess:
        ess_rec, ess_acc, ess_el = rec, nacc, el
    elif args.steps >= ESS_TRAJ:
        ess_rec, ess_acc, ess_el = rec, nacc, el
    else:
        t0 = time.perf_counter()
        ess_rec, ess_acc = e.hmc(ESS_TRAJ, 0, chmc.OMF2, DT, NSTEPS, lam=LAM, seed=13, use_graph=args.graph)
        ess_el = time.perf_counter() - t0
    u = ess_rec["u"][:, 0] / N
    nt = u.size
    tau, window, closed = stats.tau_window(u)
    informative = bool(np.isfinite(tau))
    ess_info = {"trajectories": nt, "seconds": ess_el, "acceptance": float(ess_acc[0]) / nt,
                "u_per_n": float(u.mean()), "tau_int_u": tau if informative else None,
                "sokal_window": window, "window_closed": closed, "reliable": stats.reliable(u),
                "ess_per_s": nt / (2 * tau) / ess_el if informative else None,
                "run": "the timed run" if ess_rec is rec else "a separate run after the timed one"}
    if not informative:
        ess_info["note"] = "U/N never changed (no accepted trajectory): tau_int and ESS undefined"

    # dominant kernel: the neighbour-list FP32 sweep (force + kick + drift), against the FP32 VALU peak with
    # the algorithmic work of the 27-cell sweep it replaces (8 flops per candidate pair of the 27-cell
    # stencil + 13 per interacting pair, per atom-evaluation; candidates and interacting pairs counted on
    # the device).  Its HBM side: the list entry bytes (4 B, or 2 in the 16-bit format) + SWEEP_BYTES_PER_ATOM
    # per atom algorithmic, and the rocprofv3-measured traffic when the tuned skin and lanes are the profiled
    # ones (32-bit entries).
    # (a launch sweeps every speculative copy: copies x the one chain's work)
    ms_sweep, ms_rebuild = e.time_list(reps=50)
    ncand_cell, nint_cell = e.pair_counts(0)
    flops_alg = copies * (8.0 * ncand_cell + 13.0 * nint_cell)
    flops_exec = copies * (8.0 * entries + 13.0 * nint)
    ebytes = e.list_entry_bits() / 8.0
    nbytes = copies * (ebytes * entries + SWEEP_BYTES_PER_ATOM * N)
    tflops = flops_alg / (ms_sweep * 1e-3) / 1e12
    traffic = SWEEP_TRAFFIC_PROFILED.get((round(float(skin), 2), int(lanes), copies)) if ebytes == 4.0 else None

    out = {
        "metric": "atom-steps/s (LJ force+integrator) and ESS/s of U/N",
        "value": value, "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair forces, f64 x/p/dH", "data": "synthetic (seeded lattice + jitter, Philox momenta)",
        "config": {"workload": "plain HMC N=32768 rho=0.8442 T=0.722 rc=2.5 OMF2(lambda=0.1932) 10x0.01",
                   "N": N, "trajectories": args.steps, "md_steps_per_traj": NSTEPS, "force_evals_per_traj": 2 * NSTEPS + 1,
                   "thermalisation": args.therm, "md_path": "neighbour lists", "speculative_copies": copies,
                   "skin": skin, "list_lanes": lanes, "list_check": e.list_check(),
                   "list_entry_bits": e.list_entry_bits(), "start_rebuild": e.list_start_rebuild(),
                   "launches": "graph replay" if args.graph else "direct",
                   "ms_per_traj_by_skin": skins, "tuner": tuner},
        "acceptance": float(nacc[0]) / args.steps,
        "ess_per_s": ess_info["ess_per_s"], "ess": ess_info,
        "list_rebuilds_per_traj": rebuilds_per_traj,   # over all copies when speculative
        "roofline": {"bound": "valu", "kernel": "k_lsweep<force|kick|drift> (FP32 list sweep)", "achieved": tflops,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": tflops / PEAK_FP32_TFLOPS,
                     "traffic": traffic, "flops_per_launch": flops_alg, "ms_per_launch": ms_sweep,
                     "work": "8 * N_cand + 13 * N_int per atom-evaluation, N_cand of the 27-cell stencil",
                     "atoms_per_launch": copies * N,
                     "cell_candidates_per_atom": ncand_cell / N, "n_int_per_atom": nint_cell / N,
                     "executed_flops_per_launch": flops_exec, "list_entries_per_atom": entries / N,
                     "hbm": {"bytes_per_launch": nbytes, "achieved_gbs": nbytes / (ms_sweep * 1e-3) / 1e9,
                             "frac": nbytes / (ms_sweep * 1e-3) / 1e9 / PEAK_HBM_GBS,
                             "measured_traffic_gbs": traffic / (ms_sweep * 1e-3) / 1e9 if traffic else None},
                     "ms_rebuild": ms_rebuild},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
