"""chemhmc_amd -- MI355X-native Lennard-Jones HMC engine (Python front-end of libchmc.so).

Everything below calls the C ABI of ``include/chmc.h`` through ctypes; there is no Python or CPU
compute path.  Loading fails loudly when ``libchmc.so`` has not been built or when no GPU is
visible.  Arrays cross the boundary as numpy float64 (N, 3) C-contiguous (AoS per particle, the
layout of the reference's ``config<coordinate>``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libchmc.so")

LEAPFROG, LP_LEAPFROG, OMF4, LP_OMF4, EULER, RUTH, OMF2 = 0, 1, 2, 3, 4, 5, 6
FORCE_FP32, FORCE_FP64, ORDER_ATOM_ID, CELL_MD = 0, 1, 2, 4
OMF2_LAMBDA_REFERENCE = 0.1938       # code/modules/integrator.cpp:72

_D, _I, _U64 = C.c_double, C.c_int, C.c_uint64
_PD = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_PI = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_PU64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class ChainDesc(C.Structure):
    _fields_ = [("N", _I), ("L", _D * 3), ("cutoff", _D), ("sigma", _D), ("eps", _D), ("mass", _D), ("beta", _D)]


class HmcParams(C.Structure):
    _fields_ = [("ntraj", _I), ("therm", _I), ("scheme", _I), ("steps", _I), ("dt", _D), ("lam", _D),
                ("seed", _U64), ("use_graph", _I), ("randomize", _I), ("surrogate", _I),
                ("fallback_window", _I), ("fallback_acc", _D), ("uniforms", C.c_void_p), ("traj_steps", C.c_void_p)]


HMC_RECORD_DTYPE = np.dtype([("dh", "f8"), ("u", "f8"), ("k", "f8"), ("w", "f8"), ("accepted", "i4"),
                             ("surrogate", "i4")])


class ChmcError(RuntimeError):
    pass


def build(verbose: bool = False) -> None:
    """Compile libchmc.so for gfx950 with hipcc (in-tree)."""
    out = None if verbose else subprocess.DEVNULL
    subprocess.run(["make", "-s", "-C", _HERE, "-j8", "libchmc.so"], check=True, stdout=out)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ChmcError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    P = C.c_void_p
    sig = {
        "chmc_create": ([C.POINTER(P), _I, C.POINTER(ChainDesc), _I, _I], _I),
        "chmc_destroy": ([P], _I),
        "chmc_last_error": ([P], C.c_char_p),
        "chmc_sync": ([P], _I),
        "chmc_get_geometry": ([P, _I, C.POINTER(_I), C.POINTER(_D), C.POINTER(_I)], _I),
        "chmc_num_atoms": ([P, C.POINTER(_I), C.POINTER(_I)], _I),
        "chmc_set_positions": ([P, _I, _PD], _I),
        "chmc_get_positions": ([P, _I, _PD], _I),
        "chmc_set_momenta": ([P, _I, _PD], _I),
        "chmc_get_momenta": ([P, _I, _PD], _I),
        "chmc_init_lattice": ([P, _U64, _D], _I),
        "chmc_bin": ([P], _I),
        "chmc_get_bins": ([P, _I, _PI, _PI, _PI], _I),
        "chmc_compute_force": ([P, _I], _I),
        "chmc_get_force": ([P, _I, _PD], _I),
        "chmc_compute_potential": ([P, _PD], _I),
        "chmc_evaluate_potential": ([P, _PD], _I),
        "chmc_virial": ([P, _PD], _I),
        "chmc_kinetic": ([P, _PD], _I),
        "chmc_pair_counts": ([P, _I, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)], _I),
        "chmc_update_momenta": ([P, _D], _I),
        "chmc_update_positions": ([P, _D], _I),
        "chmc_kick": ([P, _D], _I),
        "chmc_integrate": ([P, _I, _D, _I, _D], _I),
        "chmc_set_lp_exponent": ([P, _I], _I),
        "chmc_get_lp_exponent": ([P, C.POINTER(_I)], _I),
        "chmc_heatbath": ([P, _U64, _U64], _I),
        "chmc_save_positions": ([P], _I),
        "chmc_restore_positions": ([P], _I),
        "chmc_hmc_run": ([P, C.POINTER(HmcParams), C.c_void_p, C.c_void_p], _I),
        "chmc_rdf": ([P, _I, _I, _D, C.c_void_p, C.c_void_p], _I),
        "chmc_potential_energy_open": ([_PD, _I, _I, _D, _D, C.POINTER(_D)], _I),
        "chmc_potential_energy_open_noalloc": ([C.c_void_p, _I, _I, _D, _D, C.c_void_p, C.c_void_p], _I),
        "chmc_set_force_tiling": ([P, _I], _I),
        "chmc_get_force_tiling": ([P, C.POINTER(_I)], _I),
        "chmc_time_force": ([P, _I, C.POINTER(_D), C.POINTER(_D)], _I),
        "chmc_set_skin": ([P, _D], _I),
        "chmc_get_skin": ([P, C.POINTER(_D)], _I),
        "chmc_set_chain_skins": ([P, _PD], _I),
        "chmc_get_chain_skin": ([P, _I, C.POINTER(_D)], _I),
        "chmc_list_stats": ([P, _I, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)], _I),
        "chmc_time_list": ([P, _I, C.POINTER(_D), C.POINTER(_D)], _I),
        "chmc_time_surrogate": ([P, _I, C.POINTER(_D), C.POINTER(_D)], _I),
        "chmc_set_list_lanes": ([P, _I], _I),
        "chmc_set_list_check": ([P, _I], _I),
        "chmc_get_list_check": ([P, C.POINTER(_I)], _I),
        "chmc_set_list_start_rebuild": ([P, _I], _I),
        "chmc_get_list_start_rebuild": ([P, C.POINTER(_I)], _I),
        "chmc_set_speculation": ([P, _I], _I),
        "chmc_get_speculation": ([P, C.POINTER(_I)], _I),
        "chmc_get_list_lanes": ([P, C.POINTER(_I)], _I),
        "chmc_set_list_entry_bits": ([P, _I], _I),
        "chmc_get_list_entry_bits": ([P, C.POINTER(_I)], _I),
        "chmc_set_trajectory": ([P, _I], _I),
        "chmc_get_trajectory": ([P, C.POINTER(_I)], _I),
        "chmc_surrogate_set_params": ([P, C.c_void_p, _I], _I),
        "chmc_surrogate_get_params": ([P, C.c_void_p, _I], _I),
        "chmc_surrogate_eval": ([P, C.c_void_p, _I], _I),
        "chmc_surrogate_features": ([P, _I, C.c_void_p, C.c_void_p], _I),
        "chmc_surrogate_reserve": ([P, _I], _I),
        "chmc_surrogate_collect": ([P, C.POINTER(C.c_longlong)], _I),
        "chmc_surrogate_training_data": ([P, C.c_longlong, _I, C.c_void_p, C.c_void_p], _I),
        "chmc_surrogate_loss_grad": ([P, C.c_longlong, _I, C.c_void_p], _I),
        "chmc_surrogate_train": ([P, _I, _I, C.c_float, _U64, C.c_void_p], _I),
        "chmc_surrogate_get_adam": ([P, C.c_void_p, C.c_void_p, C.POINTER(_I)], _I),
        "chmc_surrogate_set_adam": ([P, C.c_void_p, C.c_void_p, _I], _I),
        "chmc_surrogate_fm_loss_grad": ([P, _D, C.c_void_p, C.POINTER(_D), C.POINTER(_D)], _I),
        "chmc_surrogate_reserve_configs": ([P, _I], _I),
        "chmc_surrogate_collect_config": ([P, C.POINTER(_I)], _I),
        "chmc_surrogate_train_fm": ([P, _I, _D, C.c_float, _U64, C.c_void_p, C.c_void_p], _I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


EXPORTED = (
    "chmc_create", "chmc_destroy", "chmc_last_error", "chmc_sync", "chmc_get_geometry", "chmc_num_atoms",
    "chmc_set_positions", "chmc_get_positions", "chmc_set_momenta", "chmc_get_momenta", "chmc_init_lattice",
    "chmc_bin", "chmc_get_bins", "chmc_compute_force", "chmc_get_force", "chmc_compute_potential",
    "chmc_evaluate_potential", "chmc_virial", "chmc_pair_counts", "chmc_kinetic", "chmc_update_momenta", "chmc_update_positions",
    "chmc_kick", "chmc_integrate", "chmc_set_lp_exponent", "chmc_get_lp_exponent", "chmc_heatbath", "chmc_save_positions", "chmc_restore_positions",
    "chmc_hmc_run", "chmc_rdf", "chmc_potential_energy_open", "chmc_potential_energy_open_noalloc",
    "chmc_set_force_tiling",
    "chmc_get_force_tiling", "chmc_time_force", "chmc_set_skin", "chmc_get_skin", "chmc_set_chain_skins",
    "chmc_get_chain_skin", "chmc_list_stats",
    "chmc_time_list", "chmc_time_surrogate", "chmc_set_list_lanes", "chmc_get_list_lanes", "chmc_set_list_entry_bits", "chmc_get_list_entry_bits", "chmc_set_list_check", "chmc_get_list_check", "chmc_set_speculation", "chmc_get_speculation", "chmc_set_list_start_rebuild", "chmc_get_list_start_rebuild", "chmc_set_trajectory", "chmc_get_trajectory",
    "chmc_surrogate_set_params", "chmc_surrogate_get_params",
    "chmc_surrogate_eval", "chmc_surrogate_features", "chmc_surrogate_reserve", "chmc_surrogate_collect",
    "chmc_surrogate_training_data", "chmc_surrogate_loss_grad", "chmc_surrogate_train",
    "chmc_surrogate_get_adam", "chmc_surrogate_set_adam",
    "chmc_surrogate_fm_loss_grad", "chmc_surrogate_reserve_configs", "chmc_surrogate_collect_config",
    "chmc_surrogate_train_fm",
)

SURR_NPARAMS, SURR_NGRAD, SURR_NFEAT, SURR_NHID = 6401, 6338, 32, 64
OPEN_NOALLOC_MAX = 65536             # CHMC_OPEN_NOALLOC_MAX: atoms the static scratch of the no-alloc sum holds
# flat parameter layout: name, shape (W1[64][32], b1, W2[64][64], b2, w3, b3, shift, scale)
SURR_LAYOUT = (("W1", (64, 32)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("w3", (64,)), ("b3", (1,)),
               ("shift", (32,)), ("scale", (32,)))


def pack_surrogate(p):
    """Flat float32 vector of a parameter dict in the library's layout."""
    return np.concatenate([np.asarray(p[k], np.float64).reshape(-1) for k, _ in SURR_LAYOUT]).astype(np.float32)


def unpack_surrogate(v, trained_only=False):
    """Parameter dict of a flat vector (trained_only: a gradient vector of SURR_NGRAD - 1 entries)."""
    out, o = {}, 0
    for k, shp in SURR_LAYOUT:
        if trained_only and k in ("shift", "scale"):
            break
        n = int(np.prod(shp))
        out[k] = np.asarray(v[o:o + n], np.float64).reshape(shp)
        o += n
    return out


@dataclass
class Chain:
    """One LJ system (particles.{N,mass,beta,cutoff,eps,sigma}, geometry.{Lx,Ly,Lz})."""
    N: int
    L: float | Sequence[float]
    cutoff: float = 2.5
    sigma: float = 1.0
    eps: float = 1.0
    mass: float = 1.0
    beta: float = 1.0

    def box(self):
        return tuple(float(v) for v in np.broadcast_to(np.asarray(self.L, dtype=np.float64), (3,)))

    def to_c(self):
        return ChainDesc(int(self.N), (_D * 3)(*self.box()), self.cutoff, self.sigma, self.eps, self.mass, self.beta)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class Engine:
    """A context of one or more chains on one GPU (all calls act on every chain at once).

    speculate=K (K > 1, one chain given): the context holds K copies of the chain and hmc() runs them as
    speculative trajectories of that one chain (chmc_set_speculation): the records and counts are the
    single chain's, identical to a run without speculation."""

    def __init__(self, chains: Sequence[Chain] | Chain, device: int = -1, fp64_forces: bool = False,
                 atom_id_order: bool = False, cell_md: bool = False, speculate: int = 0):
        if isinstance(chains, Chain):
            chains = [chains]
        if speculate > 1:
            if len(chains) != 1:
                raise ValueError("speculate: give the one chain to copy")
            chains = list(chains) * int(speculate)
        self._spec = 0
        self.chains = list(chains)
        arr = (ChainDesc * len(self.chains))(*[c.to_c() for c in self.chains])
        flags = ((FORCE_FP64 if fp64_forces else 0) | (ORDER_ATOM_ID if atom_id_order else 0)
                 | (CELL_MD if cell_md else 0))
        self._h = C.c_void_p()
        rc = lib().chmc_create(C.byref(self._h), len(self.chains), arr, int(device), flags)
        if rc != 0:
            raise ChmcError(f"chmc_create failed ({rc}): {lib().chmc_last_error(None).decode()}")
        if speculate > 1:
            self.set_speculation(int(speculate))

    # -- plumbing ----------------------------------------------------------------------------
    def _ck(self, rc, what):
        if rc != 0:
            raise ChmcError(f"{what} failed ({rc}): {lib().chmc_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().chmc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def nchains(self):
        return len(self.chains)

    def geometry(self, chain=0):
        nb, sb, bt = (_I * 3)(), (_D * 3)(), _I()
        self._ck(lib().chmc_get_geometry(self._h, chain, nb, sb, C.byref(bt)), "get_geometry")
        return [nb[i] for i in range(3)], [sb[i] for i in range(3)], bt.value

    # -- state ---------------------------------------------------------------------------------
    def set_positions(self, x, chain=0):
        self._ck(lib().chmc_set_positions(self._h, chain, _c(x)), "set_positions")

    def positions(self, chain=0):
        out = np.zeros((self.chains[chain].N, 3))
        self._ck(lib().chmc_get_positions(self._h, chain, out), "get_positions")
        return out

    def set_momenta(self, p, chain=0):
        self._ck(lib().chmc_set_momenta(self._h, chain, _c(p)), "set_momenta")

    def momenta(self, chain=0):
        out = np.zeros((self.chains[chain].N, 3))
        self._ck(lib().chmc_get_momenta(self._h, chain, out), "get_momenta")
        return out

    def init_lattice(self, seed, jitter=0.05):
        self._ck(lib().chmc_init_lattice(self._h, int(seed), float(jitter)), "init_lattice")

    # -- hot path ------------------------------------------------------------------------------
    def bin(self):
        self._ck(lib().chmc_bin(self._h), "bin")

    def bins(self, chain=0):
        _, _, bt = self.geometry(chain)
        cnt = np.zeros(bt, np.int32)
        off = np.zeros(bt, np.int32)
        perm = np.zeros(self.chains[chain].N, np.int32)
        self._ck(lib().chmc_get_bins(self._h, chain, cnt, off, perm), "get_bins")
        return perm, cnt, off

    def compute_force(self, fp64=False):
        self._ck(lib().chmc_compute_force(self._h, int(bool(fp64))), "compute_force")

    def force(self, chain=0):
        out = np.zeros((self.chains[chain].N, 3))
        self._ck(lib().chmc_get_force(self._h, chain, out), "get_force")
        return out

    def compute_potential(self):
        out = np.zeros(self.nchains)
        self._ck(lib().chmc_compute_potential(self._h, out), "compute_potential")
        return out

    def evaluate_potential(self):
        out = np.zeros(self.nchains)
        self._ck(lib().chmc_evaluate_potential(self._h, out), "evaluate_potential")
        return out

    def virial(self):
        out = np.zeros(self.nchains)
        self._ck(lib().chmc_virial(self._h, out), "virial")
        return out

    def pair_counts(self, chain=0):
        nc, ni = C.c_longlong(), C.c_longlong()
        self._ck(lib().chmc_pair_counts(self._h, chain, C.byref(nc), C.byref(ni)), "pair_counts")
        return nc.value, ni.value

    def kinetic(self):
        out = np.zeros(self.nchains)
        self._ck(lib().chmc_kinetic(self._h, out), "kinetic")
        return out

    def update_momenta(self, dt):
        self._ck(lib().chmc_update_momenta(self._h, float(dt)), "update_momenta")

    def update_positions(self, dt):
        self._ck(lib().chmc_update_positions(self._h, float(dt)), "update_positions")

    def kick(self, dt):
        self._ck(lib().chmc_kick(self._h, float(dt)), "kick")

    def integrate(self, scheme, dt, steps, lam=OMF2_LAMBDA_REFERENCE):
        self._ck(lib().chmc_integrate(self._h, int(scheme), float(dt), int(steps), float(lam)), "integrate")

    def set_lp_exponent(self, n_prec):
        """Rounding grid 10^-n_prec of the LP_LEAPFROG / LP_OMF4 schemes (set_integrator's exponent)."""
        self._ck(lib().chmc_set_lp_exponent(self._h, int(n_prec)), "set_lp_exponent")

    def lp_exponent(self):
        n = _I()
        self._ck(lib().chmc_get_lp_exponent(self._h, C.byref(n)), "get_lp_exponent")
        return n.value

    def heatbath(self, seed, traj):
        self._ck(lib().chmc_heatbath(self._h, int(seed), int(traj)), "heatbath")

    def save_positions(self):
        self._ck(lib().chmc_save_positions(self._h), "save_positions")

    def restore_positions(self):
        self._ck(lib().chmc_restore_positions(self._h), "restore_positions")

    def set_trajectory(self, istart):
        self._ck(lib().chmc_set_trajectory(self._h, int(istart)), "set_trajectory")

    def trajectory(self):
        v = _I()
        self._ck(lib().chmc_get_trajectory(self._h, C.byref(v)), "get_trajectory")
        return v.value

    def hmc(self, ntraj, therm, scheme, dt, steps, lam=OMF2_LAMBDA_REFERENCE, seed=1, use_graph=True,
            randomize=False, surrogate=False, fallback_window=0, fallback_acc=0.0):
        hp = HmcParams(int(ntraj), int(therm), int(scheme), int(steps), float(dt), float(lam), int(seed),
                       int(bool(use_graph)), int(bool(randomize)), int(bool(surrogate)), int(fallback_window),
                       float(fallback_acc))
        rec = np.zeros(ntraj * self.nchains, HMC_RECORD_DTYPE)
        nacc = np.zeros(self.nchains, np.int32)
        self._ck(lib().chmc_hmc_run(self._h, C.byref(hp), rec.ctypes.data, nacc.ctypes.data), "hmc_run")
        if self._spec > 1:                     # speculative copies of one chain: its records only
            return rec.reshape(ntraj, self.nchains)[:, :1].copy(), nacc[:1].copy()
        return rec.reshape(ntraj, self.nchains), nacc

    def rdf(self, nbins, size_bin, chain=0):
        counts = np.zeros(nbins, np.uint64)
        g = np.zeros(nbins)
        self._ck(lib().chmc_rdf(self._h, chain, int(nbins), float(size_bin), counts.ctypes.data, g.ctypes.data), "rdf")
        return counts, g

    # -- learned surrogate ----------------------------------------------------------------------
    def surrogate_set_params(self, params):
        v = np.ascontiguousarray(params if not isinstance(params, dict) else pack_surrogate(params), np.float32)
        self._ck(lib().chmc_surrogate_set_params(self._h, v.ctypes.data, v.size), "surrogate_set_params")

    def surrogate_params(self):
        v = np.zeros(SURR_NPARAMS, np.float32)
        self._ck(lib().chmc_surrogate_get_params(self._h, v.ctypes.data, v.size), "surrogate_get_params")
        return v

    def surrogate_eval(self, store_force=False):
        e = np.zeros(self.nchains)
        self._ck(lib().chmc_surrogate_eval(self._h, e.ctypes.data, int(bool(store_force))), "surrogate_eval")
        return e

    def surrogate_features(self, chain=0):
        n = self.chains[chain].N
        G = np.zeros((n, SURR_NFEAT), np.float32)
        g = np.zeros((n, SURR_NFEAT), np.float32)
        self._ck(lib().chmc_surrogate_features(self._h, int(chain), G.ctypes.data, g.ctypes.data), "surrogate_features")
        return G, g

    def surrogate_reserve(self, nconfigs):
        self._ck(lib().chmc_surrogate_reserve(self._h, int(nconfigs)), "surrogate_reserve")

    def surrogate_collect(self):
        r = C.c_longlong()
        self._ck(lib().chmc_surrogate_collect(self._h, C.byref(r)), "surrogate_collect")
        return r.value

    def surrogate_training_data(self, row0, nrows):
        G = np.zeros((nrows, SURR_NFEAT), np.float32)
        u = np.zeros(nrows, np.float32)
        self._ck(lib().chmc_surrogate_training_data(self._h, int(row0), int(nrows), G.ctypes.data, u.ctypes.data),
                 "surrogate_training_data")
        return G, u

    def surrogate_loss_grad(self, row0, nrows):
        g = np.zeros(SURR_NGRAD, np.float32)
        self._ck(lib().chmc_surrogate_loss_grad(self._h, int(row0), int(nrows), g.ctypes.data), "surrogate_loss_grad")
        return g

    def surrogate_train(self, nsteps, batch_rows, lr=1e-3, seed=1):
        h = np.zeros(nsteps, np.float32)
        self._ck(lib().chmc_surrogate_train(self._h, int(nsteps), int(batch_rows), float(lr), int(seed), h.ctypes.data),
                 "surrogate_train")
        return h

    def surrogate_adam_state(self):
        """(m, v, step) of the Adam optimiser: with surrogate_params() a training checkpoint."""
        m = np.zeros(SURR_NGRAD - 1, np.float32)
        v = np.zeros(SURR_NGRAD - 1, np.float32)
        st = _I()
        self._ck(lib().chmc_surrogate_get_adam(self._h, m.ctypes.data, v.ctypes.data, C.byref(st)), "surrogate_get_adam")
        return m, v, st.value

    def surrogate_set_adam_state(self, m, v, step):
        m = np.ascontiguousarray(m, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        self._ck(lib().chmc_surrogate_set_adam(self._h, m.ctypes.data, v.ctypes.data, int(step)), "surrogate_set_adam")

    def surrogate_fm_loss_grad(self, w_force):
        """Energy + force-matching loss gradient at the current positions (no update): (gradient of the
        trained parameters, mean energy loss, mean force loss)."""
        g = np.zeros(SURR_NGRAD - 1, np.float32)
        le, lf = _D(), _D()
        self._ck(lib().chmc_surrogate_fm_loss_grad(self._h, float(w_force), g.ctypes.data, C.byref(le), C.byref(lf)),
                 "surrogate_fm_loss_grad")
        return g, le.value, lf.value

    def surrogate_reserve_configs(self, nconfigs):
        self._ck(lib().chmc_surrogate_reserve_configs(self._h, int(nconfigs)), "surrogate_reserve_configs")

    def surrogate_collect_config(self):
        n = _I()
        self._ck(lib().chmc_surrogate_collect_config(self._h, C.byref(n)), "surrogate_collect_config")
        return n.value

    def surrogate_train_fm(self, nsteps, w_force, lr=1e-3, seed=1):
        """nsteps Adam steps of energy + force matching, one stored configuration each; returns the mean
        energy and force losses before every step."""
        le = np.zeros(nsteps, np.float32)
        lf = np.zeros(nsteps, np.float32)
        self._ck(lib().chmc_surrogate_train_fm(self._h, int(nsteps), float(w_force), float(lr), int(seed), le.ctypes.data,
                                               lf.ctypes.data), "surrogate_train_fm")
        return le, lf

    # -- tuning --------------------------------------------------------------------------------
    def set_force_tiling(self, islots):
        self._ck(lib().chmc_set_force_tiling(self._h, int(islots)), "set_force_tiling")

    def force_tiling(self):
        v = _I()
        self._ck(lib().chmc_get_force_tiling(self._h, C.byref(v)), "get_force_tiling")
        return v.value

    def set_skin(self, skin):
        self._ck(lib().chmc_set_skin(self._h, float(skin)), "set_skin")

    def skin(self):
        v = _D()
        self._ck(lib().chmc_get_skin(self._h, C.byref(v)), "get_skin")
        return v.value

    def set_chain_skins(self, skins):
        s = np.ascontiguousarray(skins, np.float64)
        if s.shape != (self.nchains,):
            raise ValueError(f"set_chain_skins: {self.nchains} skins expected, got shape {s.shape}")
        self._ck(lib().chmc_set_chain_skins(self._h, s), "set_chain_skins")

    def chain_skin(self, chain=0):
        v = _D()
        self._ck(lib().chmc_get_chain_skin(self._h, int(chain), C.byref(v)), "get_chain_skin")
        return v.value

    def list_stats(self, chain=0):
        a, b, r = C.c_longlong(), C.c_longlong(), C.c_longlong()
        self._ck(lib().chmc_list_stats(self._h, int(chain), C.byref(a), C.byref(b), C.byref(r)), "list_stats")
        return a.value, b.value, r.value

    def set_list_lanes(self, lanes=0):
        self._ck(lib().chmc_set_list_lanes(self._h, int(lanes)), "set_list_lanes")

    def list_lanes(self):
        v = _I()
        self._ck(lib().chmc_get_list_lanes(self._h, C.byref(v)), "get_list_lanes")
        return v.value

    def set_list_entry_bits(self, bits=0):
        """List entry format: 16 (segment-relative), 32 (slot), or 0 (16 where the rows exceed 256 MiB)."""
        self._ck(lib().chmc_set_list_entry_bits(self._h, int(bits)), "set_list_entry_bits")

    def list_entry_bits(self):
        v = _I()
        self._ck(lib().chmc_get_list_entry_bits(self._h, C.byref(v)), "get_list_entry_bits")
        return v.value

    def set_list_check(self, every):
        """Rebuild checks of list MD before every `every`-th force evaluation (1 to 4)."""
        self._ck(lib().chmc_set_list_check(self._h, int(every)), "set_list_check")

    def list_check(self):
        v = _I()
        self._ck(lib().chmc_get_list_check(self._h, C.byref(v)), "get_list_check")
        return v.value

    def set_list_start_rebuild(self, on):
        """Rebuild a chain's lists at the start of the trajectory after an accepted one (list MD)."""
        self._ck(lib().chmc_set_list_start_rebuild(self._h, int(bool(on))), "set_list_start_rebuild")

    def list_start_rebuild(self):
        v = _I()
        self._ck(lib().chmc_get_list_start_rebuild(self._h, C.byref(v)), "get_list_start_rebuild")
        return bool(v.value)

    def set_speculation(self, copies):
        """Run hmc() as speculative trajectories of chain 0 over the context's `copies` identical chains
        (copies <= 1: off)."""
        self._ck(lib().chmc_set_speculation(self._h, int(copies)), "set_speculation")
        self._spec = self.speculation()

    def speculation(self):
        v = _I()
        self._ck(lib().chmc_get_speculation(self._h, C.byref(v)), "get_speculation")
        return v.value

    def time_list(self, reps=20):
        a, b = _D(), _D()
        self._ck(lib().chmc_time_list(self._h, int(reps), C.byref(a), C.byref(b)), "time_list")
        return a.value, b.value

    def time_surrogate(self, reps=20):
        """(ms per descriptor + MLP launch, ms per surrogate force sweep) at the current positions."""
        a, b = _D(), _D()
        self._ck(lib().chmc_time_surrogate(self._h, int(reps), C.byref(a), C.byref(b)), "time_surrogate")
        return a.value, b.value

    def time_force(self, reps=20):
        b, f = _D(), _D()
        self._ck(lib().chmc_time_force(self._h, int(reps), C.byref(b), C.byref(f)), "time_force")
        return b.value, f.value


def potential_energy_open(x, n=12, a=1e-12, b=1e-6):
    """code-carsten potential_energy() on the GPU (open boundaries, all pairs)."""
    out = _D()
    rc = lib().chmc_potential_energy_open(_c(x), int(np.asarray(x).shape[0]), int(n), float(a), float(b), C.byref(out))
    if rc != 0:
        raise ChmcError(f"potential_energy_open failed ({rc}): {lib().chmc_last_error(None).decode()}")
    return out.value
