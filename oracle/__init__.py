"""ctypes front-end of the FP64 CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``, as the checker or the timed CPU baseline -- never by the
product library.  Arrays are numpy float64 (N, 3) C-contiguous, the AoS layout of the
reference's ``config<coordinate>`` (code-carsten/config.hh:15-94).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_carsten.so")

_D = C.c_double
_I = C.c_int
_U64 = C.c_uint64
_PD = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_PI = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_PU64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_L3 = _D * 3
_I3 = _I * 3

LEAPFROG, LP_LEAPFROG, OMF4, LP_OMF4, EULER, RUTH, OMF2 = 0, 1, 2, 3, 4, 5, 6
BIN_COUNTING, BIN_REFALGO = 0, 1


def build() -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is mounted)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference/code-carsten"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class MT64(C.Structure):
    _fields_ = [("mt", _U64 * 312), ("mti", _I)]


class System(C.Structure):
    _fields_ = [
        ("N", _I), ("L", _L3), ("cutoff", _D), ("sigma", _D), ("eps", _D), ("mass", _D), ("beta", _D),
        ("nbin", _I3), ("bintot", _I), ("sizebin", _L3), ("bin_mode", _I),
        ("x", C.POINTER(_D)), ("p", C.POINTER(_D)), ("f", C.POINTER(_D)), ("x_saved", C.POINTER(_D)),
        ("perm", C.POINTER(_I)), ("perm_tmp", C.POINTER(_I)), ("bincount", C.POINTER(_I)),
        ("binoffsets", C.POINTER(_I)), ("force_evals", C.c_longlong), ("lp_scale", _D), ("lp_active", _I),
    ]


class HmcParams(C.Structure):
    _fields_ = [("ntraj", _I), ("therm", _I), ("scheme", _I), ("steps", _I), ("randomize", _I),
                ("dt", _D), ("lam", _D), ("seed", _U64), ("chain", _I)]


def _declare(L):
    L.orc_philox4x64.argtypes = [_PU64, _PU64, _PU64]
    L.orc_mt64_seed.argtypes = [C.POINTER(MT64), _U64]
    L.orc_mt64_next.argtypes = [C.POINTER(MT64)]
    L.orc_mt64_next.restype = _U64
    L.orc_mt64_uniform.argtypes = [C.POINTER(MT64)]
    L.orc_mt64_uniform.restype = _D
    L.orc_binomial_steps.argtypes = [C.POINTER(MT64), _I]
    L.orc_binomial_steps.restype = _I
    L.orc_lattice_init.argtypes = [_I, _L3, _U64, _I, _D, _PD]
    L.orc_heatbath.argtypes = [_I, _D, _D, _U64, _I, _U64, _PD]
    L.orc_cutoff_binning.argtypes = [_L3, _D, _I3, _L3]
    L.orc_cutoff_binning.restype = _I
    L.orc_binning.argtypes = [_I, _PD, _I3, _L3, _PI, _PI, _PI]
    L.orc_binning_refalgo.argtypes = [_I, _PD, _I3, _L3, _PI, _PI, _PI, _PI]
    bins = [_PI, _PI, _PI]
    L.orc_force_binning.argtypes = [_I, _PD, _L3, _I3] + bins + [_D, _D, _D, _PD]
    L.orc_potential_binning.argtypes = [_I, _PD, _L3, _I3] + bins + [_D, _D, _D]
    L.orc_potential_binning.restype = _D
    L.orc_virial_binning.argtypes = [_I, _PD, _L3, _I3] + bins + [_D, _D, _D]
    L.orc_virial_binning.restype = _D
    L.orc_pair_counts.argtypes = [_I, _PD, _L3, _I3] + bins + [_D, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.orc_potential_all.argtypes = [_I, _PD, _L3, _D, _D, _D]
    L.orc_potential_all.restype = _D
    L.orc_force_all.argtypes = [_I, _PD, _L3, _D, _D, _D, _PD]
    L.orc_lennard_jones.argtypes = [_D, _I, _D, _D]
    L.orc_lennard_jones.restype = _D
    L.orc_potential_energy_open.argtypes = [_I, _PD, _I, _D, _D]
    L.orc_potential_energy_open.restype = _D
    L.orc_kinetic.argtypes = [_I, _PD, _D]
    L.orc_kinetic.restype = _D
    L.orc_kick.argtypes = [_I, _PD, _PD, _D, _D]
    L.orc_drift.argtypes = [_I, _PD, _PD, _D, _D, _L3]
    L.orc_lp_scale.argtypes = [_I]
    L.orc_lp_scale.restype = _D
    L.orc_kick_lp.argtypes = [_I, _PD, _PD, _D, _D, _D]
    L.orc_drift_lp.argtypes = [_I, _PD, _PD, _D, _D, _L3, _D]
    L.orc_rdf.argtypes = [_I, _PD, _L3, _I, _D, _PU64, _PD]
    L.orc_system_init.argtypes = [C.POINTER(System), _I, _L3, _D, _D, _D, _D, _D, _I]
    L.orc_system_init.restype = _I
    L.orc_system_free.argtypes = [C.POINTER(System)]
    L.orc_sys_create_binning.argtypes = [C.POINTER(System)]
    L.orc_sys_compute_force.argtypes = [C.POINTER(System)]
    L.orc_sys_compute_potential.argtypes = [C.POINTER(System)]
    L.orc_sys_compute_potential.restype = _D
    L.orc_sys_update_momenta.argtypes = [C.POINTER(System), _D]
    L.orc_sys_update_positions.argtypes = [C.POINTER(System), _D]
    L.orc_sys_integrate.argtypes = [C.POINTER(System), _I, _D, _I, _D]
    L.orc_hmc_run.argtypes = [C.POINTER(System), C.POINTER(HmcParams), _PD, _PI, _PD, _PD, _PD]
    L.orc_hmc_run.restype = _I


def _l3(L):
    return _L3(*[float(v) for v in np.broadcast_to(np.asarray(L, dtype=np.float64), (3,))])


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------------------------------------
def philox4x64(ctr, key):
    out = np.zeros(4, np.uint64)
    lib().orc_philox4x64(_c(ctr, np.uint64), _c(key, np.uint64), out)
    return out


class MT19937_64:
    """mt19937_64 as HMC.cpp:38 uses it (seeded by ``seed``)."""

    def __init__(self, seed):
        self._g = MT64()
        lib().orc_mt64_seed(C.byref(self._g), int(seed) & (2**64 - 1))

    def next(self):
        return lib().orc_mt64_next(C.byref(self._g))

    def uniform(self):
        return lib().orc_mt64_uniform(C.byref(self._g))

    def binomial_steps(self, average_steps):
        return lib().orc_binomial_steps(C.byref(self._g), int(average_steps))


def lattice_init(N, L, seed, chain=0, jitter=0.05):
    x = np.zeros((N, 3))
    lib().orc_lattice_init(int(N), _l3(L), int(seed), int(chain), float(jitter), x)
    return x


def heatbath(N, mass, beta, seed, chain, traj):
    p = np.zeros((N, 3))
    lib().orc_heatbath(int(N), float(mass), float(beta), int(seed), int(chain), int(traj), p)
    return p


def cutoff_binning(L, cutoff):
    nb, sb = _I3(), _L3()
    rc = lib().orc_cutoff_binning(_l3(L), float(cutoff), nb, sb)
    if rc != 0:
        raise ValueError(f"cutoff_binning: box {list(np.broadcast_to(L, 3))} incompatible with cutoff {cutoff}")
    return [nb[i] for i in range(3)], [sb[i] for i in range(3)]


def binning(x, L, cutoff, perm=None, refalgo=False):
    """Stable counting sort of ``perm`` (identity by default) by cell; returns (perm, count, offsets)."""
    x = _c(x)
    N = x.shape[0]
    nbin, sizebin = cutoff_binning(L, cutoff)
    bintot = nbin[0] * nbin[1] * nbin[2]
    perm = np.arange(N, dtype=np.int32) if perm is None else _c(perm, np.int32).copy()
    cnt = np.zeros(bintot, np.int32)
    off = np.zeros(bintot, np.int32)
    if refalgo:
        tmp = np.zeros(N, np.int32)
        lib().orc_binning_refalgo(N, x, _I3(*nbin), _L3(*sizebin), perm, tmp, cnt, off)
    else:
        lib().orc_binning(N, x, _I3(*nbin), _L3(*sizebin), perm, cnt, off)
    return perm, cnt, off


def _bins(x, L, cutoff, bins):
    if bins is None:
        bins = binning(x, L, cutoff)
    nbin, _ = cutoff_binning(L, cutoff)
    return _I3(*nbin), bins


def force_binning(x, L, cutoff, sigma=1.0, eps=1.0, bins=None):
    x = _c(x)
    nb, (perm, cnt, off) = _bins(x, L, cutoff, bins)
    f = np.zeros_like(x)
    lib().orc_force_binning(x.shape[0], x, _l3(L), nb, perm, cnt, off, cutoff, sigma, eps, f)
    return f


def potential_binning(x, L, cutoff, sigma=1.0, eps=1.0, bins=None):
    x = _c(x)
    nb, (perm, cnt, off) = _bins(x, L, cutoff, bins)
    return lib().orc_potential_binning(x.shape[0], x, _l3(L), nb, perm, cnt, off, cutoff, sigma, eps)


def virial_binning(x, L, cutoff, sigma=1.0, eps=1.0, bins=None):
    x = _c(x)
    nb, (perm, cnt, off) = _bins(x, L, cutoff, bins)
    return lib().orc_virial_binning(x.shape[0], x, _l3(L), nb, perm, cnt, off, cutoff, sigma, eps)


def pair_counts(x, L, cutoff, bins=None):
    x = _c(x)
    nb, (perm, cnt, off) = _bins(x, L, cutoff, bins)
    nc, ni = C.c_longlong(), C.c_longlong()
    lib().orc_pair_counts(x.shape[0], x, _l3(L), nb, perm, cnt, off, cutoff, C.byref(nc), C.byref(ni))
    return nc.value, ni.value


def potential_all(x, L, cutoff, sigma=1.0, eps=1.0):
    x = _c(x)
    return lib().orc_potential_all(x.shape[0], x, _l3(L), cutoff, sigma, eps)


def force_all(x, L, cutoff, sigma=1.0, eps=1.0):
    x = _c(x)
    f = np.zeros_like(x)
    lib().orc_force_all(x.shape[0], x, _l3(L), cutoff, sigma, eps, f)
    return f


def lennard_jones(r, n=12, a=1e-12, b=1e-6):
    return lib().orc_lennard_jones(float(r), int(n), float(a), float(b))


def potential_energy_open(x, n=12, a=1e-12, b=1e-6):
    x = _c(x)
    return lib().orc_potential_energy_open(x.shape[0], x, int(n), float(a), float(b))


def kinetic(p, mass):
    p = _c(p)
    return lib().orc_kinetic(p.shape[0], p, float(mass))


def kick(p, f, dt, coeff_p):
    p = _c(p).copy()
    lib().orc_kick(p.shape[0], p, _c(f), float(dt), float(coeff_p))
    return p


def drift(x, p, dt, coeff_x, L):
    x = _c(x).copy()
    lib().orc_drift(x.shape[0], x, _c(p), float(dt), float(coeff_x), _l3(L))
    return x


def rdf(x, L, nbins, size_bin):
    x = _c(x)
    counts = np.zeros(nbins, np.uint64)
    g = np.zeros(nbins)
    lib().orc_rdf(x.shape[0], x, _l3(L), int(nbins), float(size_bin), counts, g)
    return counts, g


class Sim:
    """One periodic LJ system, the oracle counterpart of particles_type + integrator_type."""

    def __init__(self, N, L, cutoff=2.5, sigma=1.0, eps=1.0, mass=1.0, beta=1.0, refalgo=False):
        self._s = System()
        rc = lib().orc_system_init(C.byref(self._s), int(N), _l3(L), cutoff, sigma, eps, mass, beta,
                                   BIN_REFALGO if refalgo else BIN_COUNTING)
        if rc != 0:
            raise ValueError("orc_system_init failed")
        self.N = int(N)

    def __del__(self):
        try:
            lib().orc_system_free(C.byref(self._s))
        except Exception:
            pass

    def _view(self, ptr, shape, dtype=np.float64):
        return np.ctypeslib.as_array(ptr, shape=shape).view(dtype)

    @property
    def x(self):
        return np.ctypeslib.as_array(self._s.x, shape=(self.N, 3))

    @property
    def p(self):
        return np.ctypeslib.as_array(self._s.p, shape=(self.N, 3))

    @property
    def f(self):
        return np.ctypeslib.as_array(self._s.f, shape=(self.N, 3))

    @property
    def perm(self):
        return np.ctypeslib.as_array(self._s.perm, shape=(self.N,))

    @property
    def bincount(self):
        return np.ctypeslib.as_array(self._s.bincount, shape=(self._s.bintot,))

    @property
    def binoffsets(self):
        return np.ctypeslib.as_array(self._s.binoffsets, shape=(self._s.bintot,))

    @property
    def force_evals(self):
        return self._s.force_evals

    def create_binning(self):
        lib().orc_sys_create_binning(C.byref(self._s))

    def compute_force(self):
        lib().orc_sys_compute_force(C.byref(self._s))

    def compute_potential(self):
        return lib().orc_sys_compute_potential(C.byref(self._s))

    def update_momenta(self, dt):
        lib().orc_sys_update_momenta(C.byref(self._s), float(dt))

    def update_positions(self, dt):
        lib().orc_sys_update_positions(C.byref(self._s), float(dt))

    def set_lp_exponent(self, n_prec):
        """Rounding grid 10^-n_prec of the lp_* integrators (set_integrator's exponent)."""
        self._s.lp_scale = lib().orc_lp_scale(int(n_prec))

    def integrate(self, scheme, dt, steps, lam=0.1938):
        lib().orc_sys_integrate(C.byref(self._s), int(scheme), float(dt), int(steps), float(lam))

    def hmc(self, ntraj, therm, scheme, dt, steps, lam=0.1938, seed=1, chain=0, randomize=False):
        hp = HmcParams(int(ntraj), int(therm), int(scheme), int(steps), int(bool(randomize)),
                       float(dt), float(lam), int(seed), int(chain))
        dh = np.zeros(ntraj)
        acc = np.zeros(ntraj, np.int32)
        u = np.zeros(ntraj)
        k = np.zeros(ntraj)
        w = np.zeros(ntraj)
        nacc = lib().orc_hmc_run(C.byref(self._s), C.byref(hp), dh, acc, u, k, w)
        return dict(dh=dh, acc=acc, u=u, k=k, w=w, nacc=nacc)


# ---------------------------------------------------------------------------------------------
_ref = None


def ref_lib():
    """The reference's code-carsten LennardJones/potential_energy, compiled from its own sources;
    None where the reference tree was never present to build it."""
    global _ref
    if _ref is None and os.path.exists(_REF_PATH):
        _ref = C.CDLL(_REF_PATH)
        _ref.ref_lennard_jones.argtypes = [_D, _I, _D, _D]
        _ref.ref_lennard_jones.restype = _D
        _ref.ref_potential_energy.argtypes = [_PD, _I]
        _ref.ref_potential_energy.restype = _D
    return _ref
