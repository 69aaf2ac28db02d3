"""chmc_potential_energy_open_noalloc (the entry point that allocates nothing: block partials in a
static device buffer) against chmc_potential_energy_open on the same sets, through the C ABI.  Device
buffers are made with the HIP runtime libchmc.so itself links, so the library under test makes no
allocation of its own.  Two calls in a row, the second on a smaller set (the shared buffer keeps the
first call's partials), then the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]

    def ck(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.ck(self.h.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p

    def h2d(self, dst, arr):
        self.ck(self.h.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")

    def d2h(self, arr, src):
        self.ck(self.h.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2), "hipMemcpy D2H")


def _cluster(n, seed, spacing=0.35):
    rng = np.random.default_rng(seed)
    side = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    return np.ascontiguousarray(g * spacing + rng.uniform(-0.05, 0.05, (n, 3)), dtype=np.float64)


def test_noalloc_matches_allocating_entry_point(chmc):
    lib = chmc.lib()
    hip = _Hip()
    stream = C.c_void_p()
    hip.ck(hip.h.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    nmax = 3000
    dx = hip.malloc(nmax * 3 * 8)
    dres = hip.malloc(8)
    try:
        cases = ((3000, 1, 0.35, (12, 1e-12, 1e-6)), (700, 2, 1.1, (12, 4.0, 4.0)), (1, 3, 1.1, (12, 4.0, 4.0)))
        for n, seed, spacing, (ne, a, b) in cases:
            x = _cluster(n, seed, spacing)
            print(f"noalloc case n={n} seed={seed} n_exp={ne}", flush=True)
            ref = C.c_double()
            assert lib.chmc_potential_energy_open(x, n, ne, a, b, C.byref(ref)) == 0
            hip.h2d(dx, x)
            rc = lib.chmc_potential_energy_open_noalloc(dx, n, ne, a, b, dres, stream)
            assert rc == 0, lib.chmc_last_error(None).decode()
            hip.ck(hip.h.hipStreamSynchronize(stream), "sync")
            got = np.zeros(1)
            hip.d2h(got, dres)
            assert got[0] == ref.value, (n, got[0], ref.value)        # same partials, same summation order
            if n == 1:
                assert got[0] == 0.0
        # one past the static buffer's capacity: a clear error, nothing launched
        rc = lib.chmc_potential_energy_open_noalloc(dx, chmc.OPEN_NOALLOC_MAX + 1, 12, 4.0, 4.0, dres, stream)
        assert rc == -1
        assert "exceeds" in lib.chmc_last_error(None).decode()
    finally:
        hip.h.hipFree(dx)
        hip.h.hipFree(dres)
        hip.h.hipStreamDestroy(stream)
