"""CPU-side checks of the C ABI: libchmc.so loads, exports every entry point include/chmc.h
declares, and the Python front-end binds exactly that set.  No compute call is made (no GPU here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "chmc.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(chmc_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol(chmc):
    lib = ctypes.CDLL(chmc.LIB_PATH)
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header(chmc):
    assert sorted(chmc.EXPORTED) == declared()


def test_no_device_means_clean_error(chmc):
    # on a machine without a GPU chmc_create must fail with an error code, not crash
    import pytest
    try:
        e = chmc.Engine(chmc.Chain(N=8, L=6.0))
    except chmc.ChmcError as err:
        assert "device" in str(err).lower() or "hip" in str(err).lower()
    else:
        e.close()
        pytest.skip("a GPU is present")
