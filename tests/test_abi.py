"""The C ABI library loads, exports every function include/ghx.h declares, and
reports argument/size errors without touching a GPU."""
import ctypes as C

import pytest

import ghost_amd
from ghost_amd import _lib


def test_every_declared_symbol_is_exported():
    lib = ghost_amd.load()
    names = ghost_amd.declared_symbols()
    assert len(names) >= 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_size_and_argument_errors():
    lib = ghost_amd.load()
    h = C.c_void_p()
    # 7 has no compiled schedule (x uses nx/2 = 7)
    assert lib.ghx_plan_create(C.byref(h), 14, 16, 16, -1, 64) == -2
    assert b"no compiled FFT schedule" in lib.ghx_last_error()
    assert lib.ghx_plan_create(C.byref(h), 15, 16, 16, -1, 64) == -2  # odd nx
    assert lib.ghx_plan_create(C.byref(h), 16, 16, 16, 0, 64) == -1   # bad fftdir
    assert lib.ghx_derivk3(None, None, None, 1) == -1


def test_no_device_is_an_error_not_a_fallback():
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is visible")
    lib = ghost_amd.load()
    assert lib.ghx_device_ok() == 0
    h = C.c_void_p()
    assert lib.ghx_plan_create(C.byref(h), 16, 16, 16, -1, 64) == -5
    g = C.c_void_p()
    assert lib.ghx_grid_create(C.byref(g), 32, 32, 32, 1.0, 1.0, 1.0, 0, 0.0) == -5
    with pytest.raises(_lib.GhxError):
        _lib.check(-5)


def test_randu_abi_matches_park_miller():
    s = C.c_int32(1 ^ 123459876)
    lib = ghost_amd.load()
    for _ in range(10000):
        lib.ghx_randu(C.byref(s))
    assert s.value ^ 123459876 == 1043618065
