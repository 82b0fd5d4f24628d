"""ghx_spectrumc_noalloc, the energy spectrum for callers that cannot allocate
device memory, against ghx_spectrumc (kin = 1, hel = 0) through the C ABI:
two calls in a row on one GPU, the second on a smaller grid (the shared scratch
array keeps the first call's partials), then a grid whose shell partials do not
fit the scratch, which is refused.  Shell sums: relative L2 error 1e-12 and
1e-12 relative per non-empty shell (the per-block bins are summed in LDS by
atomics, so the last bits may differ between any two calls)."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_parity import grids, rand_spec, rel, to_gpu

pytestmark = pytest.mark.gpu


def _call(lib, fn, g, fields, *extra):
    nb = g.nbins
    ek = (C.c_double * nb)()
    ptrs = [C.c_void_p(t.data_ptr()) for t in fields]
    args = list(extra)
    if fn == "ghx_spectrumc":
        hk = (C.c_double * nb)()
        rc = lib.ghx_spectrumc(g._h, *ptrs, 1, 0, C.cast(ek, C.POINTER(C.c_double)),
                               C.cast(hk, C.POINTER(C.c_double)))
    else:
        rc = lib.ghx_spectrumc_noalloc(g._h, *ptrs, C.cast(ek, C.POINTER(C.c_double)), *args)
    return rc, np.array(ek)


@pytest.mark.parametrize("first,second", [((64, 64, 64), (32, 32, 32)), ((48, 32, 24), (16, 16, 16))])
def test_noalloc_spectrum_matches_spectrumc_on_two_calls(gpu, first, second):
    import torch
    lib = gpu
    for shape, seed in ((first, 11), (second, 21)):
        g, og = grids(shape)
        v = [to_gpu(rand_spec(og, seed + s)) for s in range(3)]
        torch.cuda.synchronize()
        rc, got = _call(lib, "ghx_spectrumc_noalloc", g, v)
        assert rc == 0, lib.ghx_last_error()
        rc, want = _call(lib, "ghx_spectrumc", g, v)
        assert rc == 0, lib.ghx_last_error()
        assert got.shape == want.shape == (g.nbins,)
        assert rel(got, want) < 1e-12
        nz = want != 0
        assert np.array_equal(got != 0, nz)
        assert np.all(np.abs(got[nz] - want[nz]) <= 1e-12 * np.abs(want[nz]))
        assert want.sum() > 0


def test_noalloc_spectrum_refuses_more_shells_than_the_scratch_holds(gpu):
    from ghost_amd.pseudospec import Grid
    lib = gpu
    # 128^3 launches 1024 shell blocks; Dkk = 0.1 gives nmax = 1280 and 641
    # shells, 657,025 doubles of partials against the 524,288 of the scratch
    g = Grid(128, 128, 128, L=(1.0, 1.0, 1.0), arbsize=True, Dkk=0.1)
    assert g.nbins == 641
    a = g.czeros()
    rc, _ = _call(lib, "ghx_spectrumc_noalloc", g, [a, a, a])
    assert rc == -2
    assert b"exceed" in lib.ghx_last_error()
