"""The per-substep entry points behind the Fortran drop-ins of
include/mhd_rkstep2.f90, bouss_rkstep2.f90 and rotbouss_rkstep2.f90: the
caller's RK loop (C = state at the step start, o = ord..1) through the C ABI
against the oracle's whole steps.  Relative L2 1e-12."""
import numpy as np
import pytest

from test_gpu_bouss import bouss_ic, grids, rel, to_gpu, to_np

pytestmark = pytest.mark.gpu


def loop(g, fn, state, nord, *args, forcing=None, **kw):
    """GHOST's loop on device fields: C = v; for o = nord..1: v = rkstep2(o)."""
    import torch
    v = [to_gpu(x) for x in state]
    c = [x.clone() for x in v]
    for o in range(nord, 0, -1):
        out = [torch.empty_like(x) for x in v]
        fn(o, *args, v, c, *(forcing or ()), out, **kw)
        v = out
    return [to_np(x) for x in v]


@pytest.mark.parametrize("ord,forced", [(2, True), (1, False), (3, True)])
def test_mhd_rkstep2_loop(gpu, ord, forced):
    g, og = grids((32, 32, 16), True, (1.0, 1.0, 0.5))
    v, _ = og.initialv_k4exp(3.0, 1.0, 91)
    a, _ = og.initialv_k4exp(2.5, 0.5, 92)
    f = og.initialfv_random(2.0, 3.0, 0.3, 93)[0] if forced else None
    m = og.initialfv_random(1.0, 2.0, 0.2, 94)[0] if forced else None
    dt, nu, mu = 2e-3, 4e-3, 5e-3
    forcing = ([to_gpu(x) for x in f], [to_gpu(x) for x in m]) if forced else (None, None)
    got = loop(g, g.mhd_rkstep2, v + a, ord, dt, nu, mu, forcing=forcing)
    rv, ra = [x.copy() for x in v], [x.copy() for x in a]
    og.mhd_step(rv, ra, dt, nu, mu, f=f, m=m, ord=ord)
    for k, want in enumerate(rv + ra):
        assert rel(got[k], want) < 1e-12, k


@pytest.mark.parametrize("ord,omegaz", [(2, None), (1, None), (3, None), (2, 2.5), (3, 1.5)])
def test_bouss_rkstep2_loop(gpu, ord, omegaz):
    """ord = 3 without rotation includes the correction of
    bouss_rkstep2.f90:73-145 (ghx_bouss_ord3 after o = 1)."""
    g, og = grids((32, 32, 32))
    v0, th0 = bouss_ic(og, 95)
    f, _ = og.initialfv_random(2.0, 3.0, 0.3, 96)
    fs, _ = og.initialfv_random(1.0, 2.0, 0.2, 97)
    dt, nu, kappa, xmom, xtemp = 3e-3, 4e-3, 5e-3, 2.0, 2.0
    forcing = ([to_gpu(x) for x in f] + [to_gpu(fs[0])],)
    got = loop(g, g.bouss_rkstep2, v0 + [th0], ord, dt, nu, kappa, xmom, xtemp, forcing=forcing, omegaz=omegaz,
               ord=ord)
    rv, rth = [x.copy() for x in v0], th0.copy()
    if omegaz is None:
        og.bouss_step(rv, rth, dt, nu, kappa, xmom, xtemp, f=f, fs=fs[0], ord=ord)
    else:
        og.rotbouss_step(rv, rth, dt, nu, kappa, xmom, xtemp, omegaz, f=f, fs=fs[0], ord=ord)
    for k, want in enumerate(rv + [rth]):
        assert rel(got[k], want) < 1e-12, k
