"""Parity of the rotating-frame and passive-scalar solvers (include/
roth_rkstep2.f90, rotbouss_rkstep2.f90, phd_rkstep2.f90) on the fused HIP
path with the oracle, through the C ABI.  Relative L2 1e-12 after several
steps, 1e-13 for one substep."""
import os

import numpy as np
import pytest

from test_gpu_bouss import bouss_ic, grids, rel, to_gpu, to_np

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ord", [2, 3])
@pytest.mark.parametrize("shape,arb,L", [((32, 32, 32), False, (1, 1, 1)), ((32, 32, 16), True, (1.0, 1.0, 0.5))])
def test_roth_forced_steps(gpu, ord, shape, arb, L):
    from ghost_amd.solvers import HDSolver
    g, og = grids(shape, arb, L)
    v0, _ = og.initialv_k4exp(3.0, 1.0, 41)
    f, _ = og.initialfv_random(2.0, 3.0, 0.3, 42)
    dt, nu, omegaz = 3e-3, 4e-3, 4.0
    s = HDSolver(g, dt, nu, ord=ord)
    s.set_rotation(omegaz)
    s.set_velocity([to_gpu(x) for x in v0])
    s.set_forcing([to_gpu(x) for x in f])
    s.step(8)
    gv = s.velocity()
    rv = [x.copy() for x in v0]
    for _ in range(8):
        og.roth_step(rv, dt, nu, omegaz, f=f, ord=ord)
    for k in range(3):
        assert rel(to_np(gv[k]), rv[k]) < 1e-12
    # rotation changed the result
    hv = [x.copy() for x in v0]
    for _ in range(8):
        og.hd_step(hv, dt, nu, f=f, ord=ord)
    assert rel(rv[0], hv[0]) > 1e-4


def test_roth_rkstep2_one_substep(gpu):
    g, og = grids((24, 16, 32), True, (1.0, 1.5, 0.75))
    rng = np.random.default_rng(43)
    v = [og.r2c(rng.standard_normal(og.rshape)) for _ in range(3)]
    c = [og.r2c(rng.standard_normal(og.rshape)) for _ in range(3)]
    f = [og.r2c(rng.standard_normal(og.rshape)) for _ in range(3)]
    for o in (1, 2):
        want = [x.copy() for x in v]
        og.roth_substep(want, c, o, 1e-2, 3e-3, 1.7, f=f)
        out = [g.czeros() for _ in range(3)]
        g.hd_rkstep2(o, 1e-2, 3e-3, [to_gpu(x) for x in v], [to_gpu(x) for x in c], [to_gpu(x) for x in f], out,
                     omegaz=1.7)
        for k in range(3):
            assert rel(to_np(out[k]), want[k]) < 1e-13


@pytest.mark.parametrize("ord", [1, 2, 3])
def test_rotbouss_forced_steps(gpu, ord):
    from ghost_amd.solvers import BoussSolver
    g, og = grids((32, 32, 32))
    v0, th0 = bouss_ic(og, 44)
    f, _ = og.initialfv_random(2.0, 3.0, 0.3, 45)
    fs, _ = og.initialfv_random(1.0, 2.0, 0.2, 46)
    dt, nu, kappa, xmom, xtemp, omegaz = 3e-3, 4e-3, 5e-3, 2.0, 2.0, 3.0
    s = BoussSolver(g, dt, nu, kappa, xmom, xtemp, ord=ord)
    s.set_rotation(omegaz)
    s.set_state([to_gpu(x) for x in v0], to_gpu(th0))
    s.set_forcing([to_gpu(x) for x in f], to_gpu(fs[0]))
    s.step(8)
    gv, gth = s.state()
    rv, rth = [x.copy() for x in v0], th0.copy()
    for _ in range(8):
        og.rotbouss_step(rv, rth, dt, nu, kappa, xmom, xtemp, omegaz, f=f, fs=fs[0], ord=ord)
    for k in range(3):
        assert rel(to_np(gv[k]), rv[k]) < 1e-12
    assert rel(to_np(gth), rth) < 1e-12
    assert to_np(gth)[0, 0, 0] == th0[0, 0, 0]


def test_rotbouss_anisotropic_pruned_equals_full(gpu):
    from ghost_amd.solvers import BoussSolver
    g, og = grids((32, 32, 16), True, (1.0, 1.0, 0.5))
    v0, th0 = bouss_ic(og, 47)
    outs = []
    for env in (None, "1"):
        if env:
            os.environ["GHX_NO_PRUNE"] = env
        try:
            s = BoussSolver(g, 2e-3, 4e-3, 4e-3, 1.0, 1.0, ord=2)
            s.set_rotation(2.5)
            s.set_state([to_gpu(x) for x in v0], to_gpu(th0))
            s.set_forcing(None, None)
            s.step(3)
            gv, gth = s.state()
            outs.append([to_np(x) for x in gv] + [to_np(gth)])
        finally:
            os.environ.pop("GHX_NO_PRUNE", None)
    rv, rth = [x.copy() for x in v0], th0.copy()
    for _ in range(3):
        og.rotbouss_step(rv, rth, 2e-3, 4e-3, 4e-3, 1.0, 1.0, 2.5, ord=2)
    for k in range(4):
        assert rel(outs[0][k], outs[1][k]) < 1e-14
        assert rel(outs[0][k], (rv + [rth])[k]) < 1e-12


def test_passive_scalar(gpu):
    """PHD (include/phd_rkstep2.f90): the scalar is carried by the flow and
    does not act on it."""
    from ghost_amd.solvers import HDSolver, PHDSolver
    g, og = grids((32, 32, 32))
    v0, th0 = bouss_ic(og, 48)
    f, _ = og.initialfv_random(2.0, 3.0, 0.3, 49)
    fs, _ = og.initialfv_random(1.0, 2.0, 0.2, 50)
    dt, nu, kappa = 3e-3, 4e-3, 2e-3
    s = PHDSolver(g, dt, nu, kappa, ord=2)
    s.set_state([to_gpu(x) for x in v0], to_gpu(th0))
    s.set_forcing([to_gpu(x) for x in f], to_gpu(fs[0]))
    s.step(6)
    gv, gth = s.state()
    rv, rth = [x.copy() for x in v0], th0.copy()
    for _ in range(6):
        og.bouss_step(rv, rth, dt, nu, kappa, 0.0, 0.0, f=f, fs=fs[0], ord=2)
    for k in range(3):
        assert rel(to_np(gv[k]), rv[k]) < 1e-12
    assert rel(to_np(gth), rth) < 1e-12
    h = HDSolver(g, dt, nu, ord=2)
    h.set_velocity([to_gpu(x) for x in v0])
    h.set_forcing([to_gpu(x) for x in f])
    h.step(6)
    for k, x in enumerate(h.velocity()):
        assert rel(to_np(x), to_np(gv[k])) < 1e-14
