"""Parity of the HIP Boussinesq path (advect3 and the fused solver, through
the C ABI) with the CPU oracle's restatement of pseudospec3D_phd.f90 and
include/bouss_rkstep2.f90 on seeded inputs.  Tolerances as for HD: relative
L2 1e-13 for one operator, 1e-12 after several steps."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def to_gpu(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def to_np(t):
    return t.detach().cpu().numpy()


def grids(shape, arbsize=False, L=(1.0, 1.0, 1.0)):
    from ghost_amd.pseudospec import Grid
    from oracle import oracle as O
    nx, ny, nz = shape
    return Grid(nx, ny, nz, L=L, arbsize=arbsize), O.Grid(nx, ny, nz, L=L, arbsize=arbsize)


@pytest.mark.parametrize("shape,arb,L", [((16, 16, 16), False, (1, 1, 1)), ((32, 24, 16), True, (1.0, 1.5, 0.5)),
                                          ((64, 64, 64), False, (1, 1, 1))])
def test_advect3(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    rng = np.random.default_rng(3)
    f = [og.r2c(rng.standard_normal(og.rshape)) for _ in range(4)]
    want = og.advect3(*f)
    out = g.czeros()
    g.advect3(*[to_gpu(x) for x in f], out)
    assert rel(to_np(out), want) < 1e-13


def bouss_ic(og, seed, mean=0.4):
    v, _ = og.initialv_k4exp(3.0, 1.0, seed)
    t, _ = og.initialv_k4exp(2.5, 0.5, seed + 1)
    th = t[0].copy()
    th[0, 0, 0] = mean * og.nx * og.ny * og.nz  # a mean temperature, carried by the k = 0 rule
    return v, th


@pytest.mark.parametrize("ord", [1, 2])
def test_bouss_forced_steps(gpu, ord):
    from ghost_amd.solvers import BoussSolver
    g, og = grids((32, 32, 32))
    v0, th0 = bouss_ic(og, 11)
    f, _ = og.initialfv_random(2.0, 3.0, 0.3, 12)
    fs, _ = og.initialfv_random(1.0, 2.0, 0.2, 13)
    dt, nu, kappa, xmom, xtemp = 3e-3, 4e-3, 5e-3, 2.0, 2.0
    s = BoussSolver(g, dt, nu, kappa, xmom, xtemp, ord=ord)
    s.set_state([to_gpu(x) for x in v0], to_gpu(th0))
    s.set_forcing([to_gpu(x) for x in f], to_gpu(fs[0]))
    s.step(8)
    gv, gth = s.state()
    rv, rth = [x.copy() for x in v0], th0.copy()
    for _ in range(8):
        og.bouss_step(rv, rth, dt, nu, kappa, xmom, xtemp, f=f, fs=fs[0], ord=ord)
    for k in range(3):
        assert rel(to_np(gv[k]), rv[k]) < 1e-12
    assert rel(to_np(gth), rth) < 1e-12
    assert to_np(gth)[0, 0, 0] == rth[0, 0, 0] == th0[0, 0, 0]


def test_bouss_anisotropic_not_dealiased_start(gpu):
    """Every mode populated at the start (full first step), anisotropic box."""
    from ghost_amd.solvers import BoussSolver
    L = (1.0, 1.0, 0.5)
    g, og = grids((32, 32, 16), True, L)
    rng = np.random.default_rng(21)
    v0 = [og.r2c(rng.standard_normal(og.rshape)) * 1e-3 for _ in range(3)]
    th0 = og.r2c(rng.standard_normal(og.rshape)) * 1e-3
    dt, nu, kappa, xmom, xtemp = 2e-3, 5e-3, 5e-3, 1.5, 1.5
    s = BoussSolver(g, dt, nu, kappa, xmom, xtemp, ord=2)
    s.set_state([to_gpu(x) for x in v0], to_gpu(th0))
    s.set_forcing(None, None)
    s.step(4)
    gv, gth = s.state()
    rv, rth = [x.copy() for x in v0], th0.copy()
    for _ in range(4):
        og.bouss_step(rv, rth, dt, nu, kappa, xmom, xtemp, ord=2)
    for k in range(3):
        assert rel(to_np(gv[k]), rv[k]) < 1e-12
    assert rel(to_np(gth), rth) < 1e-12


def test_bouss_pruned_equals_full(gpu):
    from ghost_amd.solvers import BoussSolver
    g, og = grids((48, 48, 48))
    v0, th0 = bouss_ic(og, 31)
    outs = []
    for env in (None, "1"):
        if env:
            os.environ["GHX_NO_PRUNE"] = env
        try:
            s = BoussSolver(g, 2e-3, 4e-3, 4e-3, 1.0, 1.0, ord=2)
            s.set_state([to_gpu(x) for x in v0], to_gpu(th0))
            s.set_forcing(None, None)
            s.step(3)
            gv, gth = s.state()
            outs.append([to_np(x) for x in gv] + [to_np(gth)])
        finally:
            os.environ.pop("GHX_NO_PRUNE", None)
    for k in range(4):
        assert rel(outs[0][k], outs[1][k]) < 1e-14


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 32, 16)])
def test_bouss_ord3_steps_match_oracle(gpu, shape):
    """ord = 3 adds bouss_rkstep2.f90:73-145's correction after the last
    substep.  A mean temperature makes its buoyancy land in the k = 0 mode of
    v_z; the correction is not masked, so the modes outside the dealias set
    are non-zero after each step and the next one runs unpruned."""
    from ghost_amd.solvers import BoussSolver
    g, og = grids(shape)
    v0, th0 = bouss_ic(og, 41)
    f, _ = og.initialfv_random(2.0, 3.0, 0.3, 42)
    fs, _ = og.initialfv_random(1.0, 2.0, 0.2, 43)
    dt, nu, kappa, xmom, xtemp = 3e-3, 4e-3, 5e-3, 2.0, 2.0
    s = BoussSolver(g, dt, nu, kappa, xmom, xtemp, ord=3)
    s.set_state([to_gpu(x) for x in v0], to_gpu(th0))
    s.set_forcing([to_gpu(x) for x in f], to_gpu(fs[0]))
    s.step(5)
    gv, gth = s.state()
    rv, rth = [x.copy() for x in v0], th0.copy()
    for _ in range(5):
        og.bouss_step(rv, rth, dt, nu, kappa, xmom, xtemp, f=f, fs=fs[0], ord=3)
    for k in range(3):
        assert rel(to_np(gv[k]), rv[k]) < 1e-12, k
    assert rel(to_np(gth), rth) < 1e-12
    # the correction is there: its k = 0 buoyancy term on v_z is non-zero
    assert abs(rv[2][0, 0, 0]) > 0
    assert abs(to_np(gv[2])[0, 0, 0] - rv[2][0, 0, 0]) <= 1e-12 * abs(rv[2][0, 0, 0])
