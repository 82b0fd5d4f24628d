"""Parity of the HIP Lagrangian-tracer path (through the C ABI) with the CPU
oracle's restatement of GHOST's lagpart module on seeded inputs.

The HIP path takes the spline control points from a spectral prefilter, the
oracle from GHOST's cyclic tridiagonal sweeps; on grids with n >= 32 per axis
both are the exact spline to rounding.  Tolerances: relative L2 1e-12 on
interpolated values and fields, 1e-10 grid units (periodic distance) on
positions after several RK steps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def pdist(a, b, n):
    """Largest periodic distance between two position arrays on [0, n)."""
    d = np.mod(np.asarray(a) - np.asarray(b) + n / 2, n) - n / 2
    return float(np.max(np.abs(d))) if d.size else 0.0


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


def seed_positions(shape, n, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.0, shape[c], n) for c in range(3)]


def edge_positions(shape):
    """Nodes, cell edges and the top z cell that GHOST's stencil clamps."""
    nx, ny, nz = shape
    px = np.array([0.0, 1.0, nx - 1.0, nx - 1e-12, 0.5, 3.25, 0.0, nx / 2])
    py = np.array([0.0, 2.0, ny - 1e-12, 0.0, ny - 1.0, 0.75, 5.5, ny / 3])
    pz = np.array([0.0, nz - 1.0, nz - 0.5, nz - 1e-12, 1.0, 2.0, nz - 1.5, 0.25])
    return [px, py, pz]


@pytest.mark.parametrize("zcompat", [True, False])
@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 48, 32), (64, 64, 64), (32, 32, 512)])
def test_euler_to_lag(gpu, shape, zcompat):
    """GPart_EulerToLag of a random real field: spline control points + Interp3D."""
    from ghost_amd.lagpart import GPart
    g, og = grids(shape)
    field = np.random.default_rng(1).standard_normal(og.rshape)
    pos = seed_positions(shape, 3000, 2)
    for c in range(3):
        pos[c] = np.concatenate([pos[c], edge_positions(shape)[c]])
    parts = GPart(g, len(pos[0]), zcompat=zcompat)
    parts.set_positions(*[to_gpu(p) for p in pos])
    got = to_np(parts.euler_to_lag(to_gpu(field)))
    want = og.spline_interp(og.spline_coef(field, ghost=True), *pos, zcompat=zcompat)
    assert rel(got, want) < 1e-12


@pytest.mark.parametrize("zcompat", [True, False])
def test_tracer_stage(gpu, zcompat):
    """One stage (main3D.fpp:3233-3267 + GPart_StepRKK) on a k^4 exp field."""
    from ghost_amd.lagpart import GPart
    shape = (64, 64, 64)
    g, og = grids(shape)
    v, _ = og.initialv_k4exp(3.0, 1.0, 17)
    pos = seed_positions(shape, 5000, 3)
    parts = GPart(g, 5000, zcompat=zcompat)
    parts.set_positions(*[to_gpu(p) for p in pos])
    parts.set_step()
    parts.step([to_gpu(x) for x in v], 3e-3, 0.5)
    ref = [p.copy() for p in pos]
    lv = og.tracer_stage(v, ref, [p.copy() for p in pos], 3e-3, 0.5, zcompat=zcompat)
    got = [to_np(p) for p in parts.positions()]
    gu = np.array([to_np(u) for u in parts.velocities()])
    assert rel(gu, lv) < 1e-12
    for c in range(3):
        assert pdist(got[c], ref[c], shape[c]) < 1e-10
        assert np.all((got[c] >= 0) & (got[c] < shape[c]))


@pytest.mark.parametrize("zcompat", [True, False])
def test_hd_with_tracers_forced(gpu, zcompat):
    """HD RK2 with forcing and tracers attached: one tracer stage after each
    substep on the velocity it produced; fields and positions after 6 steps."""
    from ghost_amd.lagpart import GPart
    from ghost_amd.solvers import HDSolver
    shape = (32, 32, 32)
    g, og = grids(shape)
    v0, _ = og.initialv_k4exp(3.0, 1.0, 21)
    f0, _ = og.initialfv_random(2.0, 3.0, 0.5, 22)
    pos = seed_positions(shape, 2000, 4)
    dt, nu = 4e-3, 3e-3
    s = HDSolver(g, dt=dt, nu=nu, ord=2)
    s.set_velocity([to_gpu(x) for x in v0])
    s.set_forcing([to_gpu(x) for x in f0])
    parts = GPart(g, 2000, zcompat=zcompat)
    parts.set_positions(*[to_gpu(p) for p in pos])
    s.set_tracers(parts)
    s.step(6)
    ref_v, ref_p = [x.copy() for x in v0], [p.copy() for p in pos]
    for _ in range(6):
        og.hd_step_tracers(ref_v, ref_p, dt, nu, f=f0, ord=2, zcompat=zcompat)
    got_v = s.velocity()
    for k in range(3):
        assert rel(to_np(got_v[k]), ref_v[k]) < 1e-12
    got_p = [to_np(p) for p in parts.positions()]
    for c in range(3):
        assert pdist(got_p[c], ref_p[c], shape[c]) < 1e-10
    # the tracers moved: the check above is not a no-op
    assert max(pdist(got_p[c], pos[c], shape[c]) for c in range(3)) > 1e-2


def test_hd_with_tracers_anisotropic_box(gpu):
    """Tracers on a 32x48x64 box with L = (1, 1.5, 2) (units of 2 pi), RK3."""
    from ghost_amd.lagpart import GPart
    from ghost_amd.solvers import HDSolver
    shape, L = (32, 48, 64), (1.0, 1.5, 2.0)
    g, og = grids(shape, True, L)
    v0, _ = og.initialv_k4exp(3.0, 1.0, 31)
    pos = seed_positions(shape, 1500, 5)
    dt, nu = 3e-3, 4e-3
    s = HDSolver(g, dt=dt, nu=nu, ord=3)
    s.set_velocity([to_gpu(x) for x in v0])
    s.set_forcing(None)
    parts = GPart(g, 1500, zcompat=True)
    parts.set_positions(*[to_gpu(p) for p in pos])
    s.set_tracers(parts)
    s.step(4)
    ref_v, ref_p = [x.copy() for x in v0], [p.copy() for p in pos]
    for _ in range(4):
        og.hd_step_tracers(ref_v, ref_p, dt, nu, ord=3)
    got_p = [to_np(p) for p in parts.positions()]
    for c in range(3):
        assert pdist(got_p[c], ref_p[c], shape[c]) < 1e-10


def test_hd_with_tracers_128_pruned_and_edges(gpu):
    """128^3 with the pruned passes, random tracers plus the edge positions."""
    from ghost_amd.lagpart import GPart
    from ghost_amd.solvers import HDSolver
    shape = (128, 128, 128)
    g, og = grids(shape)
    v0, _ = og.initialv_k4exp(3.0, 1.0, 41)
    pos = seed_positions(shape, 20000, 6)
    for c in range(3):
        pos[c] = np.concatenate([pos[c], edge_positions(shape)[c]])
    n = len(pos[0])
    dt, nu = 2e-3, 1e-3
    s = HDSolver(g, dt=dt, nu=nu, ord=2)
    s.set_velocity([to_gpu(x) for x in v0])
    s.set_forcing(None)
    parts = GPart(g, n, zcompat=True)
    parts.set_positions(*[to_gpu(p) for p in pos])
    s.set_tracers(parts)
    s.step(2)
    ref_v, ref_p = [x.copy() for x in v0], [p.copy() for p in pos]
    for _ in range(2):
        og.hd_step_tracers(ref_v, ref_p, dt, nu, ord=2)
    got_p = [to_np(p) for p in parts.positions()]
    for c in range(3):
        assert pdist(got_p[c], ref_p[c], shape[c]) < 1e-10


def test_mhd_with_tracers_leaves_the_flow_unchanged(gpu):
    """Tracers attached to the MHD solver advect with its velocity and do not
    change the fields; on a steady uniform flow they translate exactly."""
    from ghost_amd.lagpart import GPart
    from ghost_amd.solvers import MHDSolver
    shape = (32, 32, 32)
    g, og = grids(shape)
    U = (0.3, -0.2, 0.5)
    v0 = [og.czeros() for _ in range(3)]
    a0 = [og.czeros() for _ in range(3)]
    # a uniform flow is the k = 0 mode, which the dealias set excludes; use a
    # weak shear in a instead and compare fields with and without tracers
    a0[0][0, 1, 0] = 1e-3 * 32 ** 3
    a0[0][0, 31, 0] = 1e-3 * 32 ** 3
    v1, _ = og.initialv_k4exp(3.0, 0.5, 51)
    outs = []
    for attach in (False, True):
        s = MHDSolver(g, dt=2e-3, nu=3e-3, mu=3e-3, ord=2)
        s.set_state([to_gpu(x) for x in v1], [to_gpu(x) for x in a0])
        s.set_forcing(None, None)
        if attach:
            parts = GPart(g, 500)
            parts.set_positions(*[to_gpu(p) for p in seed_positions(shape, 500, 7)])
            s.set_tracers(parts)
        s.step(3)
        outs.append([to_np(x) for x in s.state()[0]])
    for k in range(3):
        assert np.array_equal(outs[0][k], outs[1][k])
    # standalone stage on a uniform flow: exact translation
    parts = GPart(g, 64, zcompat=True)
    pos = seed_positions(shape, 64, 8)
    parts.set_positions(*[to_gpu(p) for p in pos])
    parts.set_step()
    for c in range(3):
        v0[c][0, 0, 0] = U[c] * 32 ** 3
    parts.step([to_gpu(x) for x in v0], 0.01, 1.0)
    got = [to_np(p) for p in parts.positions()]
    for c in range(3):
        want = np.mod(pos[c] + 0.01 * 32 / (2 * np.pi) * U[c], 32)
        assert pdist(got[c], want, 32) < 1e-12


def test_hd_512_step_with_tracers_matches_oracle(gpu):
    """The benchmark configuration at its full size: one forced RK2 step at
    512^3 with 20000 tracers, against the oracle on the same inputs."""
    from ghost_amd.lagpart import GPart
    from ghost_amd.solvers import HDSolver
    from oracle import oracle as O
    O.set_threads(16)
    shape = (512, 512, 512)
    g, og = grids(shape)
    v0, _ = og.initialv_k4exp(3.0, 1.0, 1000)
    f0, _ = og.initialfv_random(2.0, 3.0, 0.4, 2000)
    pos = seed_positions(shape, 20000, 3000)
    dt, nu = 2.5e-4, 2e-4
    s = HDSolver(g, dt=dt, nu=nu, ord=2)
    s.set_velocity([to_gpu(x) for x in v0])
    s.set_forcing([to_gpu(x) for x in f0])
    parts = GPart(g, len(pos[0]), zcompat=True)
    parts.set_positions(*[to_gpu(p) for p in pos])
    s.set_tracers(parts)
    s.step(1)
    got_v = [to_np(x) for x in s.velocity()]
    got_p = [to_np(p) for p in parts.positions()]
    del s, parts
    og.hd_step_tracers(v0, pos, dt, nu, f=f0, ord=2)
    for k in range(3):
        assert rel(got_v[k], v0[k]) < 1e-12
    for c in range(3):
        assert pdist(got_p[c], pos[c], shape[c]) < 1e-10
