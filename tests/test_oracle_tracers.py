"""The oracle's Lagrangian-tracer restatement (cubic-spline interpolation and
the RK stage of lagpart/), pinned without the oracle's own arithmetic: numpy's
FFT solves the periodic spline system exactly (the circulant (1/6, 2/3, 1/6)
is diagonal in Fourier space), the closed-form cardinal cubic B-spline gives
the interpolant, and steady flows give closed-form tracer paths."""
import numpy as np
import pytest

from oracle import oracle as O


def bspline(u):
    """Centred cardinal cubic B-spline, closed form."""
    a = np.abs(u)
    return np.where(a < 1, 2.0 / 3.0 - a * a + 0.5 * a ** 3, np.where(a < 2, (2 - a) ** 3 / 6.0, 0.0))


def spline_coef_fft(f):
    """Exact periodic spline control points: c_hat = f_hat / (b(kx) b(ky) b(kz)),
    b(k) = 2/3 + cos(2 pi k / n) / 3, the symbol of the (1/6, 2/3, 1/6) stencil."""
    nz, ny, nx = f.shape
    b = [2.0 / 3.0 + np.cos(2 * np.pi * np.arange(n) / n) / 3.0 for n in (nz, ny, nx)]
    sym = b[0][:, None, None] * b[1][None, :, None] * b[2][None, None, :]
    return np.real(np.fft.ifftn(np.fft.fftn(f) / sym))


def interp_bspline(c, p, zshift=0.0):
    """sum_ijk c[k,j,i] B(x-i) B(y-j) B(z+zshift-k) over the periodic images."""
    nz, ny, nx = c.shape
    out = []
    for x, y, z in zip(*p):
        z = z + zshift
        acc = 0.0
        for k in range(int(np.floor(z)) - 2, int(np.floor(z)) + 3):
            wz = bspline(z - k)
            for j in range(int(np.floor(y)) - 2, int(np.floor(y)) + 3):
                wy = bspline(y - j)
                for i in range(int(np.floor(x)) - 2, int(np.floor(x)) + 3):
                    acc += c[k % nz, j % ny, i % nx] * bspline(x - i) * wy * wz
        out.append(acc)
    return np.array(out)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape", [(8, 8, 8), (6, 10, 12), (32, 32, 32), (16, 24, 40)])
def test_exact_spline_solve_matches_fft(shape):
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz)
    f = np.random.default_rng(1).standard_normal((nz, ny, nx))
    assert rel(g.spline_coef(f, ghost=False), spline_coef_fft(f)) < 1e-14


def test_ghost_solve_row_n_slip():
    """GHOST's row-n update uses xx(n-2)*s(n-1) where the elimination needs
    xx(n-2)*s(n-2) (gpsplineint_mod.fpp:1240); xx decays like (2-sqrt 3)^n, so
    the solve is exact to roundoff from n = 32 on and visibly off at n = 8."""
    f8 = np.random.default_rng(2).standard_normal((8, 8, 8))
    g8 = O.Grid(8)
    assert rel(O.Grid(8).spline_coef(f8, ghost=True), spline_coef_fft(f8)) > 1e-5
    assert rel(g8.spline_coef(f8, ghost=False), spline_coef_fft(f8)) < 1e-14
    f = np.random.default_rng(3).standard_normal((32, 32, 32))
    assert rel(O.Grid(32).spline_coef(f, ghost=True), spline_coef_fft(f)) < 1e-14


def test_interp_is_the_cardinal_bspline_sum():
    nx, ny, nz = 8, 12, 10
    g = O.Grid(nx, ny, nz)
    rng = np.random.default_rng(4)
    c = rng.standard_normal((nz, ny, nx))
    n = 40
    p = [rng.uniform(0, nx, n), rng.uniform(0, ny, n), rng.uniform(0, nz - 1, n)]
    p[0][:3] = [0.0, nx - 1e-9, 3.0]  # cell edges and a node
    assert rel(g.spline_interp(c, *p, zcompat=False), interp_bspline(c, p)) < 1e-14
    # GHOST's z stencil interpolates at z + 1 below the top cell
    assert rel(g.spline_interp(c, *p, zcompat=True), interp_bspline(c, p, zshift=1.0)) < 1e-14
    # top cell: the clamp keeps planes nz-2..nz+1; at z = nz-1 it still equals z + 1
    q = [np.array([2.5]), np.array([3.5]), np.array([nz - 1.0])]
    assert abs(g.spline_interp(c, *q, zcompat=True)[0] - interp_bspline(c, q, zshift=1.0)[0]) < 1e-14
    # and inside the top cell it is the extrapolated cubic, not the spline at z + 1
    q[2] = np.array([nz - 0.5])
    assert abs(g.spline_interp(c, *q, zcompat=True)[0] - interp_bspline(c, q, zshift=1.0)[0]) > 1e-6


def test_interp_reproduces_nodes_and_converges_fourth_order():
    g = O.Grid(16)
    f = np.random.default_rng(5).standard_normal(g.rshape)
    c = g.spline_coef(f, ghost=False)
    k, j, i = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    got = g.spline_interp(c, i.ravel() * 1.0, j.ravel() * 1.0, k.ravel() * 1.0, zcompat=False)
    assert np.max(np.abs(got - f.ravel())) < 1e-13
    errs = []
    for n in (32, 64):
        g = O.Grid(n)
        z, y, x = np.meshgrid(*(np.arange(n) * 2 * np.pi / n,) * 3, indexing="ij")
        f = np.sin(x) * np.cos(2 * y) * np.sin(z + 0.3)
        c = g.spline_coef(f)
        rng = np.random.default_rng(6)
        p = [rng.uniform(0, n, 500) for _ in range(3)]
        exact = np.sin(2 * np.pi * p[0] / n) * np.cos(4 * np.pi * p[1] / n) * np.sin(2 * np.pi * p[2] / n + 0.3)
        errs.append(np.max(np.abs(g.spline_interp(c, *p, zcompat=False) - exact)))
    assert 12 < errs[0] / errs[1] < 20  # O(h^4)


def test_tracer_stage_against_fft_spline_and_bspline_sum():
    n = 32
    g = O.Grid(n)
    rng = np.random.default_rng(7)
    v, _ = g.initialv_k4exp(3.0, 1.0, 11)
    npart = 25
    pos = [rng.uniform(0, n, npart) for _ in range(3)]
    pos[2] = np.minimum(pos[2], n - 1.01)  # below the clamped top cell
    x0 = [p.copy() for p in pos]
    dt, xk = 2e-3, 0.5
    lv = g.tracer_stage(v, pos, x0, dt, xk)
    u = [g.c2r(c / n ** 3) for c in v]
    want = np.array([interp_bspline(spline_coef_fft(uc), x0, zshift=1.0) for uc in u])
    assert rel(lv, want) < 1e-13
    for c in range(3):
        moved = np.mod(x0[c] + dt * xk * n / (2 * np.pi) * want[c], n)
        assert np.max(np.abs(pos[c] - moved)) < 1e-11


def test_tracers_in_uniform_flow_translate_exactly():
    n = 32
    g = O.Grid(n)
    U = (0.3, -0.2, 0.7)
    v = [g.czeros() for _ in range(3)]
    for c in range(3):
        v[c][0, 0, 0] = U[c] * n ** 3  # spectral fields are N^3 x Fourier coefficients
    rng = np.random.default_rng(8)
    pos = [rng.uniform(0, n, 30) for _ in range(3)]
    x0 = [p.copy() for p in pos]
    g.tracer_stage(v, pos, x0, 0.01, 1.0)
    for c in range(3):
        want = np.mod(x0[c] + 0.01 * n / (2 * np.pi) * U[c], n)
        assert np.max(np.abs(pos[c] - want)) < 1e-12


@pytest.mark.parametrize("zcompat", [False, True])
def test_tracers_in_steady_shear_flow(zcompat):
    """u = (A sin z, 0, 0) is a steady solution of the Euler equations (its
    Lamb vector is a gradient), so with nu = 0 each tracer moves along x by
    dt A sin(z_phys) per step; GHOST's z stencil samples sin at z + 1 cell."""
    n, A, dt, steps = 32, 0.8, 5e-3, 4
    g = O.Grid(n)
    z = np.arange(n) * 2 * np.pi / n
    ux = np.broadcast_to((A * np.sin(z))[:, None, None], g.rshape).copy()
    v = [g.r2c(ux), g.czeros(), g.czeros()]
    rng = np.random.default_rng(9)
    pos = [rng.uniform(0, n, 50), rng.uniform(0, n, 50), rng.uniform(0, n - 1.5, 50)]
    start = [p.copy() for p in pos]
    for _ in range(steps):
        g.hd_step_tracers(v, pos, dt, 0.0, zcompat=zcompat)
    shift = 1.0 if zcompat else 0.0
    disp = steps * dt * A * np.sin(2 * np.pi * (start[2] + shift) / n) * n / (2 * np.pi)
    assert np.max(np.abs(np.mod(pos[0] - start[0] - disp + n / 2, n) - n / 2)) < 1e-4 * np.max(np.abs(disp))
    assert np.max(np.abs(pos[1] - start[1])) < 1e-12
    assert np.max(np.abs(pos[2] - start[2])) < 1e-12


# The stage coupling of main3D.fpp:3229-3262 with GPart_StepRKK
# (gparts_mod.f90:2046-2145): after RK substep o the fluid holds the substep's
# output v_o, and the tracers move to x = x0 + dt (1/o) u_o(x) with u_o the
# spline interpolant of c2r(v_o / N^3) at their current positions x and x0
# their positions at the start of the step.  For RK2 (o = 2, then 1) and a
# linear mode that the nonlinear term leaves alone, v_2 = (1 + z/2) v and
# v_1 = (1 + z + z^2/2) v with z = -nu |k|^2 dt, so the paths are closed forms.
# A tracer that saw the step's input field, the midpoint field in both stages,
# advanced from x* instead of x0, or was sampled at x0 in the second stage
# would land elsewhere by far more than the tolerance.

def test_tracer_coupling_decaying_shear_at_nodes():
    """u = A sin(z) x: an exact Navier-Stokes solution (its Lamb vector is a
    gradient).  With the tracers on z nodes the spline interpolant is exact, so
    x1 = x0 + dt c A (1 + z + z^2/2) sin(z_node + 1 cell) (GHOST's z stencil)."""
    n, A, nu, dt = 32, 0.8, 2.0, 0.05
    g = O.Grid(n)
    zc = np.arange(n) * 2 * np.pi / n
    ux = np.broadcast_to((A * np.sin(zc))[:, None, None], g.rshape).copy()
    v = [g.r2c(ux), g.czeros(), g.czeros()]
    rng = np.random.default_rng(12)
    npart = 40
    zn = rng.integers(0, n - 2, npart).astype(float)  # nodes below the clamped top cell
    pos = [rng.uniform(0, n, npart), rng.uniform(0, n, npart), zn]
    start = [p.copy() for p in pos]
    g.hd_step_tracers(v, pos, dt, nu, ord=2, zcompat=True)
    lam = -nu * dt
    c = n / (2 * np.pi)
    s = np.sin(2 * np.pi * (zn + 1.0) / n)
    want = np.mod(start[0] + dt * c * A * (1 + lam + lam * lam / 2) * s, n)
    assert np.max(np.abs(pos[0] - want)) < 1e-11
    assert np.max(np.abs(pos[1] - start[1])) < 1e-12 and np.max(np.abs(pos[2] - start[2])) < 1e-12
    # the alternatives differ from GHOST's coupling by far more than the tolerance
    for factor in (1.0, 1 + lam / 2):  # input field in stage 2, midpoint field in stage 2
        alt = np.mod(start[0] + dt * c * A * factor * s, n)
        assert np.max(np.abs(alt - want)) > 1e-4
    alt = np.mod(start[0] + dt * c * A * ((1 + lam / 2) / 2 + (1 + lam + lam * lam / 2)) * s, n)  # from x*
    assert np.max(np.abs(alt - want)) > 1e-2


def test_tracer_coupling_decaying_taylor_green_2d():
    """u = A (sin x cos y, -cos x sin y, 0), an exact Navier-Stokes solution
    that decays at |k|^2 = 2.  Stage 1: x* = x0 + (dt/2) c s(x0) (1 + z/2);
    stage 2: x1 = x0 + dt c s(x*) (1 + z + z^2/2), with s the exact periodic
    spline interpolant of the sampled field (numpy FFT solve, closed-form
    B-spline sum).  Sampling stage 2 at x0 moves the tracers by O(dt^2)."""
    n, A, nu, dt = 32, 1.0, 1.0, 0.05
    g = O.Grid(n)
    zc, yc, xc = np.meshgrid(*(np.arange(n) * 2 * np.pi / n,) * 3, indexing="ij")
    ur = [A * np.sin(xc) * np.cos(yc), -A * np.cos(xc) * np.sin(yc), np.zeros(g.rshape)]
    v = [g.r2c(u) for u in ur]
    coef = [spline_coef_fft(u) for u in ur[:2]]
    rng = np.random.default_rng(13)
    npart = 12
    pos = [rng.uniform(0, n, npart), rng.uniform(0, n, npart), rng.uniform(0, n - 1.5, npart)]
    x0 = [p.copy() for p in pos]
    g.hd_step_tracers(v, pos, dt, nu, ord=2, zcompat=True)
    lam = -2.0 * nu * dt
    c = n / (2 * np.pi)
    f2, f1 = 1 + lam / 2, 1 + lam + lam * lam / 2
    s0 = [interp_bspline(cf, x0, zshift=1.0) for cf in coef]
    xs = [np.mod(x0[q] + 0.5 * dt * c * f2 * s0[q], n) for q in range(2)] + [x0[2].copy()]
    s1 = [interp_bspline(cf, xs, zshift=1.0) for cf in coef]
    for q in range(2):
        want = np.mod(x0[q] + dt * c * f1 * s1[q], n)
        err = np.abs(np.mod(pos[q] - want + n / 2, n) - n / 2)
        assert np.max(err) < 1e-11, q
        at_x0 = np.mod(x0[q] + dt * c * f1 * s0[q], n)  # stage 2 sampled at x0
        assert np.max(np.abs(np.mod(at_x0 - want + n / 2, n) - n / 2)) > 1e-4
    assert np.max(np.abs(pos[2] - x0[2])) < 1e-12
