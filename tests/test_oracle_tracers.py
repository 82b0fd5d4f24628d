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
