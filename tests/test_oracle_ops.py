"""The oracle's pseudospectral operators and time step, pinned by closed forms,
analytic real-space fields transformed with numpy, and invariants that a
dropped term, a wrong sign or a transposed index would break."""
import numpy as np
import pytest

from oracle import oracle as O


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb else 1.0)


def np_r2c(r):
    return np.ascontiguousarray(np.transpose(np.fft.rfftn(r), (2, 1, 0)))


def coords(nx, ny, nz, L=(1.0, 1.0, 1.0)):
    z, y, x = np.meshgrid(np.arange(nz) * 2 * np.pi * L[2] / nz, np.arange(ny) * 2 * np.pi * L[1] / ny,
                          np.arange(nx) * 2 * np.pi * L[0] / nx, indexing="ij")
    return x, y, z


def test_grid_constants_cubic_and_box():
    g = O.Grid(64)
    assert g.anis == 0 and g.nmax == 64 and g.Dkk == 1.0
    assert g.kmax == (1.0 / 9.0) * 64.0 ** 2
    assert g.tiny == min(1e-5, 0.1 / 64 ** 2)
    assert list(g.kx[:3]) == [0.0, 1.0, 2.0] and g.kx[32] == -32.0 and g.kx[-1] == -1.0
    # the spectral arrays stop at index nx/2, where GHOST's kx carries the negative sign
    assert g.kk2[32, 0, 0] == 32.0 ** 2
    # 2/3 rule with spherical truncation: |k|^2 <= 455 kept for 64^3
    kept = (g.kn2 <= g.kmax) & (g.kn2 >= g.tiny)
    assert kept[21, 0, 0] and not kept[22, 0, 0] and kept[15, 15, 0] and not kept[16, 15, 0]
    b = O.Grid(32, 32, 16, L=(1.0, 1.0, 0.5), arbsize=True)
    assert b.anis == 1 and b.kmax == 1.0 / 9.0 and b.Dkk == 1.0
    assert b.nmax == 32
    assert b.kz[1] == 2.0 and b.kx[1] == 1.0  # kz scaled by Dkz = 1/Lz
    assert b.kn2[0, 0, 1] == (1.0 / 16 ** 2) * 1.0  # kn2 in index units, before box scaling
    assert b.kk2[0, 0, 1] == 4.0


@pytest.mark.parametrize("dir_", [1, 2, 3])
def test_derivk3_single_mode(dir_):
    nx, ny, nz = 16, 12, 8
    g = O.Grid(nx, ny, nz)
    x, y, z = coords(nx, ny, nz)
    a, b, c = 3, -2, 1
    u = np.cos(a * x + b * y + c * z) + 0.5 * np.sin(2 * x - y)
    du = {1: -a * np.sin(a * x + b * y + c * z) + np.cos(2 * x - y),
          2: -b * np.sin(a * x + b * y + c * z) - 0.5 * np.cos(2 * x - y),
          3: -c * np.sin(a * x + b * y + c * z)}[dir_]
    assert rel(g.derivk3(np_r2c(u), dir_), np_r2c(du)) < 1e-14


def test_laplak3_and_rotor3_closed_forms():
    nx = ny = nz = 16
    g = O.Grid(nx)
    x, y, z = coords(nx, ny, nz)
    f = np.sin(2 * x) * np.cos(3 * y) * np.cos(z)
    assert rel(g.laplak3(np_r2c(f)), np_r2c(-(4 + 9 + 1) * f)) < 1e-14
    # A = (sin y, sin z, sin x) -> curl A = (-cos z, -cos x, -cos y)
    A = [np_r2c(np.sin(y)), np_r2c(np.sin(z)), np_r2c(np.sin(x))]
    cu = g.curl(*A)
    for got, want in zip(cu, (-np.cos(z), -np.cos(x), -np.cos(y))):
        assert rel(got, np_r2c(want)) < 1e-14


def test_nonlhd3_is_minus_the_leray_projection():
    g = O.Grid(16, 12, 10)
    rng = np.random.default_rng(1)
    a = [np_r2c(rng.standard_normal(g.rshape)) for _ in range(3)]
    gx, gy, gz = (g.nonlhd3(*a, d) for d in (1, 2, 3))
    KX, KY, KZ = np.meshgrid(g.kx[: g.nxh], g.ky, g.kz, indexing="ij")
    div = KX * gx + KY * gy + KZ * gz
    scale = np.max(np.abs(a[0])) * 16
    assert np.max(np.abs(div)) < 1e-12 * scale
    # a pure gradient a = k phi is removed entirely (away from k = 0)
    phi = np_r2c(rng.standard_normal(g.rshape))
    grad = [KX * phi, KY * phi, KZ * phi]
    gg = [g.nonlhd3(*grad, d) for d in (1, 2, 3)]
    gg[0][0, 0, 0] = gg[1][0, 0, 0] = gg[2][0, 0, 0] = 0
    assert max(np.max(np.abs(q)) for q in gg) < 1e-12 * np.max(np.abs(KX * phi))
    # idempotence of the projection P = -nonlhd3
    again = [g.nonlhd3(-gx, -gy, -gz, d) for d in (1, 2, 3)]
    for p, q in zip(again, (gx, gy, gz)):
        assert rel(p, q) < 1e-14
    # nonlin3(a, d) = nonlhd3(a - d); gauge3(a) = -nonlhd3(a)
    d = [np_r2c(rng.standard_normal(g.rshape)) for _ in range(3)]
    for k in (1, 2, 3):
        diff = [a[0] - d[0], a[1] - d[1], a[2] - d[2]]
        assert rel(g.nonlin3(*a, *d, k), g.nonlhd3(*diff, k)) < 1e-14
        assert rel(g.gauge3(*a, k), -g.nonlhd3(*a, k)) < 1e-15


def tg_fields(n):
    x, y, z = coords(n, n, n)
    u = [np.sin(x) * np.cos(y) * np.cos(z), -np.cos(x) * np.sin(y) * np.cos(z), np.zeros_like(x)]
    w = [-np.cos(x) * np.sin(y) * np.sin(z), -np.sin(x) * np.cos(y) * np.sin(z),
         2 * np.sin(x) * np.sin(y) * np.cos(z)]
    return u, w


def test_prodre3_taylor_green_closed_form():
    """prodre3 of the TG field = FFT of (curl u) x u evaluated analytically."""
    n = 16
    g = O.Grid(n)
    u, w = tg_fields(n)
    vhat = [np_r2c(c) for c in u]
    cross = [w[1] * u[2] - u[1] * w[2], w[2] * u[0] - u[2] * w[0], w[0] * u[1] - u[0] * w[1]]
    got = g.prodre3(*vhat)
    for k in range(3):
        assert rel(got[k], np_r2c(cross[k])) < 1e-14


def test_vector3_matches_real_space_product():
    g = O.Grid(12, 16, 8)
    rng = np.random.default_rng(2)
    A = [rng.standard_normal(g.rshape) for _ in range(3)]
    B = [rng.standard_normal(g.rshape) for _ in range(3)]
    cross = np.cross(np.stack(A, -1), np.stack(B, -1))
    got = g.vector3(*[np_r2c(a) for a in A], *[np_r2c(b) for b in B])
    for k in range(3):
        assert rel(got[k], np_r2c(cross[..., k])) < 1e-14


def test_nonlinear_term_conserves_energy_pointwise():
    """u . (w x u) = 0 on every grid point, so sum_k Re(u_hat . conj(N_hat)) = 0."""
    g = O.Grid(24)
    v, _ = g.initialv_k4exp(3.0, 1.0, 99)
    d = g.prodre3(*v)
    w = np.full(g.nxh, 2.0)
    w[0] = 1.0
    s = sum(np.sum(w[:, None, None] * np.real(v[k] * np.conj(d[k]))) for k in range(3))
    scale = sum(np.sum(np.abs(v[k]) * np.abs(d[k])) for k in range(3))
    assert abs(s) < 1e-13 * scale


def test_taylor_green_diagnostics():
    g = O.Grid(32)
    v = g.initialv_tg(1.0, 1.0, 0.5)
    u, w = tg_fields(32)
    # u0 = 0.5 reproduces the unit-amplitude TG field: <u^2> = 1/4, <w^2> = 3/4, H = 0
    assert abs(g.energy(*v, 1) - 0.25) < 1e-15
    assert abs(g.energy(*v, 0) - 0.75) < 1e-14
    assert abs(g.helicity(*v)) < 1e-15
    for k in range(3):
        assert rel(v[k], np_r2c(u[k])) < 1e-14 if k < 2 else np.max(np.abs(v[k])) == 0
    # all energy in the shell int(sqrt(3) + .501) = 2
    Ek, _ = g.spectrumc(*v, kin=1, hel=0)
    assert abs(Ek[1] - 0.25) < 1e-15 and abs(np.sum(Ek) - Ek[1]) < 1e-15
    # max |u| of TG is 1 at x = pi/2, y = z = 0 (on the grid)
    assert abs(g.maxabs(*v, 2) - 1.0) < 1e-14
    assert g.divergence(*v) < 1e-28
    # lap energy (kin=2) = <|lap u|^2> = 9 <u^2>
    assert abs(g.energy(*v, 2) - 9 * 0.25) < 1e-13


def test_cross_and_energy_consistency():
    g = O.Grid(16, 12, 10)
    rng = np.random.default_rng(3)
    a = [np_r2c(rng.standard_normal(g.rshape)) for _ in range(3)]
    b = [np_r2c(rng.standard_normal(g.rshape)) for _ in range(3)]
    assert abs(g.cross(*a, *a, 1) - g.energy(*a, 1)) < 1e-13 * g.energy(*a, 1)
    # GHOST's weights double the kx=Nyquist plane, so compare with the same weights
    N = 16 * 12 * 10
    w = np.full(g.nxh, 2.0)
    w[0] = 1.0
    ref = sum(np.sum(w[:, None, None] * np.real(a[k] * np.conj(b[k]))) for k in range(3)) / N ** 2
    assert abs(g.cross(*a, *b, 1) - ref) < 1e-13 * abs(ref) + 1e-15


def test_spectrum_bins_and_offset():
    """Shell index int(sqrt(kk2)/Dkk + .501) with the single-precision .501."""
    g = O.Grid(16)
    c = [g.czeros() for _ in range(3)]
    # put energy on single modes and check their bins
    for (i, j, k), shell in (((3, 0, 0), 3), ((2, 2, 0), 3), ((1, 1, 1), 2), ((1, 2, 2), 3), ((2, 3, 1), 4)):
        c[0][:] = 0
        c[0][i, j, k] = 16 ** 3
        Ek, _ = g.spectrumc(*c, kin=1, hel=0)
        assert np.argmax(Ek) == shell - 1, (i, j, k)
    assert abs(float(np.float32(0.501)) - 0.501) > 1e-9  # the offset is not the double 0.501


def test_rk2_linear_decay_is_exact():
    """A shear mode u = (sin(kz z), 0, 0) has a gradient nonlinear term, removed by
    the projection: RK2 multiplies it by 1 + q + q^2/2 with q = -nu kz^2 dt."""
    n = 16
    g = O.Grid(n)
    x, y, z = coords(n, n, n)
    kz, nu, dt = 3, 0.05, 0.01
    v = [np_r2c(np.sin(kz * z)), g.czeros(), g.czeros()]
    v0 = v[0].copy()
    g.hd_step(v, dt, nu, ord=2)
    q = -nu * kz ** 2 * dt
    assert rel(v[0], v0 * (1 + q + q * q / 2)) < 1e-14
    assert np.max(np.abs(v[1])) < 1e-12 * np.max(np.abs(v0)) and np.max(np.abs(v[2])) < 1e-12 * np.max(np.abs(v0))


def test_ord_matches_low_storage_definition():
    """ord = 3 substeps: v = C + dt (L v + N(v) + f)/o for o = 3, 2, 1 (hd_rkstep2.f90)."""
    g = O.Grid(12)
    v, _ = g.initialv_k4exp(2.0, 1.0, 5)
    ref = [x.copy() for x in v]
    g.hd_step(v, 1e-3, 1e-2, ord=3)
    c = [x.copy() for x in ref]
    for o in (3, 2, 1):
        g.hd_substep(ref, c, o, 1e-3, 1e-2)
    for k in range(3):
        assert np.array_equal(v[k], ref[k])


def test_inviscid_invariants():
    """nu = 0, no forcing: with the spherical truncation the semi-discrete system
    conserves energy and helicity; RK2 drifts at O(dt^3) per step."""
    g = O.Grid(24)
    v, _ = g.initialv_k4exp(3.0, 1.0, 123)
    e0, h0 = g.energy(*v, 1), g.helicity(*v)
    drifts = []
    for dt in (1e-3, 5e-4):
        w = [x.copy() for x in v]
        for _ in range(5):
            g.hd_step(w, dt, 0.0, ord=2)
        drifts.append(abs(g.energy(*w, 1) - e0) / e0)
    assert drifts[0] < 1e-9
    assert drifts[1] < drifts[0] / 4  # third order per step: halving dt divides it by ~8
    w = [x.copy() for x in v]
    for _ in range(5):
        g.hd_step(w, 1e-3, 0.0, ord=2)
    assert abs(g.helicity(*w) - h0) < 1e-5 * max(abs(h0), e0)


def test_energy_budget_balance():
    """GHOST's balance.txt rule: d<u^2>/dt = -2 nu <w^2> + 2 <u.f> (forced, viscous)."""
    g = O.Grid(16)
    v, _ = g.initialv_k4exp(2.0, 1.0, 7)
    f, _ = g.initialfv_random(1.0, 2.0, 0.5, 11)
    nu, dt = 2e-2, 1e-4
    e0 = g.energy(*v, 1)
    rate = -2 * nu * g.energy(*v, 0) + 2 * g.cross(*v, *f, 1)
    w = [x.copy() for x in v]
    g.hd_step(w, dt, nu, f=f, ord=2)
    de = (g.energy(*w, 1) - e0) / dt
    assert abs(de - rate) < 2e-3 * abs(rate)


def test_mhd_step_ideal_invariants():
    """nu = mu = 0: total energy <u^2 + b^2> and cross helicity <u.b> conserved to O(dt^2)."""
    g = O.Grid(16, 16, 8, L=(1.0, 1.0, 0.5), arbsize=True)
    v, _ = g.initialv_k4exp(2.0, 1.0, 3)
    a, _ = g.initialv_k4exp(2.0, 0.3, 4)

    def invariants(v, a):
        b = g.curl(*a)
        return g.energy(*v, 1) + g.energy(*a, 0), g.cross(*v, *b, 1)

    e0, x0 = invariants(v, a)
    for _ in range(4):
        g.mhd_step(v, a, 5e-4, 0.0, 0.0, ord=2)
    e1, x1 = invariants(v, a)
    assert abs(e1 - e0) < 1e-5 * e0
    assert abs(x1 - x0) < 1e-5 * e0


def test_sharp_filter():
    g = O.Grid(16)
    a = np_r2c(np.random.default_rng(4).standard_normal(g.rshape))
    b = g.sharp(a, np.pi / 3.0)
    keep = g.kn2 <= 9.0
    assert np.array_equal(b[keep], a[keep]) and np.all(b[~keep] == 0)


def test_truncate_regrids_band_limited_field_exactly():
    """trunc3D: a field whose modes all satisfy the new-grid bound comes back
    sampled on the coarse grid (every (n/nt)-th point); modes beyond the bound
    are removed, as a direct evaluation of the kept Fourier series shows."""
    n, nt = 32, 16
    g, t = O.Grid(n), O.Grid(nt)
    z, y, x = np.meshgrid(*(np.arange(n) * 2 * np.pi / n,) * 3, indexing="ij")
    low = np.sin(x) * np.cos(2 * y) + 0.3 * np.cos(3 * z + 1.0) + 0.2 * np.sin(2 * x + y - z)
    high = 0.5 * np.cos(7 * x) * np.sin(6 * z)  # kn2_new = (49 + 36)/256 > 1/9
    out = t.c2r(g.truncate(g.r2c(low + high), t))
    assert np.max(np.abs(out - low[::2, ::2, ::2])) < 1e-13
