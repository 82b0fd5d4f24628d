"""The oracle's Boussinesq restatement (advect3 and include/bouss_rkstep2.f90),
pinned by closed forms: the advection of a single product of Fourier modes,
and a single-mode internal gravity wave, which is an exact solution of the
nonlinear equations (u . grad u = 0 and u . grad th = 0 for one Fourier mode),
so the RK step must reproduce the polynomial of the 2x2 linear operator."""
import numpy as np
import pytest

from oracle import oracle as O


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_advect3_closed_form():
    n = 16
    g = O.Grid(n)
    z, y, x = np.meshgrid(*(np.arange(n) * 2 * np.pi / n,) * 3, indexing="ij")
    ax, ay, az = np.sin(z), np.cos(x) * np.sin(y), np.cos(2 * y)
    d = np.cos(x) + np.sin(3 * z) * np.cos(y)
    grad = (-np.sin(x), -np.sin(3 * z) * np.sin(y), 3 * np.cos(3 * z) * np.cos(y))
    want = g.r2c(-(ax * grad[0] + ay * grad[1] + az * grad[2]))
    e = g.advect3(g.r2c(ax), g.r2c(ay), g.r2c(az), g.r2c(d))
    assert rel(e, want) < 1e-13


def gravity_wave(g, k, amp_v, amp_t):
    """Single mode (kx,ky,kz) with velocity along P(z_hat) and temperature."""
    kx, ky, kz = k
    kv = np.array(k, float)
    pz = np.array([0, 0, 1.0]) - kv * kz / kv.dot(kv)
    e = pz / np.linalg.norm(pz)
    v = [g.czeros() for _ in range(3)]
    th = g.czeros()
    N3 = g.nx * g.ny * g.nz
    for c in range(3):
        v[c][kx, ky % g.ny, kz % g.nz] = amp_v * e[c] * N3
    th[kx, ky % g.ny, kz % g.nz] = amp_t * N3
    return v, th, e, np.linalg.norm(pz)


@pytest.mark.parametrize("ord", [2, 3])
def test_single_mode_gravity_wave(ord):
    n = 16
    g = O.Grid(n)
    k = (2, 1, 3)
    k2 = float(np.dot(k, k))
    nu, kappa, xmom, xtemp, dt = 0.02, 0.03, 1.7, 1.7, 0.01
    a0, t0 = 0.3 + 0.1j, -0.2 + 0.4j
    v, th, e, s = gravity_wave(g, k, a0, t0)
    mean = 0.7 * n ** 3  # the k = 0 temperature is carried unchanged
    th[0, 0, 0] = mean
    v_before = [x.copy() for x in v]
    g.bouss_step(v, th, dt, nu, kappa, xmom, xtemp, ord=ord)
    # d/dt [a, th] = M [a, th]: buoyancy enters the momentum as -P(xmom th z_hat),
    # the heat current as +xtemp v_z (bouss_rkstep2.f90:4-26)
    M = np.array([[-nu * k2, -xmom * s], [xtemp * s, -kappa * k2]])
    A = dt * M
    if ord == 2:
        R = np.eye(2) + A + A @ A / 2
    else:
        R = np.eye(2) + A + A @ A / 2 + A @ A @ A / 6
    want = R @ np.array([a0, t0])
    if ord == 3:
        # bouss_rkstep2.f90:73-145: u += 2 N(L(C) + N(C)) dt^3/24 with N(single mode)
        # = -P(xmom th0 z_hat): only the buoyancy of the step-start temperature
        want[0] += 2 * (-xmom * t0 * s) * dt ** 3 / 24
    N3 = n ** 3
    idx = (k[0], k[1] % n, k[2] % n)
    got_a = sum(v[c][idx] * e[c] for c in range(3)) / N3
    assert abs(got_a - want[0]) < 1e-14
    assert abs(th[idx] / N3 - want[1]) < 1e-14
    # the velocity stays along P(z_hat) and every other mode stays zero, except
    # that the ord = 3 correction is added without a dealias mask
    # (bouss_rkstep2.f90:134-143): the buoyancy of the mean temperature lands
    # in the k = 0 mode of vz (nonlhd3 returns -a_z on the kz = 0 plane)
    if ord == 3:
        assert abs(v[2][0, 0, 0] - 2 * (-xmom * mean) * dt ** 3 / 24) < 1e-12
        v[2][0, 0, 0] = 0
    for c in range(3):
        rest = v[c].copy()
        rest[idx] = 0
        assert np.max(np.abs(rest)) < 1e-10
        assert abs(v[c][idx] / N3 - got_a * e[c]) < 1e-14
    assert th[0, 0, 0] == mean
    assert np.all(v_before[0][0, 0, 0] == 0)


def test_bouss_reduces_to_hd_without_temperature():
    """th = 0, xmom = xtemp = 0: the velocity follows og_hd_step exactly."""
    g = O.Grid(16)
    v, _ = g.initialv_k4exp(3.0, 1.0, 5)
    v2 = [x.copy() for x in v]
    th = g.czeros()
    for _ in range(3):
        g.bouss_step(v, th, 5e-3, 1e-2, 2e-2, 0.0, 0.0, ord=2)
        g.hd_step(v2, 5e-3, 1e-2, ord=2)
    for c in range(3):
        assert np.array_equal(v[c], v2[c])
    assert np.all(th == 0)


def solenoidal_basis(k):
    """Two orthonormal real vectors perpendicular to k."""
    kv = np.asarray(k, float)
    a = np.cross(kv, [0.3, 0.5, 0.8])
    a /= np.linalg.norm(a)
    b = np.cross(kv, a)
    return a, b / np.linalg.norm(b)


def coriolis_operator(k, omegaz):
    """The projected Coriolis force -P_k(2 Omega z_hat x v) on the plane
    perpendicular to k, in the basis of solenoidal_basis."""
    a, b = solenoidal_basis(k)
    kv = np.asarray(k, float)
    P = np.eye(3) - np.outer(kv, kv) / kv.dot(kv)
    Om = np.array([0, 0, 2.0 * omegaz])
    B = np.stack([a, b], axis=1)
    return B.T @ (-P @ np.stack([np.cross(Om, B[:, q]) for q in range(2)], axis=1))


@pytest.mark.parametrize("k", [(2, 1, 3), (1, -2, 2), (3, 0, 1)])
def test_coriolis_operator_gives_inertial_waves(k):
    """Textbook dispersion relation of inertial waves, omega = +-2 Omega kz/|k|,
    for the operator the next tests use."""
    omegaz = 0.8
    ev = np.linalg.eigvals(coriolis_operator(k, omegaz))
    want = 2 * omegaz * k[2] / np.linalg.norm(k)
    assert np.allclose(sorted(ev.imag), sorted([-want, want]), atol=1e-14)
    assert np.allclose(ev.real, 0, atol=1e-14)


@pytest.mark.parametrize("ord", [1, 2, 3])
@pytest.mark.parametrize("k", [(2, 1, 3), (1, -2, 2)])
def test_single_mode_inertial_wave(ord, k):
    """One Fourier mode is an exact solution in a rotating frame (the
    nonlinear term is a gradient); the RK step must apply the Taylor
    polynomial of nu*lap + Coriolis to its two solenoidal amplitudes."""
    n = 16
    g = O.Grid(n)
    nu, omegaz, dt = 0.02, 0.9, 0.05
    a, b = solenoidal_basis(k)
    amp = np.array([0.3 + 0.1j, -0.2 + 0.25j])
    idx = (k[0], k[1] % n, k[2] % n)
    N3 = n ** 3
    v = [g.czeros() for _ in range(3)]
    for c in range(3):
        v[c][idx] = (amp[0] * a[c] + amp[1] * b[c]) * N3
    g.roth_step(v, dt, nu, omegaz, ord=ord)
    A = dt * (-nu * np.dot(k, k) * np.eye(2) + coriolis_operator(k, omegaz))
    R = np.eye(2)
    term = np.eye(2)
    for p in range(1, ord + 1):
        term = term @ A / p
        R = R + term
    want = R @ amp
    got = np.array([sum(v[c][idx] * a[c] for c in range(3)), sum(v[c][idx] * b[c] for c in range(3))]) / N3
    assert np.max(np.abs(got - want)) < 1e-14
    for c in range(3):
        rest = v[c].copy()
        rest[idx] = 0
        assert np.max(np.abs(rest)) < 1e-10


def test_rotation_off_is_hd_and_bouss():
    g = O.Grid(16)
    v, _ = g.initialv_k4exp(3.0, 1.0, 5)
    th = g.r2c(np.cos(2 * np.pi * np.arange(16) / 16)[None, None, :] * np.ones((16, 16, 16)))
    v1, v2, v3, v4 = ([x.copy() for x in v] for _ in range(4))
    t3, t4 = th.copy(), th.copy()
    for _ in range(2):
        g.roth_step(v1, 5e-3, 1e-2, 0.0)
        g.hd_step(v2, 5e-3, 1e-2)
        g.rotbouss_step(v3, t3, 5e-3, 1e-2, 2e-2, 0.5, -0.5, 0.0)
        g.bouss_step(v4, t4, 5e-3, 1e-2, 2e-2, 0.5, -0.5)
    for c in range(3):
        assert np.array_equal(v1[c], v2[c])
        assert np.array_equal(v3[c], v4[c])
    assert np.array_equal(t3, t4)


@pytest.mark.parametrize("ord", [2, 3])
def test_single_mode_inertia_gravity_wave(ord):
    """Rotation and stratification together: the 3x3 linear operator of one
    mode (two solenoidal amplitudes and th) has the textbook inertia-gravity
    frequencies omega^2 = (N^2 k_perp^2 + f^2 kz^2)/k^2 with f = 2 Omega and
    N^2 = xmom xtemp, and the RK step applies its Taylor polynomial."""
    n = 16
    g = O.Grid(n)
    k = (2, 1, 3)
    k2 = float(np.dot(k, k))
    nu, kappa, omegaz, bv, dt = 0.02, 0.03, 0.7, 1.3, 0.02
    xmom, xtemp = bv, bv
    a, b = solenoidal_basis(k)
    kv = np.array(k, float)
    pz = np.array([0, 0, 1.0]) - kv * k[2] / k2
    M = np.zeros((3, 3))
    M[:2, :2] = coriolis_operator(k, omegaz)
    # buoyancy -xmom th P(z_hat) (the products enter the projection of
    # nonlhd3 with a minus sign), heat current +xtemp v_z
    M[0, 2], M[1, 2] = -xmom * a.dot(pz), -xmom * b.dot(pz)
    M[2, 0], M[2, 1] = xtemp * a[2], xtemp * b[2]
    freq = np.sort(np.abs(np.linalg.eigvals(M).imag))
    kp2 = k[0] ** 2 + k[1] ** 2
    w = np.sqrt((bv ** 2 * kp2 + (2 * omegaz) ** 2 * k[2] ** 2) / k2)
    assert np.allclose(freq, [0, w, w], atol=1e-13)
    D = np.diag([-nu * k2, -nu * k2, -kappa * k2])
    A = dt * (D + M)
    R = np.eye(3)
    term = np.eye(3)
    for p in range(1, ord + 1):
        term = term @ A / p
        R = R + term
    amp = np.array([0.3 + 0.1j, -0.2 + 0.25j, 0.15 - 0.3j])
    idx = (k[0], k[1], k[2])
    N3 = n ** 3
    v = [g.czeros() for _ in range(3)]
    th = g.czeros()
    for c in range(3):
        v[c][idx] = (amp[0] * a[c] + amp[1] * b[c]) * N3
    th[idx] = amp[2] * N3
    g.rotbouss_step(v, th, dt, nu, kappa, xmom, xtemp, omegaz, ord=ord)
    want = R @ amp
    got = np.array([sum(v[c][idx] * a[c] for c in range(3)) / N3, sum(v[c][idx] * b[c] for c in range(3)) / N3,
                    th[idx] / N3])
    assert np.max(np.abs(got - want)) < 1e-14
