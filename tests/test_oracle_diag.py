"""The oracle's transfer spectra, 1D spectra, cross-helicity spectrum and
mhdcheck (pseudospec3D_hd.f90:1418-1949, pseudospec3D_mhd.f90:339-712),
pinned by identities with the already-pinned spectrumc/energy/helicity/cross
and by single-mode bin placement."""
import numpy as np
import pytest

from oracle import oracle as O


def test_entrans_of_a_field_with_itself_is_its_spectrum():
    og = O.Grid(32)
    v, _ = og.initialv_k4exp(3.0, 1.0, 5)
    Ek, _ = og.spectrumc(*v, kin=1)
    assert np.allclose(og.entrans(*v, *v, kin=1), Ek, rtol=1e-13, atol=0)
    # kk2 |v|^2 is the enstrophy spectrum for a solenoidal field
    Ew, _ = og.spectrumc(*v, kin=0)
    assert np.allclose(og.entrans(*v, *v, kin=0), Ew, rtol=1e-11, atol=1e-14 * Ew.max())


def test_heltrans_of_a_field_with_itself_is_its_helicity_spectrum():
    og = O.Grid(32)
    v, _ = og.initialv_k4exp(3.0, 1.0, 6)
    _, Hk = og.spectrumc(*v, kin=1, hel=1)
    assert np.allclose(og.heltrans(*v, *v), Hk, rtol=1e-12, atol=1e-14 * np.abs(Hk).max())


def test_crosspecc_of_curl_A_with_A_is_the_magnetic_spectrum():
    og = O.Grid(32)
    a, _ = og.initialv_k4exp(3.0, 1.0, 7)
    b = og.curl(*a)
    Eb, _ = og.spectrumc(*b, kin=1)
    Cb = og.crosspecc(*b, *a)
    assert np.allclose(Cb, Eb, rtol=1e-12, atol=1e-14 * Eb.max())
    # the shift moves the shell boundaries: one shell up for shift = 1
    C1 = og.crosspecc(*b, *a, shift=1.0)
    assert np.allclose(C1[1:], Cb[:-1], rtol=1e-12, atol=1e-14 * Eb.max())


def test_spectr1d_single_mode_bins_and_parseval():
    n = 32
    og = O.Grid(n)
    a = og.czeros()
    a[3, 5, 7] = 2.0 + 1.0j     # kx = 3, ky = 5, kz = 7
    a[0, 4, n - 6] = 1.5         # kx = 0, ky = 4, kz = -6
    for d, (b1, b2) in zip((1, 2, 3), ((3, 0), (5, 4), (7, 6))):
        e = og.spectr1d(a, d)
        nz = np.nonzero(e)[0]
        assert sorted(nz.tolist()) == sorted({b1, b2})  # bin index = |k_dir| (shell |k_dir| + 1)
    v, _ = og.initialv_k4exp(3.0, 1.0, 8)
    z = og.czeros()
    for d in (1, 2, 3):
        assert abs(np.sum(og.spectr1d(v[0], d)) - og.energy(v[0], z, z, 1)) < 1e-13 * og.energy(v[0], z, z, 1)


def test_mhdcheck_is_the_composition_of_the_diagnostics():
    og = O.Grid(32)
    v, _ = og.initialv_k4exp(3.0, 1.0, 9)
    a, _ = og.initialv_k4exp(2.5, 0.7, 10)
    out = og.mhdcheck(*v, *a, hel=1, crs=1, chk=1)
    b = og.curl(*a)
    want = [og.energy(*v, 1) + og.energy(*b, 1), og.energy(*v, 0), og.energy(*b, 0), og.energy(*v, 1),
            og.energy(*b, 1), og.helicity(*v), og.helicity(*a), og.cross(*v, *b, 1), og.energy(*a, 1),
            og.divergence(*v), og.divergence(*a)]
    assert np.allclose(out, want, rtol=1e-14, atol=1e-300)
    assert out[9] < 1e-20 * out[3]  # solenoidal fields


# ---- scalar diagnostics (pseudospec3D_phd.f90) ------------------------------

def test_variance_product_parseval():
    og = O.Grid(16, 16, 32, L=(1.0, 1.0, 2.0), arbsize=True)
    rng = np.random.default_rng(3)
    # no kx = nx/2 content: GHOST's Hermitian weights count that plane twice
    a, b = og.r2c(rng.standard_normal(og.rshape)), og.r2c(rng.standard_normal(og.rshape))
    a[-1], b[-1] = 0, 0
    n3 = og.nx * og.ny * og.nz
    ra, rb = og.c2r(a) / n3, og.c2r(b) / n3
    assert og.variance(a, 1) == pytest.approx(np.mean(ra * ra), rel=1e-13)
    assert og.product(a, b) == pytest.approx(np.mean(ra * rb), rel=1e-12)
    # <|grad th|^2> from the derivatives (derivk3 is pinned by closed forms)
    gx, gy, gz = (og.derivk3(a, d) for d in (1, 2, 3))
    assert og.variance(a, 0) == pytest.approx(og.energy(gx, gy, gz, 1), rel=1e-13)
    assert np.allclose(og.pscheck(a, b), [og.variance(a, 1), og.variance(a, 0), og.product(a, b)], rtol=0)


def test_scalar_spectra():
    og = O.Grid(32)
    v, _ = og.initialv_k4exp(3.0, 1.0, 7)
    a = v[0]
    z = og.czeros()
    S = og.spectrscc(a)
    assert np.sum(S) == pytest.approx(og.variance(a, 1), rel=1e-13)  # dealiased: every mode in a shell
    Ek, _ = og.spectrumc(a, z, z, kin=1)
    assert np.allclose(S, Ek, rtol=1e-13, atol=0)
    assert np.allclose(og.spectrscc(a, shift=1.0)[1:], S[:-1], rtol=1e-13, atol=0)
    assert np.allclose(og.sctrans(a, a), S, rtol=1e-13, atol=0)
    m = og.czeros()
    m[3, 4, 0] = 1.0  # |k| = 5
    assert np.nonzero(og.spectrscc(m))[0].tolist() == [4]


def test_difuc_closed_forms():
    n = 16
    og = O.Grid(n)
    zz, yy, xx = np.meshgrid(*(2 * np.pi * np.arange(n) / n,) * 3, indexing="ij")
    v = og.r2c(np.cos(xx) + 0.5)
    th = og.r2c(np.cos(xx) + np.sin(zz))
    x = 2 * np.pi * np.arange(n) / n
    px = og.difuc(v, th, 1)
    assert np.allclose(px[0], np.cos(x) + 0.5, atol=1e-14)
    assert np.allclose(px[1], np.cos(x), atol=1e-14)
    assert np.allclose(px[2], (np.cos(x) + 0.5) * np.cos(x), atol=1e-14)
    pz = og.difuc(v, th, 3)
    assert np.allclose(pz[0], 0.5, atol=1e-14)
    assert np.allclose(pz[1], np.sin(x), atol=1e-14)
    assert np.allclose(pz[2], 0.5 + 0.5 * np.sin(x), atol=1e-14)


# ---- axisymmetric spectra (pseudospec3D_rot.f90) ----------------------------

@pytest.mark.parametrize("shape,L,arb", [((32, 32, 32), (1, 1, 1), False), ((32, 32, 16), (1.0, 1.0, 0.5), True)])
def test_axisymmetric_sums(shape, L, arb):
    og = O.Grid(*shape, L=L, arbsize=arb)
    v, _ = og.initialv_k4exp(3.0, 1.0, 9)
    z = og.czeros()
    eng = og.energy(*v, 1)
    Ep, Hp = og.specpara(*v, kin=1, hel=1)
    Ec, Hc = og.specperp(*v, kin=1, hel=1)
    for E in (Ep, Ec):
        assert np.sum(E[0]) == pytest.approx(eng, rel=1e-13)  # dealiased: no mode is dropped
    assert np.sum(Ep[1]) == pytest.approx(og.energy(v[0], v[1], z, 1), rel=1e-13)
    assert np.sum(Ep[2]) == pytest.approx(og.energy(z, z, v[2], 1), rel=1e-13)
    # cylinders: the xy and z columns hold the kz = 0 plane only
    plane = [x.copy() for x in v]
    for x in plane:
        x[:, :, 1:] = 0
    assert np.sum(Ec[1]) == pytest.approx(og.energy(plane[0], plane[1], z, 1), rel=1e-13)
    assert np.sum(Ec[2]) == pytest.approx(og.energy(z, z, plane[2], 1), rel=1e-13)
    hel = og.helicity(*v)
    for H in (Hp, Hc):
        assert np.sum(H[0]) == pytest.approx(hel, rel=1e-12, abs=1e-14 * eng)
    # the 2D spectrum's marginals are the two reduced spectra
    E2, H2 = og.spec2d(*v, kin=1, hel=1)
    assert np.allclose(E2.sum(axis=1), Ep[0], rtol=1e-12, atol=1e-15 * eng)
    assert np.allclose(E2.sum(axis=0), Ec[0], rtol=1e-12, atol=1e-15 * eng)
    assert np.allclose(H2.sum(axis=1), Hp[0], rtol=1e-11, atol=1e-14 * eng)
    # magnetic spectra (kin 0) are the spectra of the curl
    w = og.curl(*v)
    assert np.allclose(og.specpara(*v, kin=0)[0], og.specpara(*w, kin=1)[0], rtol=0, atol=0)


def test_axisymmetric_transfers():
    og = O.Grid(32)
    v, _ = og.initialv_k4exp(3.0, 1.0, 11)
    eng = og.energy(*v, 1)
    Ep, Hp = og.specpara(*v, kin=1, hel=1)
    Ec, Hc = og.specperp(*v, kin=1, hel=1)
    assert np.allclose(og.entpara(*v, *v, kin=1), Ep[0], rtol=1e-13, atol=0)
    assert np.allclose(og.entperp(*v, *v, kin=1), Ec[0], rtol=1e-13, atol=0)
    # kk2 |v|^2 = |k x v|^2 for a solenoidal field: the curl's spectra
    assert np.allclose(og.entpara(*v, *v, kin=0), og.specpara(*v, kin=0)[0][0], rtol=1e-11,
                       atol=1e-14 * og.energy(*v, 0))
    assert np.allclose(og.heltpara(*v, *v), Hp[0], rtol=1e-11, atol=1e-14 * eng)
    assert np.allclose(og.heltperp(*v, *v), Hc[0], rtol=1e-11, atol=1e-14 * eng)


def test_axisymmetric_single_mode_bins():
    og = O.Grid(32, 32, 16, L=(1.0, 1.0, 0.5), arbsize=True)
    m = og.czeros()
    m[3, 4, 16 - 2] = 1.0  # kx = 3, ky = 4, kz index -2 (kz = -4 in a box of length 0.5)
    z = og.czeros()
    Ep, _ = og.specpara(m, z, z, kin=1)
    Ec, _ = og.specperp(m, z, z, kin=1)
    assert np.nonzero(Ep[0])[0].tolist() == [2]  # int(|kz| Lz + 1) = 3
    assert np.nonzero(Ec[0])[0].tolist() == [5]  # int(sqrt(9 + 16)/Dkk + 1) = 6
    assert not np.any(Ec[1]) and not np.any(Ec[2])  # kz != 0
    assert og.nbpara == 9 and og.nbperp == 17


# ---- Boussinesq diagnostics (pseudospec3D_bouss.f90) ------------------------

def bouss_fields(og, A=0.7, B=-0.4, C=0.3, W0=0.25, F=0.5, D=0.6, E=-0.35):
    """u = A sin z + F cos y, v = B cos z, w = W0 + C cos x, th = D sin z + E sin y,
    whose derivatives and horizontal averages are known in closed form."""
    n = og.nx
    zz, yy, xx = np.meshgrid(*(2 * np.pi * np.arange(n) / n,) * 3, indexing="ij")
    real = (A * np.sin(zz) + F * np.cos(yy), B * np.cos(zz) + 0 * xx, W0 + C * np.cos(xx) + 0 * zz,
            D * np.sin(zz) + E * np.sin(yy) + 0 * xx)
    return [og.r2c(r) for r in real], (xx, yy, zz), dict(A=A, B=B, C=C, W0=W0, F=F, D=D, E=E)


def test_havgcomp_closed_forms():
    n = 16
    og = O.Grid(n)
    (u, v, w, s), (xx, yy, zz), p = bouss_fields(og)
    A, B, C, W0, F, D, E = (p[k] for k in ("A", "B", "C", "W0", "F", "D", "E"))
    fo, bv = 1.3, 0.8
    z = 2 * np.pi * np.arange(n) / n
    want = {
        0: A ** 2 * np.cos(z) ** 2 + B ** 2 * np.sin(z) ** 2,
        1: D ** 2 * np.cos(z) ** 2,
        2: W0 * D * np.cos(z),
        3: A ** 2 * np.sin(z) ** 2 + F ** 2 / 2 + B ** 2 * np.cos(z) ** 2,
        4: np.full(n, W0 ** 2 + C ** 2 / 2),
        5: np.full(n, A * B),
        6: np.full(n, F * E / 2),
        8: np.full(n, A * B),
        10: W0 * D * np.sin(z),
    }
    want[9] = want[0]  # GHOST's first branch serves itype 0 and 9
    # potential vorticity from the analytic derivatives
    om = (B * np.sin(zz), A * np.cos(zz) + C * np.sin(xx), F * np.sin(yy))
    gth = (0 * xx, E * np.cos(yy), D * np.cos(zz))
    pv = fo * D * np.cos(zz) - bv * om[2] + sum(o * g_ for o, g_ in zip(om, gth)) - fo * bv
    want[7] = np.mean(pv ** 2, axis=(1, 2))
    for it, prof in want.items():
        got = og.havgcomp(u, v, w, s, fo, bv, it)
        assert np.allclose(got, prof, rtol=1e-12, atol=1e-13), it
    avg, mx, mn = og.tbouss(u, v, w, s, fo, bv)
    assert np.allclose(avg[[4, 5]], [W0 ** 2 + C ** 2 / 2, A * B], rtol=1e-12)
    assert mx[0] == pytest.approx(max(A, B, key=abs) ** 2, rel=1e-12) and mn[10] == pytest.approx(-abs(W0 * D), rel=1e-12)


def test_spectpv_closed_form():
    og = O.Grid(16)
    (u, v, w, s), _, p = bouss_fields(og)
    A, C, F, D, E = (p[k] for k in ("A", "C", "F", "D", "E"))
    S = og.spectpv(u, v, w, s)
    # omega . grad th = E C sin x cos y (kz = 0) + (E A cos y + F D sin y) cos z (|kz| = 1)
    assert S[0] == pytest.approx(E ** 2 * C ** 2 / 4, rel=1e-12)
    assert S[1] == pytest.approx((E ** 2 * A ** 2 + F ** 2 * D ** 2) / 4, rel=1e-12)
    assert np.all(np.abs(S[2:]) < 1e-25)


def exact_shell_weights(nx, ny, nz, nb):
    """Hermitian-weighted mode count per shell of spectrumc on the 2 pi cube
    (Dkk = 1), computed in exact integer and rational arithmetic.

    A mode with integer |k|^2 = n lies in shell s = int(sqrt(n) + 0.501) with
    GHOST's offset the single-precision literal .501 (pseudospec3D_hd.f90:1190-
    1415).  With r = isqrt(n), s is r + 1 exactly when n >= (r + 1 - off)^2 as
    rationals, and r otherwise, so no floating-point rounding enters."""
    from fractions import Fraction
    from math import isqrt
    off = Fraction(float(np.float32(0.501)))
    ky = np.fft.fftfreq(ny, 1.0 / ny).astype(np.int64)
    kz = np.fft.fftfreq(nz, 1.0 / nz).astype(np.int64)
    yz = np.bincount((ky[:, None] ** 2 + kz[None, :] ** 2).ravel())
    cnt = np.zeros((nx // 2) ** 2 + yz.size, dtype=np.int64)
    for kx in range(nx // 2 + 1):
        cnt[kx * kx:kx * kx + yz.size] += (1 if kx == 0 else 2) * yz
    out = np.zeros(nb, dtype=np.int64)
    for n in np.nonzero(cnt)[0].tolist():
        r = isqrt(n)
        s = r + 1 if Fraction(n) >= (r + 1 - off) ** 2 else r
        if 1 <= s <= nb:
            out[s - 1] += cnt[n]
    return out


def test_exact_shell_weights_small_cases():
    # shell 1 holds |k| = 1 and |k| = sqrt 2 (1.414 + .501 < 2); shell 2
    # starts at |k| = sqrt 3.  kx = 0: 4 + 4 modes, weight 1; kx = 1: the
    # axis mode and (1, +-1, 0), (1, 0, +-1), weight 2.
    w = exact_shell_weights(8, 8, 8, 5)
    assert w[0] == 1 * (4 + 4) + 2 * (1 + 4)
    # shell 2: |k|^2 in {3, 4, 5, 6} (sqrt 6 + .501 = 2.95); shell 3 from 7
    # (2.646 + .501 = 3.147).  Count |k|^2 = 3..6 by brute force.
    k = np.fft.fftfreq(8, 1 / 8).astype(int)
    n2 = k[:5, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    wt = np.where(np.arange(5) == 0, 1, 2)[:, None, None]
    assert w[1] == int(np.sum(wt * ((n2 >= 3) & (n2 <= 6))))
    # every mode of the half space with 0 < |k| < 4.499 lands in shells 1..4
    assert w[:4].sum() == int(np.sum(wt * ((n2 > 0) & (np.sqrt(n2) + 0.501 < 5))))


@pytest.mark.parametrize("n", [64, 512])
def test_spectrumc_bins_match_exact_lattice_counts(n):
    """spectrumc of a field with |a|^2 = N^6 on every mode returns, per shell,
    the Hermitian-weighted number of modes in it (exact in double), which must
    equal the exact rational count, at 64^3 and at the 512^3 workload size."""
    og = O.Grid(n)
    a = np.full(og.cshape, float(n) ** 3, dtype=np.complex128)
    z = np.zeros(og.cshape, dtype=np.complex128)
    Ek, _ = og.spectrumc(a, z, z, kin=1)
    del a
    want = exact_shell_weights(n, n, n, og.nbins)
    assert np.array_equal(Ek, want.astype(np.float64))
