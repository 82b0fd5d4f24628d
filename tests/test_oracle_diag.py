"""The oracle's transfer spectra, 1D spectra, cross-helicity spectrum and
mhdcheck (pseudospec3D_hd.f90:1418-1949, pseudospec3D_mhd.f90:339-712),
pinned by identities with the already-pinned spectrumc/energy/helicity/cross
and by single-mode bin placement."""
import numpy as np

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
