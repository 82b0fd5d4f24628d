"""GPU parity of the transfer spectra, 1D spectra, cross-helicity spectrum and
mhdcheck with the oracle (pseudospec3D_hd.f90:1418-1949,
pseudospec3D_mhd.f90:339-712), on cubic, non-cubic and anisotropic boxes.
Shell sums: relative L2 error 1e-12; scalar totals 1e-13 relative."""
import numpy as np
import pytest

from test_gpu_parity import grids, rand_spec, rel, to_gpu

pytestmark = pytest.mark.gpu

CASES = [((32, 32, 32), False, (1, 1, 1)), ((32, 32, 16), True, (1.0, 1.0, 0.5)), ((24, 16, 12), False, (1, 1, 1)),
         ((64, 64, 64), False, (1, 1, 1))]


@pytest.mark.parametrize("shape,arb,L", CASES)
def test_transfer_spectra(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    v = [rand_spec(og, s) for s in (41, 42, 43)]
    w = [rand_spec(og, s) for s in (44, 45, 46)]
    V, W = [to_gpu(x) for x in v], [to_gpu(x) for x in w]
    for kin in (0, 1):
        assert rel(g.entrans(*V, *W, kin=kin), og.entrans(*v, *w, kin=kin)) < 1e-12
    assert rel(g.heltrans(*V, *W), og.heltrans(*v, *w)) < 1e-12
    for shift in (0.0, 1.0):
        assert rel(g.crosspecc(*V, *W, shift=shift), og.crosspecc(*v, *w, shift=shift)) < 1e-12


@pytest.mark.parametrize("shape,arb,L", CASES)
def test_spectr1d(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    a = rand_spec(og, 47)
    A = to_gpu(a)
    for d in (1, 2, 3):
        got, want = g.spectr1d(A, d), og.spectr1d(a, d)
        assert rel(got, want) < 1e-12
        assert np.nonzero(np.asarray(got))[0].tolist() == np.nonzero(want)[0].tolist()


@pytest.mark.parametrize("shape,arb,L", CASES)
def test_mhdcheck(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    v = [rand_spec(og, s) for s in (51, 52, 53)]
    a = [rand_spec(og, s) for s in (54, 55, 56)]
    V, A = [to_gpu(x) for x in v], [to_gpu(x) for x in a]
    for hel, crs, chk in ((1, 1, 1), (0, 0, 0), (1, 0, 1)):
        got = g.mhdcheck(*V, *A, hel=hel, crs=crs, chk=chk)
        want = og.mhdcheck(*v, *a, hel=hel, crs=crs, chk=chk)
        scale = max(abs(x) for x in want)
        for x, y in zip(got, want):
            assert x == pytest.approx(y, rel=1e-12, abs=1e-13 * scale)


@pytest.mark.parametrize("shape,arb,L", CASES[:3])
def test_hdcheck(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    v = [rand_spec(og, s) for s in (61, 62, 63)]
    f = [rand_spec(og, s) for s in (64, 65, 66)]
    V, F = [to_gpu(x) for x in v], [to_gpu(x) for x in f]
    for hel, chk in ((1, 1), (0, 0)):
        got = g.hdcheck(*V, *F, hel=hel, chk=chk)
        want = og.hdcheck(*v, *f, hel=hel, chk=chk)
        scale = max(abs(x) for x in want)
        for x, y in zip(got, want):
            assert x == pytest.approx(y, rel=1e-12, abs=1e-13 * scale)


@pytest.mark.parametrize("shape,arb,L", CASES)
def test_scalar_diagnostics(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    a, b = rand_spec(og, 71), rand_spec(og, 72)
    A, B = to_gpu(a), to_gpu(b)
    for kin in (0, 1):
        assert g.variance(A, kin) == pytest.approx(og.variance(a, kin), rel=1e-13)
    assert g.product(A, B) == pytest.approx(og.product(a, b), rel=1e-12, abs=1e-14 * og.variance(a, 1))
    assert rel(g.pscheck(A, B), og.pscheck(a, b)) < 1e-13
    for shift in (0.0, 1.0):
        assert rel(g.spectrscc(A, shift), og.spectrscc(a, shift)) < 1e-13
    assert rel(g.sctrans(A, B), og.sctrans(a, b)) < 1e-12
    for d in (1, 3):
        got, want = np.asarray(g.difuc(A, B, d)), og.difuc(a, b, d)
        for q in range(3):
            assert rel(got[q], want[q]) < 1e-12


@pytest.mark.parametrize("shape,arb,L", CASES)
def test_axisymmetric_spectra(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    v = [rand_spec(og, s) for s in (73, 74, 75)]
    w = [rand_spec(og, s) for s in (76, 77, 78)]
    V, W = [to_gpu(x) for x in v], [to_gpu(x) for x in w]
    assert g.axis_bins() == (og.nbpara, og.nbperp)
    for name in ("specpara", "specperp"):
        for kin, hel in ((1, 1), (0, 1), (2, 1), (1, 0)):
            E, H = getattr(g, name)(*V, kin=kin, hel=hel)
            Er, Hr = getattr(og, name)(*v, kin=kin, hel=hel)
            assert (E is None) == (Er is None) and (H is None) == (Hr is None)
            if E is not None:
                for q in range(3):
                    assert rel(E[q], Er[q]) < 1e-12, (name, kin, q)
            if H is not None:
                for q in range(3):
                    assert rel(H[q], Hr[q]) < 1e-11, (name, kin, q)
    for name in ("entpara", "entperp"):
        for kin in (0, 1, 2):
            assert rel(getattr(g, name)(*V, *W, kin=kin), getattr(og, name)(*v, *w, kin=kin)) < 1e-12
    for name in ("heltpara", "heltperp"):
        assert rel(getattr(g, name)(*V, *W), getattr(og, name)(*v, *w)) < 1e-12
    for kin, hel in ((1, 1), (0, 0)):
        E, H = g.spec2d(*V, kin=kin, hel=hel)
        Er, Hr = og.spec2d(*v, kin=kin, hel=hel)
        assert rel(np.asarray(E), Er) < 1e-12
        if hel:
            assert rel(np.asarray(H), Hr) < 1e-11


@pytest.mark.parametrize("shape,arb,L", CASES)
def test_boussinesq_diagnostics(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    f = [rand_spec(og, s) for s in (81, 82, 83, 84)]
    F = [to_gpu(x) for x in f]
    fo, bv = 1.7, 0.9
    for it in range(11):
        assert rel(g.havgcomp(*F, fo, bv, it), og.havgcomp(*f, fo, bv, it)) < 1e-12, it
    for got, want in zip(g.tbouss(*F, fo, bv), og.tbouss(*f, fo, bv)):
        assert rel(got, want) < 1e-12
    assert rel(g.spectpv(*F), og.spectpv(*f)) < 1e-12


@pytest.mark.parametrize("n", [64, 512])
def test_spectrumc_bins_match_exact_lattice_counts(gpu, n):
    """A field with |a|^2 = N^6 on every mode: each shell of the GPU spectrum is
    the Hermitian-weighted mode count, equal to the exact rational count (the
    binning of pseudospec3D_hd.f90:1190-1415 at the 512^3 workload size)."""
    import torch
    from ghost_amd.pseudospec import Grid
    from test_oracle_diag import exact_shell_weights
    g = Grid(n, n, n)
    a = torch.full(g.cshape, float(n) ** 3, dtype=torch.complex128, device="cuda")
    z = torch.zeros(g.cshape, dtype=torch.complex128, device="cuda")
    Ek, _ = g.spectrumc(a, z, z, kin=1, hel=0)
    del a, z
    want = exact_shell_weights(n, n, n, g.nbins)
    assert np.array_equal(np.array(Ek), want.astype(np.float64))
