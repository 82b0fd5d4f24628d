"""The oracle's fftp transforms, pinned without the oracle's own code:
numpy's pocketfft (a library routine), brute-force DFT sums, Parseval,
linearity and the unnormalised round trip."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O

SHAPES = [(8, 8, 8), (6, 4, 10), (12, 12, 12), (16, 8, 32), (10, 6, 4), (20, 24, 6), (32, 32, 32), (48, 48, 48)]


def ghost_layout(c):
    """numpy rfftn output c[kz][ky][kx] -> GHOST's transposed c[kx][ky][kz]."""
    return np.ascontiguousarray(np.transpose(c, (2, 1, 0)))


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_matches_pocketfft(shape):
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz)
    r = np.random.default_rng(3).standard_normal((nz, ny, nx))
    ref = ghost_layout(np.fft.rfftn(r))  # sign -1, unnormalised
    assert rel(g.r2c(r), ref) < 2e-15


@pytest.mark.parametrize("shape", SHAPES)
def test_c2r_matches_pocketfft_and_round_trip(shape):
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz)
    r = np.random.default_rng(4).standard_normal((nz, ny, nx))
    c = g.r2c(r)
    out = g.c2r(c)
    assert rel(out, r * nx * ny * nz) < 2e-15
    # a spectrum that is not Hermitian-consistent: GHOST's c2r (FFTW) does c2c
    # along z and y and ignores Im of the kx=0 and Nyquist points of each x row,
    # as pocketfft's irfftn does
    rng = np.random.default_rng(5)
    cc = rng.standard_normal(g.cshape) + 1j * rng.standard_normal(g.cshape)
    ref = np.fft.irfftn(np.transpose(cc, (2, 1, 0)), s=(nz, ny, nx), axes=(0, 1, 2)) * nx * ny * nz
    assert rel(g.c2r(cc), ref) < 2e-15


def test_c2r_preserves_input():
    g = O.Grid(16)
    c = g.r2c(np.random.default_rng(6).standard_normal(g.rshape))
    keep = c.copy()
    g.c2r(c)
    assert np.array_equal(c, keep)


@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 4, 2), (2, 6, 6)])
def test_r2c_brute_force_dft(shape):
    """X(kx,ky,kz) = sum_xyz r exp(-2 pi i (kx x/nx + ky y/ny + kz z/nz)), direct sum."""
    nx, ny, nz = shape
    g = O.Grid(nx, ny, nz)
    r = np.random.default_rng(7).standard_normal((nz, ny, nx))
    c = g.r2c(r)
    for kx, ky, kz in itertools.product(range(nx // 2 + 1), range(ny), range(nz)):
        s = 0j
        for x, y, z in itertools.product(range(nx), range(ny), range(nz)):
            s += r[z, y, x] * np.exp(-2j * np.pi * (kx * x / nx + ky * y / ny + kz * z / nz))
        assert abs(c[kx, ky, kz] - s) < 1e-12 * (1 + abs(s))


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 24, 25, 32, 48, 64, 96, 7, 11])
def test_fft1d_brute_force(n):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    k = np.arange(n)
    F = np.exp(-2j * np.pi * (np.outer(k, k) % n) / n)  # exact phases: reduce jk mod n first
    assert rel(O.fft1d(x, -1), F @ x) < 1e-14
    assert rel(O.fft1d(x, +1), F.conj() @ x) < 1e-14


def test_parseval_linearity_and_hermitian_plane():
    g = O.Grid(16, 12, 10)
    rng = np.random.default_rng(8)
    a, b = rng.standard_normal(g.rshape), rng.standard_normal(g.rshape)
    A, B = g.r2c(a), g.r2c(b)
    assert rel(g.r2c(2.5 * a - b), 2.5 * A - B) < 1e-15
    # Parseval with GHOST's weights: kx=0 plane once, kx>0 planes twice (no
    # separate Nyquist plane here because nx/2 is counted twice as GHOST does)
    N = 16 * 12 * 10
    w = np.full(g.nxh, 2.0)
    w[0] = 1.0
    w[-1] = 1.0  # the Nyquist kx plane is its own mirror
    lhs = np.sum(w[:, None, None] * np.abs(A) ** 2) / N
    assert abs(lhs - np.sum(a * a)) < 1e-12 * np.sum(a * a)
    # the kx=0 plane of a real field is Hermitian: A(0,-ky,-kz) = conj A(0,ky,kz)
    P = A[0]
    mirror = np.conj(P[(-np.arange(12)) % 12][:, (-np.arange(10)) % 10])
    assert np.max(np.abs(P - mirror)) < 1e-12 * np.max(np.abs(P))
