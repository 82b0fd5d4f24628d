"""GHOST's binary field files (mpiio/binary_io.f90) through the C ABI, host
arrays only (no GPU needed): the file is the global array in Fortran order as
raw native FP64 with no header, so numpy's raw dump of a C-ordered (nz, ny, nx)
array is the expected file byte for byte."""
import os

import numpy as np
import pytest

from ghost_amd import binio
from ghost_amd._lib import GhxError


def rand_real(shape, seed=0):
    return np.random.default_rng(seed).standard_normal(shape)


def test_write_is_raw_fortran_order(tmp_path):
    nx, ny, nz = 8, 6, 4
    a = rand_real((nz, ny, nx))
    p = binio.IOPlan(0, (nx, ny, nz), 1, nz)
    binio.io_write(1, str(tmp_path), "vx", binio.ext_label(7), p, a)
    f = tmp_path / "vx.0007.out"
    assert f.stat().st_size == nx * ny * nz * 8
    assert f.read_bytes() == a.tobytes()
    # Fortran view of the same file: element (i, j, k) = a[k, j, i]
    fa = np.fromfile(f).reshape((nx, ny, nz), order="F")
    assert fa[3, 2, 1] == a[1, 2, 3]
    b = np.zeros_like(a)
    binio.io_read(1, str(tmp_path), "vx", "0007", p, b)
    assert np.array_equal(a, b)


def test_slabs_share_one_file(tmp_path):
    """Two ranks' slabs (ksta:kend) land at their offsets in one file; the
    second open neither truncates nor moves the first slab."""
    nx, ny, nz = 4, 3, 6
    a = rand_real((nz, ny, nx), 1)
    for ksta, kend in ((4, 6), (1, 3)):
        p = binio.IOPlan(0, (nx, ny, nz), ksta, kend)
        binio.io_write(1, str(tmp_path), "th", "0001", p, np.ascontiguousarray(a[ksta - 1:kend]))
    assert (tmp_path / "th.0001.out").read_bytes() == a.tobytes()
    p = binio.IOPlan(1, (nx, ny, nz), 2, 5)
    b = np.zeros((4, ny, nx))
    binio.io_read(1, str(tmp_path), "th", "0001", p, b)
    assert np.array_equal(b, a[1:5])


def test_disp_and_unmangled_names(tmp_path, monkeypatch):
    nx, ny, nz = 4, 4, 2
    a = rand_real((nz, ny, nx), 2)
    p = binio.IOPlan(0, (nx, ny, nz), 1, nz, bmangle=0, disp=24)
    # io_write with bmangle = 0 opens fname alone (relative to the working directory)
    monkeypatch.chdir(tmp_path)
    binio.io_write(1, "ignored_dir", "raw.bin", "0001", p, a)
    data = (tmp_path / "raw.bin").read_bytes()
    assert len(data) == 24 + a.nbytes and data[24:] == a.tobytes()
    # io_read with bmangle = 0 opens dir/fname
    b = np.zeros_like(a)
    binio.io_read(1, str(tmp_path), "raw.bin", "0001", p, b)
    assert np.array_equal(a, b)


def test_iswap_reverses_bytes_on_read(tmp_path):
    nx, ny, nz = 4, 2, 2
    a = rand_real((nz, ny, nx), 3)
    a.byteswap().tofile(tmp_path / "vy.0002.out")  # a file from a machine of the other endianness
    p = binio.IOPlan(0, (nx, ny, nz), 1, nz, iswap=1)
    b = np.zeros_like(a)
    binio.io_read(1, str(tmp_path), "vy", "0002", p, b)
    assert np.array_equal(a, b)


def test_complex_files(tmp_path):
    """io_writec/io_readc: complex pairs in Fortran order; io_readc never swaps."""
    nx, ny, nz = 8, 4, 4
    rng = np.random.default_rng(4)
    c = rng.standard_normal((nx // 2 + 1, ny, nz)) + 1j * rng.standard_normal((nx // 2 + 1, ny, nz))
    p = binio.IOPlan(0, (nz, ny, nx // 2 + 1), 1, nx // 2 + 1, complex_=True, iswap=1)
    binio.io_writec(1, str(tmp_path), "cvx", "0003", p, c)
    assert (tmp_path / "cvx.0003.out").read_bytes() == c.tobytes()
    d = np.zeros_like(c)
    binio.io_readc(1, str(tmp_path), "cvx", "0003", p, d)
    assert np.array_equal(c, d)
    # a complex plan refuses the real routines and vice versa
    with pytest.raises(GhxError):
        binio.io_read(1, str(tmp_path), "cvx", "0003", p, np.zeros(2 * c.size))


def test_errors(tmp_path):
    nx, ny, nz = 4, 4, 4
    p = binio.IOPlan(0, (nx, ny, nz), 1, nz)
    b = np.zeros((nz, ny, nx))
    with pytest.raises(GhxError, match="cannot open"):
        binio.io_read(1, str(tmp_path), "missing", "0001", p, b)
    np.zeros(10).tofile(tmp_path / "short.0001.out")
    with pytest.raises(GhxError, match="shorter"):
        binio.io_read(1, str(tmp_path), "short", "0001", p, b)
    with pytest.raises(GhxError):
        binio.IOPlan(0, (nx, ny, nz), 3, 2)
    with pytest.raises(GhxError):
        binio.IOPlan(0, (nx, ny, nz), 1, nz + 1)
    assert not os.path.exists(tmp_path / "missing.0001.out")
