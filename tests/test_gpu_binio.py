"""Binary field output and restart of device-resident fields: main3D's output
block (c2r of v/N^3 and of its curl) written through the staged writer, read
back with numpy as raw Fortran-order FP64 and compared with the CPU oracle;
the restart path reads the files and transforms them back.  Fields larger
than the 64 MiB staging chunk cover the double-buffered copy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def to_gpu(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


@pytest.mark.parametrize("shape,L", [((32, 32, 32), (1.0, 1.0, 1.0)), ((64, 48, 32), (1.0, 1.5, 2.0))])
def test_hd_output_matches_oracle_and_restarts(gpu, tmp_path, shape, L):
    from ghost_amd import binio
    from ghost_amd.pseudospec import Grid
    from oracle import oracle as O
    nx, ny, nz = shape
    arb = L != (1.0, 1.0, 1.0)
    g, og = Grid(nx, ny, nz, L=L, arbsize=arb), O.Grid(nx, ny, nz, L=L, arbsize=arb)
    v, _ = og.initialv_k4exp(3.0, 1.0, 11)
    plan = binio.IOPlan(0, (nx, ny, nz), 1, nz)
    binio.hd_output(g, plan, str(tmp_path), binio.ext_label(3), [to_gpu(x) for x in v], outs=1)
    rmp = 1.0 / (nx * ny * nz)
    C = [x * rmp for x in v]
    w = og.curl(*C)
    for name, c in zip(("vx", "vy", "vz", "wx", "wy", "wz"), list(C) + list(w)):
        got = np.fromfile(tmp_path / f"{name}.0003.out").reshape((nx, ny, nz), order="F")
        want = og.c2r(c)  # (nz, ny, nx)
        assert rel(got.transpose(2, 1, 0), want) < 1e-13, name
    # restart: read vx, vy, vz and r2c back onto the spectral state
    back = [g.czeros() for _ in range(3)]
    binio.hd_restart(g, plan, str(tmp_path), "0003", back)
    for k in range(3):
        assert rel(back[k].cpu().numpy(), v[k]) < 1e-13


def test_device_staging_multi_chunk(gpu, tmp_path):
    """A 256^3 real field (128 MiB: three staging chunks) and a spectral field
    through io_write/io_read and io_writec/io_readc, bit-exact."""
    import torch
    from ghost_amd import binio
    n = 256
    gen = torch.Generator(device="cuda").manual_seed(5)
    r = torch.randn((n, n, n), dtype=torch.float64, device="cuda", generator=gen)
    p = binio.IOPlan(0, (n, n, n), 1, n)
    binio.io_write(1, str(tmp_path), "th", "0001", p, r)
    raw = np.fromfile(tmp_path / "th.0001.out")
    assert np.array_equal(raw, r.cpu().numpy().ravel())
    back = torch.empty_like(r)
    binio.io_read(1, str(tmp_path), "th", "0001", p, back)
    assert torch.equal(back, r)
    # the byte-swapped file read with iswap through the device path
    raw.byteswap().tofile(tmp_path / "sw.0001.out")
    ps = binio.IOPlan(0, (n, n, n), 1, n, iswap=1)
    binio.io_read(1, str(tmp_path), "sw", "0001", ps, back)
    assert torch.equal(back, r)
    c = torch.randn((n // 2 + 1, n, n), dtype=torch.complex128, device="cuda", generator=gen)
    pc = binio.IOPlan(0, (n, n, n // 2 + 1), 1, n // 2 + 1, complex_=True)
    binio.io_writec(1, str(tmp_path), "cv", "0001", pc, c)
    cb = torch.empty_like(c)
    binio.io_readc(1, str(tmp_path), "cv", "0001", pc, cb)
    assert torch.equal(cb, c)
