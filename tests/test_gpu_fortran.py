"""The Fortran host side (ghost_amd/fortran: GHOST's fftp3d_* and pseudospec
interfaces over iso_c_binding) run end to end on the GPU, its results
compared with the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ghost_amd", "fortran", "test_fftp_hip")


def test_fortran_host_side(gpu, tmp_path):
    from oracle import oracle as O
    assert os.path.exists(EXE), "build the Fortran host side first (__graft_entry__.build)"
    run = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=180)
    print(run.stdout)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "ALL OK" in run.stdout
    n, nh = 32, 17
    og = O.Grid(n)
    # RK2 loop of main3D.fpp written in Fortran on resident fields: 20 steps of TG
    got = np.fromfile(tmp_path / "fortran_hd32.bin", dtype=np.complex128).reshape(3, nh, n, n)
    v = og.initialv_tg(1.0, 1.0, 0.5)
    for _ in range(20):
        og.hd_step(v, 1e-3, 1e-2, ord=2)
    for k in range(3):
        assert np.linalg.norm(got[k] - v[k]) <= 1e-12 * max(np.linalg.norm(v[k]), np.linalg.norm(v[0]))
    # GHOST's binary output of that v: the resident path (vx, vy, vz) and the
    # host-array path through io_write (hx) write the same bytes; the file is
    # c2r(v/N^3) in Fortran order
    vx_file = (tmp_path / "vx.0001.out").read_bytes()
    assert vx_file == (tmp_path / "hx.0001.out").read_bytes()
    for k, name in enumerate(("vx", "vy", "vz")):
        r = np.fromfile(tmp_path / f"{name}.0001.out").reshape((n, n, n), order="F").transpose(2, 1, 0)
        want = og.c2r(got[k] * (1.0 / n ** 3))
        assert np.linalg.norm(r - want) <= 1e-13 * max(np.linalg.norm(want), 1e-300) + 1e-16
    # resident solver with 1000 tracers attached, 5 steps, GHOST's z stencil
    raw = np.fromfile(tmp_path / "fortran_tracers32.bin", dtype=np.float64)
    p0 = raw[:3000].reshape(3, 1000)  # Fortran pos(npart, 3): column c is axis c
    p1 = raw[3000:].reshape(3, 1000)
    v = og.initialv_tg(1.0, 1.0, 0.5)
    pos = [p0[c].copy() for c in range(3)]
    for _ in range(5):
        og.hd_step_tracers(v, pos, 1e-3, 1e-2, ord=2, zcompat=True)
    for c in range(3):
        d = np.mod(p1[c] - pos[c] + n / 2, n) - n / 2
        assert np.max(np.abs(d)) < 1e-10
