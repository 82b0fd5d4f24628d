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
    run = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=180, cwd=tmp_path)
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
    # GHOST's text diagnostics (spectrum, entrans kin 1 and 0, heltrans,
    # spectr1d of the x component along z, hdcheck) of a k^4 exp field
    # against a random forcing field, both drawn in the Fortran program
    v, _ = og.initialv_k4exp(3.0, 1.0, 11)
    w, _ = og.initialfv_random(1.0, 6.0, 1.0, 29)
    nb = n // 2 + 1

    def table(name):
        return np.loadtxt(tmp_path / name)

    def close(a, b):
        a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
        return np.all(np.abs(a - b) <= 1e-13 * np.max(np.abs(b)) + 1e-300)

    Ek, Hk = og.spectrumc(*v, kin=1, hel=1)
    for name, col in (("kspectrum.0001.txt", 0.5 * Ek), ("khelicity.0001.txt", Hk),
                      ("ktransfer.0001.txt", og.entrans(*v, *w, kin=1)),
                      ("mtransfer.0001.txt", og.entrans(*v, *w, kin=0)),
                      ("hktransfer.0001.txt", og.heltrans(*v, *w))):
        t = table(name)
        assert t.shape == (nb, 2), name
        assert np.array_equal(t[:, 0], np.arange(1, nb + 1, dtype=float)), name
        assert close(t[:, 1], col), name
    t = table("kspec1dxz.0001.txt")
    assert np.array_equal(t[:, 0], np.arange(nb, dtype=float))
    assert close(t[:, 1], 0.5 * og.spectr1d(v[0], 3))
    want = og.hdcheck(*v, *w, hel=1, chk=1)
    bal = table("balance.txt")
    assert bal[0] == pytest.approx(20 * 1e-3, rel=1e-6)
    assert min(abs(want[:4])) > 1e-8 * want[0]
    assert np.allclose(bal[1:], want[:3], rtol=1e-13, atol=1e-13 * max(abs(want[:3])))
    assert table("helicity.txt")[1] == pytest.approx(want[3], rel=1e-13, abs=1e-13 * want[0])
    assert table("divergence.txt")[1] == pytest.approx(want[4], abs=1e-13 * want[0])
    # the spectra about z: planes kz = (q-1) with values times Lz, cylinders
    # kperp = q-1 over Dkk (1 here); energy columns carry 1/2
    Ep, Hp = og.specpara(*v, kin=1, hel=1)
    Ec, Hc = og.specperp(*v, kin=1, hel=1)
    for name, cols in (("kspecpara", 0.5 * Ep), ("khelipara", Hp), ("kspecperp", 0.5 * Ec), ("kheliperp", Hc)):
        t = table(name + ".0001.txt")
        assert np.array_equal(t[:, 0], np.arange(cols.shape[1], dtype=float)), name
        for q in range(3):
            assert close(t[:, 1 + q], cols[q]), (name, q)
    for name, col in (("ktranpara", og.entpara(*v, *w, kin=1)), ("mtranperp", og.entperp(*v, *w, kin=0)),
                      ("hktranpara", og.heltpara(*v, *w))):
        t = table(name + ".0001.txt")
        assert close(t[:, 1], col), name
    # spec2D: one unformatted record (4-byte markers) of E(kperp, kpara)
    raw = (tmp_path / "kspec2D.0001.out").read_bytes()
    E2, _ = og.spec2d(*v, kin=1, hel=0)
    assert len(raw) == 8 + 8 * E2.size
    assert close(np.frombuffer(raw[4:-4], dtype=np.float64), E2.ravel())
    # scalar diagnostics of v_x with f_x as the source
    sc = table("scalar.txt")
    assert close(sc[1:], og.pscheck(v[0], w[0]))
    assert close(table("sspectrum.0001.txt")[:, 1], og.spectrscc(v[0]))
    assert close(table("s2transfer.0001.txt")[:, 1], og.sctrans(v[0], w[0]))
    pz = table("profilez.0001.txt")
    assert np.allclose(pz[:, 0], 2 * np.pi * np.arange(n) / n, rtol=1e-14)
    want_pz = og.difuc(v[0], v[1], 3)
    for q in range(3):
        assert close(pz[:, 1 + q], want_pz[q])
    # Boussinesq diagnostics of (v, f_x as the temperature), fo = 1.3, bv = 0.8
    avg, mx, mn = og.tbouss(*v, w[0], 1.3, 0.8)
    for name, col in (("tboussavg", avg), ("tboussmax", mx), ("tboussmin", mn)):
        assert close(table(name + ".txt")[1:], col), name
    assert close(table("pvprof.0001.txt")[:, 1], og.havgcomp(*v, w[0], 1.3, 0.8, 7))
    assert close(table("pvspectrum.0001.txt")[:, 1], og.spectpv(*v, w[0]))
    # the MHD loop: (v, a) = (the k^4 exp field, the random field), 3 RK2 steps
    got_m = np.fromfile(tmp_path / "fortran_mhd32.bin", dtype=np.complex128).reshape(6, nh, n, n)
    mv, ma = [x.copy() for x in v], [x.copy() for x in w]
    for _ in range(3):
        og.mhd_step(mv, ma, 1e-3, 1e-2, 1e-2, ord=2)
    for k, want in enumerate(mv + ma):
        assert np.linalg.norm(got_m[k] - want) <= 1e-12 * np.linalg.norm(want), k
    # GHOST's binary output of that v: the resident path (vx, vy, vz) and the
    # host-array path through io_write (hx) write the same bytes; the file is
    # c2r(v/N^3) in Fortran order
    vx_file = (tmp_path / "vx.0001.out").read_bytes()
    assert vx_file == (tmp_path / "hx.0001.out").read_bytes()
    for k, name in enumerate(("vx", "vy", "vz")):
        r = np.fromfile(tmp_path / f"{name}.0001.out").reshape((n, n, n), order="F").transpose(2, 1, 0)
        want = og.c2r(got[k] * (1.0 / n ** 3))
        assert np.linalg.norm(r - want) <= 1e-13 * max(np.linalg.norm(want), 1e-300) + 1e-16
    # fftp3d_real_to_complex on host arrays with 512-point x and z lines
    for shape in ((512, 8, 16), (16, 8, 512)):
        mx, my, mz = shape
        raw = np.fromfile(tmp_path / f"fortran_fftp_{mx}x{my}x{mz}.bin", dtype=np.float64)
        nr = mx * my * mz
        a = raw[:nr].reshape(mz, my, mx)  # Fortran a(mx,my,mz) read in C order
        s = raw[nr:].view(np.complex128).reshape(mx // 2 + 1, my, mz)  # s(mz,my,mx/2+1)
        ogs = O.Grid(mx, my, mz)
        want = ogs.r2c(a)
        assert np.linalg.norm(s - want) <= 1e-13 * np.linalg.norm(want), shape
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
