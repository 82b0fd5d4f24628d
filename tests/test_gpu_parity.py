"""Parity of the HIP path (through the C ABI) with the CPU oracle on seeded inputs.

Tolerances: relative L2 error 1e-13 for single transforms and operators,
1e-12 for multi-step time integration (the FP64 parity bar); exact equality
for integer/structural results (dealias masks, seeds, bin counts).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(8, 8, 8), (6, 4, 10), (12, 12, 12), (16, 8, 32), (32, 32, 32), (48, 48, 48), (64, 64, 64),
          (96, 96, 96), (128, 128, 128), (256, 64, 128), (20, 24, 6), (16, 512, 8), (24, 512, 16),
          # 512-point x and z lines (the drop-in's k_xfwd2/k_xinv2<512>, k_zfwd/k_zinv<512>)
          (512, 8, 16), (16, 8, 512), (512, 16, 512),
          # every other compiled schedule: 192, 384, 768 and 1024 points on each axis
          (384, 192, 8), (16, 384, 192), (768, 8, 16), (8, 8, 384), (16, 8, 1024), (8, 1024, 8), (2048, 8, 8)]

# the two configurations' own grids (one transform each way, oracle on all threads)
FULL_SHAPES = [(512, 512, 512), (512, 512, 256)]
# one GPU's share of the 512^3 transform in an 8-way slab decomposition: the
# y and x passes on 64 z-planes, and the z pass on the 64-wide x slab
SLAB_SHAPES = [(512, 512, 64), (64, 512, 512)]


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def to_gpu(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def to_np(t):
    return t.detach().cpu().numpy()


def grids(shape, arbsize=False, L=(1.0, 1.0, 1.0)):
    from ghost_amd.pseudospec import Grid
    from oracle import oracle as O
    nx, ny, nz = shape
    return Grid(nx, ny, nz, L=L, arbsize=arbsize), O.Grid(nx, ny, nz, L=L, arbsize=arbsize)


def rand_real(og, seed):
    return np.random.default_rng(seed).standard_normal(og.rshape)


def rand_spec(og, seed):
    """A Hermitian-consistent spectral field (FFT of a random real field)."""
    return og.r2c(rand_real(og, seed))


@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_c2r_match_oracle(gpu, shape):
    g, og = grids(shape)
    r = rand_real(og, 1)
    c_ref = og.r2c(r)
    dr = to_gpu(r)
    c = g.r2c(dr)
    assert rel(to_np(c), c_ref) < 1e-13
    cin = to_gpu(c_ref)
    keep = cin.clone()
    back = g.c2r(cin)
    assert rel(to_np(back), og.c2r(c_ref)) < 1e-13
    assert rel(to_np(back) / (shape[0] * shape[1] * shape[2]), r) < 1e-13
    assert bool((cin == keep).all()), "c2r must preserve its input"


@pytest.mark.parametrize("shape", FULL_SHAPES + SLAB_SHAPES)
def test_r2c_c2r_full_size_match_oracle(gpu, shape):
    """The plain transforms at the workload sizes (1 GB per field at 512^3)
    and at the per-GPU slab shapes of the 8-way decomposition."""
    import torch
    from oracle import oracle as O
    O.set_threads(16)
    g, og = grids(shape)
    r = rand_real(og, 3)
    c_ref = og.r2c(r)
    c = g.r2c(to_gpu(r))
    assert rel(to_np(c), c_ref) < 1e-13
    del c
    back = g.c2r(to_gpu(c_ref))
    del c_ref
    torch.cuda.empty_cache()
    n3 = shape[0] * shape[1] * shape[2]
    assert rel(to_np(back) / n3, r) < 1e-13


def test_fftp_module_interface(gpu):
    """fftp3d_create_plan / real_to_complex / complex_to_real with GHOST's argument order."""
    import torch
    from ghost_amd import fftp
    from oracle import oracle as O
    n = (32, 16, 24)
    og = O.Grid(*n)
    prc = fftp.fftp3d_create_plan(n, fftp.FFTW_REAL_TO_COMPLEX, fftp.FFTW_ESTIMATE)
    pcr = fftp.fftp3d_create_plan(n, fftp.FFTW_COMPLEX_TO_REAL, fftp.FFTW_ESTIMATE)
    r = rand_real(og, 5)
    out = torch.zeros(prc.cshape, dtype=torch.complex128, device="cuda")
    fftp.fftp3d_real_to_complex(prc, to_gpu(r), out, 0)
    assert rel(to_np(out), og.r2c(r)) < 1e-13
    back = torch.zeros(pcr.rshape, dtype=torch.float64, device="cuda")
    fftp.fftp3d_complex_to_real(pcr, out, back, 0)
    assert rel(to_np(back), og.c2r(og.r2c(r))) < 1e-13
    pcr.reset_timers()
    fftp.fftp3d_complex_to_real(pcr, out, back, 0)
    t = pcr.timers()
    assert t["ffttime"] > 0 and t["tottime"] == t["ffttime"]
    fftp.fftp3d_destroy_plan(prc)
    fftp.fftp3d_destroy_plan(pcr)


def test_unsupported_size_is_an_error(gpu):
    from ghost_amd import GhxError
    from ghost_amd.pseudospec import Grid
    with pytest.raises(GhxError):
        Grid(14)  # 7 has no compiled schedule


@pytest.mark.parametrize("shape,arb,L", [((16, 12, 8), False, (1, 1, 1)), ((32, 32, 16), True, (1.0, 1.0, 0.5)),
                                         ((24, 24, 24), False, (1, 1, 1))])
def test_pointwise_operators(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    a, b, c = (rand_spec(og, s) for s in (11, 12, 13))
    d, e, f = (rand_spec(og, s) for s in (14, 15, 16))
    A, B, Cc, D, E, F = map(to_gpu, (a, b, c, d, e, f))
    out = g.czeros()
    for dr in (1, 2, 3):
        assert rel(to_np(g.derivk3(A, out, dr)), og.derivk3(a, dr)) < 1e-15
        assert rel(to_np(g.rotor3(A, B, out, dr)), og.rotor3(a, b, dr)) < 1e-15
        assert rel(to_np(g.nonlhd3(A, B, Cc, out, dr)), og.nonlhd3(a, b, c, dr)) < 1e-14
        assert rel(to_np(g.nonlin3(A, B, Cc, D, E, F, out, dr)), og.nonlin3(a, b, c, d, e, f, dr)) < 1e-14
        assert rel(to_np(g.gauge3(A, B, Cc, out, dr)), og.gauge3(a, b, c, dr)) < 1e-14
    assert rel(to_np(g.laplak3(A, out)), og.laplak3(a)) < 1e-15
    for delta in (np.pi / 3.0, np.pi / 5.5):
        assert np.array_equal(to_np(g.sharp(A, out, delta)), og.sharp(a, delta))
    cd = np.exp(1j * 0.7321)
    assert rel(to_np(g.phaseshift(A.clone(), cd)), og.phaseshift(a, cd)) < 1e-15


@pytest.mark.parametrize("shape,arb,L", [((32, 32, 32), False, (1, 1, 1)), ((16, 24, 12), False, (1, 1, 1)),
                                         ((32, 32, 16), True, (1.0, 1.0, 0.5))])
def test_products(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    v = [rand_spec(og, s) for s in (21, 22, 23)]
    b = [rand_spec(og, s) for s in (24, 25, 26)]
    V, Bv = [to_gpu(x) for x in v], [to_gpu(x) for x in b]
    out = [g.czeros() for _ in range(3)]
    ref = og.prodre3(*v)
    g.prodre3(*V, *out)
    for k in range(3):
        assert rel(to_np(out[k]), ref[k]) < 1e-13
    ref = og.vector3(*v, *b)
    g.vector3(*V, *Bv, *out)
    for k in range(3):
        assert rel(to_np(out[k]), ref[k]) < 1e-13


@pytest.mark.parametrize("shape,arb,L", [((32, 32, 32), False, (1, 1, 1)), ((32, 32, 16), True, (1.0, 1.0, 0.5)),
                                         ((24, 16, 12), False, (1, 1, 1))])
def test_diagnostics(gpu, shape, arb, L):
    g, og = grids(shape, arb, L)
    v = [rand_spec(og, s) for s in (31, 32, 33)]
    w = [rand_spec(og, s) for s in (34, 35, 36)]
    V, W = [to_gpu(x) for x in v], [to_gpu(x) for x in w]
    for kin in (0, 1, 2):
        assert g.energy(*V, kin) == pytest.approx(og.energy(*v, kin), rel=1e-13)
        assert g.maxabs(*V, kin) == pytest.approx(og.maxabs(*v, kin), rel=1e-13)
    assert g.helicity(*V) == pytest.approx(og.helicity(*v), rel=1e-12, abs=1e-14 * og.energy(*v, 1))
    for kin in (0, 1):
        assert g.cross(*V, *W, kin) == pytest.approx(og.cross(*v, *w, kin), rel=1e-12, abs=1e-13)
    assert g.divergence(*V) == pytest.approx(og.divergence(*v), rel=1e-13)
    assert g.nbins == og.nbins
    for kin in (0, 1):
        Ek, Hk = g.spectrumc(*V, kin=kin, hel=1)
        Er, Hr = og.spectrumc(*v, kin=kin, hel=1)
        assert rel(Ek, Er) < 1e-13
        assert rel(Hk, Hr) < 1e-12


def test_normalize_and_initial_conditions(gpu):
    g, og = grids((32, 32, 32))
    vt = g.initialv_tg(1.0, 1.0, 0.5)
    vr = og.initialv_tg(1.0, 1.0, 0.5)
    for k in range(3):
        assert rel(to_np(vt[k]), vr[k]) < 1e-13
    f, s1 = g.initialfv_random(2.0, 3.0, 0.25, 1357)
    fr, s2 = og.initialfv_random(2.0, 3.0, 0.25, 1357)
    assert s1 == s2
    for k in range(3):
        assert rel(to_np(f[k]), fr[k]) < 1e-13
    v, s1 = g.initialv_k4exp(3.0, 1.0, 97531)
    vr, s2 = og.initialv_k4exp(3.0, 1.0, 97531)
    assert s1 == s2
    for k in range(3):
        assert rel(to_np(v[k]), vr[k]) < 1e-13


def test_randu_matches_oracle(gpu):
    from ghost_amd.pseudospec import randu
    from oracle import oracle as O
    s1 = s2 = 12345
    for _ in range(1000):
        a, s1 = randu(s1)
        b, s2 = O.randu(s2)
        assert a == b and s1 == s2


@pytest.mark.parametrize("shape", [(32, 32, 32), (24, 16, 12), (64, 64, 64), (16, 16, 512), (8, 12, 512),
                                   (16, 512, 16), (48, 48, 48), (96, 96, 96), (96, 32, 96), (96, 48, 24),
                                   (192, 16, 16), (16, 384, 32)])
def test_hd_substep_fused(gpu, shape):
    g, og = grids(shape)
    v, _ = og.initialv_k4exp(2.5, 1.0, 2024)
    f, _ = og.initialfv_random(2.0, 3.0, 0.5, 77)
    c = [x + 0.01 * rand_spec(og, 40 + k) for k, x in enumerate(v)]
    V, Cg, Fg = [to_gpu(x) for x in v], [to_gpu(x) for x in c], [to_gpu(x) for x in f]
    for o in (2, 1):
        out = [g.czeros() for _ in range(3)]
        g.hd_rkstep2(o, 2.5e-3, 3e-3, V, Cg, Fg, out)
        vr = [x.copy() for x in v]
        og.hd_substep(vr, c, o, 2.5e-3, 3e-3, f=f)
        for k in range(3):
            assert rel(to_np(out[k]), vr[k]) < 1e-13


def test_hd_taylor_green_64_100_steps(gpu):
    """HD3D 64^3 FP64 Taylor-Green, nu=1e-2, RK2, dt=1e-3, 100 steps, no forcing."""
    from ghost_amd.solvers import HDSolver
    g, og = grids((64, 64, 64))
    v0 = og.initialv_tg(1.0, 1.0, 0.5)
    s = HDSolver(g, dt=1e-3, nu=1e-2, ord=2)
    s.set_velocity([to_gpu(x) for x in v0])
    s.set_forcing(None)
    s.step(100)
    got = s.velocity()
    ref = [x.copy() for x in v0]
    for _ in range(100):
        og.hd_step(ref, 1e-3, 1e-2, ord=2)
    for k in range(3):
        assert rel(to_np(got[k]), ref[k]) < 1e-12, k
    Ek, _ = g.spectrumc(*got, kin=1, hel=0)
    Er, _ = og.spectrumc(*ref, kin=1, hel=0)
    assert rel(Ek, Er) < 1e-12
    assert g.energy(*got, 1) == pytest.approx(og.energy(*ref, 1), rel=1e-12)


def test_hd_forced_random_32(gpu):
    """Forced k^4 exp IC with constant-amplitude random-phase forcing, phase shifted
    every 5 steps (rand=1 of main3D.fpp:2544-2556)."""
    from ghost_amd.solvers import HDSolver
    from oracle import oracle as O
    g, og = grids((32, 32, 32))
    v0, _ = og.initialv_k4exp(3.0, 1.0, 11)
    f0, seed = og.initialfv_random(2.0, 3.0, 0.3, 29)
    s = HDSolver(g, dt=2.5e-3, nu=2e-3, ord=2)
    s.set_velocity([to_gpu(x) for x in v0])
    s.set_forcing([to_gpu(x) for x in f0])
    ref, fr = [x.copy() for x in v0], [x.copy() for x in f0]
    sd = seed
    for t in range(20):
        if t % 5 == 0:
            r, sd = O.randu(sd)
            phase = 2 * np.pi * r
            cd = np.cos(phase) + 1j * np.sin(phase)
            fr = [og.phaseshift(x, cd) for x in fr]
            s.phaseshift_forcing(cd)
            s.set_forcing(s.forcing())
        s.step(1)
        og.hd_step(ref, 2.5e-3, 2e-3, f=fr, ord=2)
    got = s.velocity()
    for k in range(3):
        assert rel(to_np(got[k]), ref[k]) < 1e-12


def test_mhd_anisotropic_box(gpu):
    """MHD on a 32x32x16 box with Lz = pi (Lz=0.5 in units of 2 pi), forced, RK2."""
    from ghost_amd.solvers import MHDSolver
    L = (1.0, 1.0, 0.5)
    g, og = grids((32, 32, 16), True, L)
    v0, _ = og.initialv_k4exp(3.0, 1.0, 5)
    a0, _ = og.initialv_k4exp(3.0, 0.3, 6)
    f0, _ = og.initialfv_random(2.0, 3.0, 0.3, 7)
    s = MHDSolver(g, dt=2e-3, nu=3e-3, mu=3e-3, ord=2)
    s.set_state([to_gpu(x) for x in v0], [to_gpu(x) for x in a0])
    s.set_forcing([to_gpu(x) for x in f0], None)
    s.step(5)
    gv, ga = s.state()
    rv, ra = [x.copy() for x in v0], [x.copy() for x in a0]
    for _ in range(5):
        og.mhd_step(rv, ra, 2e-3, 3e-3, 3e-3, f=f0, ord=2)
    for k in range(3):
        assert rel(to_np(gv[k]), rv[k]) < 1e-12
        assert rel(to_np(ga[k]), ra[k]) < 1e-12


def test_hd_not_dealiased_start_then_pruned(gpu):
    """A state with energy outside the dealias set takes the full path for the
    first step (as GHOST does) and the pruned path after; both match the oracle."""
    from ghost_amd.solvers import HDSolver
    g, og = grids((32, 32, 32))
    v0 = [rand_spec(og, 60 + k) * 1e-3 for k in range(3)]  # all modes populated
    s = HDSolver(g, dt=1e-3, nu=5e-3, ord=2)
    s.set_velocity([to_gpu(x) for x in v0])
    s.set_forcing(None)
    s.step(3)
    got = s.velocity()
    ref = [x.copy() for x in v0]
    for _ in range(3):
        og.hd_step(ref, 1e-3, 5e-3, ord=2)
    for k in range(3):
        assert rel(to_np(got[k]), ref[k]) < 1e-12


def test_hd_pruned_equals_full_path(gpu):
    """Dealias pruning skips exact zeros and discarded outputs only: the pruned
    and full transforms agree to rounding on a dealiased state."""
    import os
    from ghost_amd.solvers import HDSolver
    g, og = grids((48, 48, 48))
    v0, _ = og.initialv_k4exp(3.0, 1.0, 4242)
    outs = []
    for env in (None, "1"):
        if env:
            os.environ["GHX_NO_PRUNE"] = env
        try:
            s = HDSolver(g, dt=2e-3, nu=4e-3, ord=2)
            s.set_velocity([to_gpu(x) for x in v0])
            s.set_forcing(None)
            s.step(4)
            outs.append([to_np(x) for x in s.velocity()])
        finally:
            os.environ.pop("GHX_NO_PRUNE", None)
    for k in range(3):
        assert rel(outs[0][k], outs[1][k]) < 1e-14


@pytest.mark.parametrize("big,small", [((32, 32, 32), (16, 16, 16)), ((64, 48, 32), (32, 24, 16)),
                                       ((128, 128, 128), (48, 48, 48))])
def test_truncate_matches_oracle(gpu, big, small):
    """trunc3D's spectral truncation onto a smaller grid, then the c2r there."""
    from ghost_amd.pseudospec import Grid
    from oracle import oracle as O
    g, og = Grid(*big), O.Grid(*big)
    t, ot = Grid(*small), O.Grid(*small)
    c = og.r2c(rand_real(og, 9))
    want = og.truncate(c, ot)
    got = g.truncate(to_gpu(c), t)
    assert rel(to_np(got), want) < 1e-15
    assert rel(to_np(t.c2r(got)), ot.c2r(want)) < 1e-13


def test_mhd_512x512x256_step_matches_oracle(gpu):
    """The MHD configuration at its full size (512x512x256, Lz = pi): one
    forced RK2 step from k^4 exp velocity and vector potential."""
    from ghost_amd.solvers import MHDSolver
    from oracle import oracle as O
    O.set_threads(16)
    L = (1.0, 1.0, 0.5)
    g, og = grids((512, 512, 256), True, L)
    v0, _ = og.initialv_k4exp(3.0, 1.0, 1000)
    a0, _ = og.initialv_k4exp(3.0, 1.0, 1001)
    a0 = list(og.normalize(*a0, 1.0, kin=0))
    f0, _ = og.initialfv_random(2.0, 3.0, 0.4, 2000)
    s = MHDSolver(g, dt=2e-4, nu=3e-4, mu=3e-4, ord=2)
    s.set_state([to_gpu(x) for x in v0], [to_gpu(x) for x in a0])
    s.set_forcing([to_gpu(x) for x in f0], None)
    s.step(1)
    gv, ga = s.state()
    gv, ga = [to_np(x) for x in gv], [to_np(x) for x in ga]
    del s
    og.mhd_step(v0, a0, 2e-4, 3e-4, 3e-4, f=f0, ord=2)
    for k in range(3):
        assert rel(gv[k], v0[k]) < 1e-12
        assert rel(ga[k], a0[k]) < 1e-12


def test_mhd_pruned_equals_full_path(gpu):
    """The MHD chain with and without dealias pruning on a dealiased state
    (velocity and vector potential from the k^4 exp fill), 4 forced steps."""
    import os
    from ghost_amd.solvers import MHDSolver
    g, og = grids((48, 48, 48))
    v0, _ = og.initialv_k4exp(3.0, 1.0, 515)
    a0, _ = og.initialv_k4exp(3.0, 0.5, 516)
    f0, _ = og.initialfv_random(2.0, 3.0, 0.3, 517)
    outs = []
    for env in (None, "1"):
        if env:
            os.environ["GHX_NO_PRUNE"] = env
        try:
            s = MHDSolver(g, dt=2e-3, nu=4e-3, mu=4e-3, ord=2)
            s.set_state([to_gpu(x) for x in v0], [to_gpu(x) for x in a0])
            s.set_forcing([to_gpu(x) for x in f0], None)
            s.step(4)
            gv, ga = s.state()
            outs.append([to_np(x) for x in gv + ga])
        finally:
            os.environ.pop("GHX_NO_PRUNE", None)
    for k in range(6):
        assert rel(outs[0][k], outs[1][k]) < 1e-14, k


def _dirty_device_memory(nbytes, value=1.0):
    """Allocate nbytes through the C ABI, fill them with `value`, free them:
    the next allocation of that size usually gets these bytes back, so a
    buffer that a pruned pass forgets to write holds `value`, not zero."""
    import ctypes as C
    from ghost_amd import _lib
    lib = _lib.load()
    p = C.c_void_p()
    _lib.check(lib.ghx_malloc(C.byref(p), nbytes))
    host = np.full(nbytes // 8, value)
    _lib.check(lib.ghx_memcpy_h2d(p, host.ctypes.data_as(C.c_void_p), nbytes))
    _lib.check(lib.ghx_free(p))


@pytest.mark.parametrize("solver", ["hd", "mhd"])
def test_pruned_steps_write_the_mean_mode(gpu, solver):
    """The pruned RK updates skip the modes outside the dealias set but must
    still write the k = 0 mode (GHOST's update zeroes kn2 < tiny): the next
    substep's inverse transform reads it from the intermediate state.  The
    solver is created on device memory left holding 1.0 everywhere, and its
    pruned steps must match the oracle at 1e-12."""
    from ghost_amd.solvers import HDSolver, MHDSolver
    n = 48
    g, og = grids((n, n, n))
    nc = (n // 2 + 1) * n * n
    slot = n * n * ((n // 2 + 1 + 7) & ~7)
    v, _ = og.initialv_k4exp(3.0, 1.0, 17)
    if solver == "hd":
        _dirty_device_memory((9 * nc + 12 * slot) * 16)
        s = HDSolver(g, dt=2e-3, nu=5e-3, ord=2)
        s.set_velocity([to_gpu(x) for x in v])
        s.set_forcing(None)
        s.step(3)
        got = [to_np(x) for x in s.velocity()]
        ref = [x.copy() for x in v]
        for _ in range(3):
            og.hd_step(ref, 2e-3, 5e-3, ord=2)
    else:
        a, _ = og.initialv_k4exp(3.0, 0.5, 18)
        _dirty_device_memory((18 * nc + 24 * slot) * 16)
        s = MHDSolver(g, 2e-3, 5e-3, 5e-3, ord=2)
        s.set_state([to_gpu(x) for x in v], [to_gpu(x) for x in a])
        s.set_forcing(None, None)
        s.step(3)
        gv, ga = s.state()
        got = [to_np(x) for x in list(gv) + list(ga)]
        ref = [x.copy() for x in v] + [x.copy() for x in a]
        for _ in range(3):
            og.mhd_step(ref[:3], ref[3:], 2e-3, 5e-3, 5e-3, ord=2)
    for k in range(len(ref)):
        assert rel(got[k], ref[k]) < 1e-12, k
