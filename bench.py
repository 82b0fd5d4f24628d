#!/usr/bin/env python3
"""HD3D FP64 RK2 steps/s on one MI355X (gfx950), GHOST's pseudospectral hot path.

Workload (the 512^3 forced-turbulence configuration): 512^3 cubic 2*pi box,
FP64, random-phase solenoidal initial velocity with E(k) ~ k^4 exp(-2 (k/3)^2)
normalised to total energy 0.5, constant-amplitude random-phase forcing in the
shell 2 <= |k| <= 3 whose phase is shifted every FSTEP steps (GHOST rand=1),
nu = 2e-4, RK2, dt = 2.5e-4, plus 10^6 Lagrangian tracers at uniform-random
seeded positions advanced with cubic-spline velocity interpolation (one tracer
stage after each fluid substep, GHOST's lagpart coupling).  A step is one
full RK2 step of the fused HIP path, tracers included, with all fields and
particles resident in HBM.

Prints one JSON line on stdout.  Everything else goes to stderr.
  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512] [--tracers 1000000] [--no-cpu-baseline]

--workload mhd measures the MHD configuration instead (512x512x256 box with
Lz = pi, nu = eta = 3e-4, dt = 2e-4, velocity and vector potential from the
k^4 exp IC at unit energy ratio, mechanical forcing in 2 <= |k| <= 3).
--workload les measures the learned-closure LES: a 512^3 DNS of the workload
above (without tracers, other seeds) trains a 10-64-64-6 closure on its
filtered states (kc = 42, 128^3) with a per-shell spectral part fitted from
the network's residual subgrid transfer and calibrated a posteriori over the
8000 steps that follow its spin-up; a second 512^3 DNS (the workload's seeds)
is the reference: a 2000-step 128^3 LES (dt = 4 DNS steps) from its filtered
state is judged against its filtered spectrum over the following 8000 DNS
steps, and a run-time check against the training DNS's spectrum on the shells
above the forcing band falls back to the DNS.
--workload slab times one GPU's share of an 8-way slab decomposition of the
512^3 transform at its real shapes: the y and x passes on 64 z-planes (a
512x512x64 grid) and the z pass on 33 of the 257 kx columns (a 64x512x512
grid), 20 r2c and 20 c2r each.  The all-to-all between them is not built.
--workload ensemble measures the batched ensemble: 96 unforced HD runs, 24
each at 32^3, 48^3, 64^3 and 96^3, k^4 exp(-2 (k/k0)^2) ICs with k0 drawn
uniformly in [2, 4], nu set so that k_max eta = 1.5 at the start, RK2.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "HD3D 512³ FP64 steps/s + FFT HBM GB/s; learned-LES speedup at ≤5% spectrum err"
HBM_PEAK_GBS = 8000.0
NU, DT, K0, U0, F0 = 2e-4, 2.5e-4, 3.0, 1.0, 0.4
KDN, KUP = 2.0, 3.0
FSTEP = 20
IC_SEED, F_SEED = 1000, 2000
PART_SEED = 3000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(kernel, n, npart):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary, if any.

    The summary carries one line per pass: `traffic <pass> <bytes per launch>`."""
    path = os.path.join(ROOT, "profiles", f"hd{n}_rk2_tracers_hbm_traffic.md")
    if npart != 1_000_000:
        return None  # the summary covers the workload with 10^6 tracers
    try:
        with open(path) as fh:
            for line in fh:
                f = line.split()
                if len(f) == 3 and f[0] == "traffic" and f[1] == kernel:
                    return float(f[2])
    except OSError:
        pass
    return None


def tracer_positions(n, npart):
    """Uniform-random seeded tracer positions in grid units (the seed file of GPINIT_USERLOC)."""
    import numpy as np
    rng = np.random.default_rng(PART_SEED)
    return [rng.uniform(0.0, float(n), npart) for _ in range(3)]


# MHD configuration: 512 x 512 x 256, Lx = Ly = 2 pi, Lz = pi
MHD_SHAPE, MHD_L = (512, 512, 256), (1.0, 1.0, 0.5)
MHD_NU, MHD_DT = 3e-4, 2e-4


def time_steps(one_step, steps, warmup, lib, torch):
    """warmup untimed steps, then steps timed with a sync on both sides and
    the pass profiler's HIP events on; returns (elapsed s, profile)."""
    for _ in range(warmup):
        one_step()
    torch.cuda.synchronize()
    lib.prof_reset()
    lib.prof_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib.prof_enable(False)
    return t1 - t0, lib.prof_read()


def kernel_name(slot, n):
    """The kernel a profiler slot times at size n (at 512 points the y inverse
    and the spline z inverse run the 16-points-per-lane schedule, the HD z
    forward update the four-points-per-lane one)."""
    if slot in ("ypass_inv", "ypass_inv_tracer") and n == 512:
        return "ghx::k_yinv<512, FFTCfgWide<512> >"
    if slot == "zfwd_rk_hd" and n == 512:
        return "ghx::k_zfwd_rk_hd<512, FFTCfgLow<512>, 1>"
    if slot == "zinv_spline" and n == 512:
        return "ghx::k_zinv<512, 3, FFTCfgWide<512> >"
    return KERNEL_NAME.get(slot, slot).replace("<N", f"<{n}")


# profiler slots that time the same kernel as another slot
SAME_KERNEL = {"ypass_inv_tracer": "ypass_inv"}
KERNEL_NAME = {"ypass_inv": "ghx::k_yinv<N>", "xcross": "ghx::k_xcross2<N>",
               "zinv_curl": "ghx::k_zinv<N, 1, FFTCfg<N> >", "zfwd_rk_hd": "ghx::k_zfwd_rk_hd<N, FFTCfg<N> >",
               "ypass_fwd": "ghx::k_yfwd<N, FFTCfg<N> >", "xinv_tracer": "ghx::k_xinv3i2<N>",
               "tracer_stage": "ghx::k_tracer_stage", "zinv_spline": "ghx::k_zinv<N, 3, FFTCfg<N> >"}


def stream_bandwidth(torch, gib=2, reps=5):
    """Achievable HBM bandwidth from the library's own copy, read and write
    kernels (ghx_stream_probe) over `gib` GiB buffers: bytes moved (read +
    written) over the mean time of `reps` launches, best over a few grid sizes,
    over plain and streaming stores and over grid-stride and one-chunk-per-block
    forms.  Returns {name: GB/s}."""
    from ghost_amd import _lib
    lib = _lib.load()
    nbytes = gib << 30
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.fill_(0.0)
    sink = torch.zeros(1, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    kinds = {"copy": (0, 2), "copy_nt": (3, 2), "read": (1, 1), "write": (2, 1), "write_nt": (4, 1),
             "copy_chunk": (5, 2), "copy_chunk_nt": (6, 2), "read_chunk": (7, 1), "write_chunk": (8, 1),
             "write_chunk_nt": (9, 1)}
    out = {}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for name, (kind, moved) in kinds.items():
        best = 0.0
        for blocks in (1024, 2048, 4096, 8192, 16384, 32768):
            dst = sink.data_ptr() if kind in (1, 7) else b.data_ptr()
            src = a.data_ptr()
            _lib.check(lib.ghx_stream_probe(kind, dst, src, nbytes, blocks, stream))
            torch.cuda.synchronize()
            e0.record()
            for _ in range(reps):
                lib.ghx_stream_probe(kind, dst, src, nbytes, blocks, stream)
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / reps * 1e-3
            best = max(best, moved * nbytes / t / 1e9)
        out[name] = round(best, 1)
    del a, b, sink
    torch.cuda.empty_cache()
    return out


def transform_rates(torch, n, reps=10):
    """Plain r2c and c2r of one real field at n^3 (ghx_r2c, ghx_c2r): the
    per-transform effective rate B_fft / T with B_fft = S_r + S_c."""
    from ghost_amd.pseudospec import Grid
    g = Grid(n)
    r = torch.randn(g.rshape, dtype=torch.float64, device="cuda")
    c = g.r2c(r)
    out_r = torch.empty_like(r)
    b_fft = r.numel() * 8.0 + c.numel() * 16.0
    res = {}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for name, fn in (("r2c", lambda: g.r2c(r, c)), ("c2r", lambda: g.c2r(c, out_r))):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms": round(ms, 4), "B_fft_GB": round(b_fft / 1e9, 4),
                     "effective_GBps": round(b_fft / ms / 1e6, 1)}
    del g, r, c, out_r
    torch.cuda.empty_cache()
    return res


def pass_table(prof, elapsed):
    passes = {}
    for name, (cnt, ms, by) in prof.items():
        avg_s = ms / cnt * 1e-3
        passes[name] = {"launches": cnt, "avg_ms": round(ms / cnt, 4),
                        "algorithmic_GB_per_launch": round(by / cnt / 1e9, 4),
                        "achieved_GBps": round(by / cnt / avg_s / 1e9, 1),
                        "share_of_step": round(ms * 1e-3 / elapsed, 4)}
        log(f"  {name:12s} x{cnt:4d}  avg {ms / cnt:8.3f} ms  {by / cnt / 1e9:7.3f} GB  "
            f"{by / cnt / avg_s / 1e9:8.1f} GB/s")
    # the roofline is taken per kernel: the tracer stage's Y^-1 launches run the
    # same kernel as the fluid's (3 fields instead of 6), so they are pooled the
    # way a kernel trace pools them
    groups = {}
    for name, (cnt, ms, by) in prof.items():
        key = SAME_KERNEL.get(name, name)
        c0, m0, b0 = groups.get(key, (0, 0.0, 0.0))
        groups[key] = (c0 + cnt, m0 + ms, b0 + by)
    dom = max(groups.items(), key=lambda kv: kv[1][1])[0]
    cnt, ms, by = groups[dom]
    members = [k for k in prof if SAME_KERNEL.get(k, k) == dom]
    return passes, dom, by / cnt / (ms / cnt * 1e-3) / 1e9, members, ms / cnt


def run_mhd(args, torch):
    """The MHD configuration; returns the JSON record."""
    from ghost_amd import _lib
    from ghost_amd.pseudospec import Grid, randu
    from ghost_amd.solvers import MHDSolver
    nx, ny, nz = MHD_SHAPE
    g = Grid(nx, ny, nz, L=MHD_L, arbsize=True)
    s = MHDSolver(g, dt=MHD_DT, nu=MHD_NU, mu=MHD_NU, ord=2)
    v, _ = g.initialv_k4exp(K0, U0, IC_SEED)
    a, _ = g.initialv_k4exp(K0, U0, IC_SEED + 1)
    g.normalize(*a, U0, kin=0)  # <b^2> = <u^2>: unit energy ratio
    f, fseed = g.initialfv_random(KDN, KUP, F0, F_SEED)
    s.set_state(v, a)
    s.set_forcing(f, None)
    ev, eb = g.energy(*v, 1), g.energy(*a, 0)
    del v, a, f
    torch.cuda.empty_cache()
    log(f"MHD grid {nx}x{ny}x{nz}, L = {MHD_L}, <u^2> = {ev:.6f}, <b^2> = {eb:.6f}")
    state = {"timef": FSTEP, "seed": fseed}

    def one_step():
        if state["timef"] == FSTEP:
            state["timef"] = 0
            r, state["seed"] = randu(state["seed"])
            phase = 2 * math.pi * r
            s.phaseshift_forcing(complex(math.cos(phase), math.sin(phase)))
        s.step(1)
        state["timef"] += 1

    elapsed, prof = time_steps(one_step, args.steps, args.warmup, _lib, torch)
    steps_per_s = args.steps / elapsed
    vv, aa = s.state()
    e1 = g.energy(*vv, 1)
    assert math.isfinite(e1), "energy is not finite"
    log(f"{args.steps} steps in {elapsed:.4f} s -> {steps_per_s:.3f} steps/s; <u^2> = {e1:.6f}")
    passes, dom, achieved, _, avg_ms = pass_table(prof, elapsed)
    # per substep: 12 c2r (u, w, b, j) and 9 r2c (three products), the update
    # reads N (6), C (6), v (6), f (3) and writes 6 fields
    s_r, s_c = nx * ny * nz * 8.0, (nx // 2 + 1) * ny * nz * 16.0
    b_step = 2 * (21 * (s_r + s_c) + 27 * s_c)
    out = {
        "metric": "MHD3D 512x512x256 FP64 steps/s",
        "value": round(steps_per_s, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded random-phase k^4 exp velocity and vector potential, shell-2..3 forcing)",
        "config": {"workload": f"MHD3D {nx}x{ny}x{nz} FP64, Lx=Ly=2pi, Lz=pi, RK2, dt={MHD_DT}, "
                               f"nu=eta={MHD_NU}, unit energy ratio, mechanical forcing f0={F0} in 2<=|k|<=3",
                   "grid": [nx, ny, nz], "ord": 2, "dt": MHD_DT, "nu": MHD_NU, "eta": MHD_NU},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "fft_passes": passes,
        "step_algorithmic_GB": round(b_step / 1e9, 3),
        "step_effective_GBps": round(b_step * steps_per_s / 1e9, 1),
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = cpu_threads(args)
        del s, vv, aa
        torch.cuda.empty_cache()
        O.set_threads(threads)
        og = O.Grid(nx, ny, nz, L=MHD_L, arbsize=True)
        v, _ = og.initialv_k4exp(K0, U0, IC_SEED)
        a, _ = og.initialv_k4exp(K0, U0, IC_SEED + 1)
        a = og.normalize(*a, U0, kin=0)
        f, _ = og.initialfv_random(KDN, KUP, F0, F_SEED)
        t0 = time.perf_counter()
        og.mhd_step(v, a, MHD_DT, MHD_NU, MHD_NU, f=f, ord=2)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / t, "unit": "steps/s", "cores": int(threads), "kind": "port",
                               "sample": f"one RK2 step of the MHD workload with the C/OpenMP oracle on "
                                         f"{threads} host threads ({t:.2f} s)"}
        log(f"cpu baseline: {out['cpu_baseline']['value']:.5f} steps/s")
    return out


ENS_SIZES, ENS_PER_SIZE, ENS_DT = (32, 48, 64, 96), 24, 2e-3


def ensemble_members(grids):
    """(size, k0, nu, v) of the 96 members, ICs generated on the GPU.  nu gives
    k_max eta = 1.5 with k_max = n/3, eta = (nu^3/eps)^(1/4) and eps = nu <|w|^2>
    of the IC (<u^2> = 1): nu = (1.5/k_max)^2 <|w|^2>^(1/2)."""
    import numpy as np
    rng = np.random.default_rng(4000)
    out, m = [], 0
    for n in ENS_SIZES:
        g = grids[n]
        for _ in range(ENS_PER_SIZE):
            k0 = float(rng.uniform(2.0, 4.0))
            v, _ = g.initialv_k4exp(k0, 1.0, 5000 + m)
            omega2 = g.energy(*v, 0)
            nu = (1.5 / (n / 3.0)) ** 2 * math.sqrt(omega2)
            out.append((n, k0, nu, v))
            m += 1
    return out


def run_ensemble(args, torch):
    from ghost_amd import _lib
    from ghost_amd.pseudospec import Grid
    from ghost_amd.solvers import Ensemble
    grids = {n: Grid(n) for n in ENS_SIZES}
    mem = ensemble_members(grids)
    e = Ensemble([m[0] for m in mem], [m[2] for m in mem], dt=ENS_DT, ord=2)
    for i, (n, k0, nu, v) in enumerate(mem):
        e.set_velocity(i, v)
    mem = [m[:3] for m in mem]  # drop the IC copies
    torch.cuda.empty_cache()
    e.autotune(steps=3)
    scheds = e.schedules()
    log(f"ensemble of {len(mem)} members, groups {e.groups()}, (schedule, z tile, y tile) {scheds}")

    def one_step():
        e.step(1)

    elapsed, prof = time_steps(one_step, args.steps, args.warmup, _lib, torch)
    e.sync()
    steps_per_s = args.steps / elapsed
    for i in (0, len(mem) - 1):
        v = e.velocity(i)
        en = grids[mem[i][0]].energy(*v, 1)
        assert math.isfinite(en), "energy is not finite"
    log(f"{args.steps} ensemble steps in {elapsed:.4f} s -> {steps_per_s:.3f} steps/s "
        f"({steps_per_s * len(mem):.1f} member-steps/s)")
    b_step = 0.0
    for n in ENS_SIZES:
        s_r, s_c = n ** 3 * 8.0, (n // 2 + 1) * n * n * 16.0
        b_step += ENS_PER_SIZE * (18 * (s_r + s_c) + 30 * s_c)
    out = {
        "metric": "HD3D ensemble (96 runs, 24 each of 32^3/48^3/64^3/96^3) FP64 RK2 ensemble steps/s",
        "value": round(steps_per_s, 4), "unit": "ensemble steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded k^4 exp ICs, k0 uniform in [2, 4])",
        "config": {"workload": "96 unforced HD3D runs, 24 each at 32^3, 48^3, 64^3, 96^3, RK2, "
                               f"dt={ENS_DT}, nu from k_max eta = 1.5, one launch per pass per size, "
                               "captured HIP graph, autotuned radix schedules and tile widths",
                   "members": len(mem), "sizes": list(ENS_SIZES),
                   "schedules": [{"size": n, "sched": sc, "tile_z": tz, "tile_y": ty}
                                 for (n, _), (sc, tz, ty) in zip(e.groups(), scheds)]},
        "member_steps_per_s": round(steps_per_s * len(mem), 1),
        "step_algorithmic_GB": round(b_step / 1e9, 3),
        "step_effective_GBps": round(b_step * steps_per_s / 1e9, 1),
        "roofline": {"bound": "hbm", "kernel": "ensemble step", "achieved": round(b_step * steps_per_s / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(b_step * steps_per_s / 1e9 / HBM_PEAK_GBS, 4), "traffic": None},
    }
    if not args.no_cpu_baseline:
        from oracle import oracle as O
        threads = cpu_threads(args)
        O.set_threads(threads)
        t_total = 0.0
        for n in ENS_SIZES:
            og = O.Grid(n)
            v, _ = og.initialv_k4exp(3.0, 1.0, 1)
            t0 = time.perf_counter()
            og.hd_step(v, ENS_DT, 1e-2, ord=2)
            t_total += ENS_PER_SIZE * (time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": 1.0 / t_total, "unit": "ensemble steps/s", "cores": int(threads),
                               "kind": "port",
                               "sample": "one RK2 step of one member per size with the C/OpenMP oracle, times "
                                         f"24 members per size (96 sequential runs), {threads} host threads"}
        log(f"cpu baseline: {out['cpu_baseline']['value']:.4f} ensemble steps/s")
    return out


# learned-closure LES: two 512^3 DNS of the HD workload (no tracers), one
# with other seeds that trains and calibrates the closure, one scored; the
# 128^3 LES runs 2000 steps of four DNS steps over the scored DNS's window
LES_N, LES_KC, LES_SPIN, LES_WINDOW, LES_RATIO = 128, 42, 4000, 8000, 4
LES_TOL = 0.05
LES_CHECK_KMIN = 4  # the run-time check skips the forced shells 2..3 and the shell below them


def mlp_kernel_time(torch, mlp, grad, delta, tau, reps=50):
    """Mean time of one k_mlp_fields launch: ghx_mlp_stress issued back to
    back through the C ABI on the stream the events are recorded on."""
    import ctypes as C
    from ghost_amd import _lib
    lib = _lib.load()
    n = grad[0].numel()
    gp = (C.c_void_p * 9)(*[t.data_ptr() for t in grad])
    tp = (C.c_void_p * 6)(*[t.data_ptr() for t in tau])
    stream = torch.cuda.current_stream()
    _lib.check(lib.ghx_mlp_set_stream(mlp._h, C.c_void_p(stream.cuda_stream)))
    g2, t2 = C.cast(gp, _lib._PP), C.cast(tp, _lib._PP)
    _lib.check(lib.ghx_mlp_stress(mlp._h, n, g2, float(delta), t2))
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        lib.ghx_mlp_stress(mlp._h, n, g2, float(delta), t2)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps * 1e-3


def run_les(args, torch):
    import time as _t
    import numpy as np
    from ghost_amd.les import Experiment
    N = args.n
    ex = Experiment(nfine=N, nles=LES_N, kc=LES_KC, nu=NU, dt=DT, spin=LES_SPIN,
                    snaps=[LES_SPIN - 1000, LES_SPIN - 500, LES_SPIN], window=LES_WINDOW, ratio=LES_RATIO,
                    samples=8, f0=F0, fstep=FSTEP, k0=K0, u0=U0, ic_seed=IC_SEED, f_seed=F_SEED,
                    cal_window=LES_WINDOW, train_seeds=(IC_SEED + 1, F_SEED + 1), check_kmin=LES_CHECK_KMIN,
                    log=log)
    ex.run_training_dns()
    ex.fit(steps=3000, batch=65536, lr=1e-3)
    # the spectral part: the network's residual shell transfer on the training
    # states as a per-shell viscosity (signed, down to -nu), calibrated a
    # posteriori on the training DNS; and the same method without a network
    t0 = _t.perf_counter()
    nu_fit = ex.train.spectral_viscosity(ex.mlp, floor=-NU)
    nuk = ex.calibrate(nu_fit, closure=True, iters=5)
    t_cal = _t.perf_counter() - t0
    nu_alone = ex.train.spectral_viscosity(None, floor=-NU)
    nuk_alone = ex.calibrate(nu_alone, closure=False, iters=5)
    if args.save_closure:
        np.savez(args.save_closure, params=ex.mlp.params().cpu().numpy(), norm=ex.mlp.norm(), nuk=np.asarray(nuk),
                 delta=np.pi / LES_KC, kc=LES_KC, nles=LES_N, nu=NU, dt=LES_RATIO * DT)
        log(f"closure written to {args.save_closure}")
    ex.run_scored_dns()
    with_c = ex.run_les(closure=True, nuk=nuk, label="closure (network + spectral part)", fallback_tol=LES_TOL)
    net_only = ex.run_les(closure=True, label="network alone")
    spec_only = ex.run_les(closure=False, nuk=nuk_alone, label="spectral part alone (no network)")
    without = ex.run_les(closure=False)
    # the closure's inference kernel alone on the LES grid, for its roofline
    import math as _m
    from ghost_amd import les as L
    grad = L.velocity_gradients(ex.small, ex.v_start)
    tau = [ex.small.rzeros() for _ in range(6)]
    t_mlp = mlp_kernel_time(torch, ex.mlp, grad, _m.pi / LES_KC, tau)
    npts = LES_N ** 3
    mlp_bytes = npts * (9 + 6) * 8.0
    mlp_flops = npts * 2.0 * (10 * 64 + 64 * 64 + 64 * 6)
    within = with_c["finite"] and not with_c["fallback"] and with_c["spectrum_error_mean"] <= LES_TOL
    e_c, e_n, e_s = (with_c["spectrum_error_mean"], net_only["spectrum_error_mean"],
                     spec_only["spectrum_error_mean"])
    if e_n <= LES_TOL:
        carried = f"the network: alone it reaches {e_n:.4f} (spectral part alone {e_s:.4f}, both {e_c:.4f})"
    else:
        carried = (f"the spectral part (a calibrated per-shell viscosity): the network alone reaches {e_n:.4f}, "
                   f"the spectral part alone {e_s:.4f}, both together {e_c:.4f}; the network's gain over the "
                   f"spectral part alone is {e_s - e_c:+.4f}")
    speedup = ex.t_dns / with_c["result_seconds"]
    setup = ex.t_train + t_cal
    log(f"speedup {speedup:.1f}x over the scored DNS window; spectrum error {with_c['spectrum_error_mean']:.4f} "
        f"(network alone {net_only['spectrum_error_mean']:.4f}, spectral part alone "
        f"{spec_only['spectrum_error_mean']:.4f}, no closure {without['spectrum_error_mean']:.4f}); closure kernel "
        f"{t_mlp * 1e3:.3f} ms, {mlp_bytes / t_mlp / 1e9:.0f} GB/s, {mlp_flops / t_mlp / 1e12:.1f} TFLOP/s")
    return {
        "metric": "learned-LES speedup at <=5% spectrum err (128^3 LES vs 512^3 FP64 DNS)",
        "value": round(speedup, 3), "unit": "x (DNS GPU-seconds / LES GPU-seconds over the scored window)",
        "n_gpus": 1, "steps": with_c["steps"], "warmup": 0,
        "ms_per_step": round(1e3 * with_c["seconds"] / with_c["steps"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (closure: bf16 operands, f32 accumulate)",
        "data": "synthetic (two 512^3 DNS of the HD workload generate the training states and the reference)",
        "config": {"workload": f"HD3D {N}^3 FP64 DNS (nu={NU}, dt={DT}, forcing f0={F0} in 2<=|k|<=3, "
                               f"phase shift every {FSTEP} steps); training DNS (seeds {IC_SEED + 1}, {F_SEED + 1}) "
                               f"spun up {LES_SPIN} steps: sharp filter kc={LES_KC} and truncation onto {LES_N}^3, "
                               f"10-64-64-6 GELU closure trained on its filtered states at steps {LES_SPIN - 1000}, "
                               f"{LES_SPIN - 500}, {LES_SPIN}, a per-shell spectral part fitted from the network's "
                               f"residual subgrid transfer and calibrated a posteriori on its next {LES_WINDOW} steps; "
                               f"scored DNS (seeds {IC_SEED}, {F_SEED}) spun up {LES_SPIN} steps, then LES "
                               f"{with_c['steps']} steps of dt={LES_RATIO * DT} against its next {LES_WINDOW} steps",
                   "grid_dns": N, "grid_les": LES_N, "kc": LES_KC, "les_steps": with_c["steps"]},
        "within_5pct": bool(within),
        "fallback": with_c["fallback"],
        "closure_carried_by": carried,
        "les_closure": with_c, "les_network_alone": net_only, "les_spectral_alone": spec_only,
        "les_no_closure": without,
        "spectral_part": {"nu_fit_k10_20_30_40": [float(nu_fit[k - 1]) for k in (10, 20, 30, 40)],
                          "nu_calibrated_k10_20_30_40": [float(nuk[k - 1]) for k in (10, 20, 30, 40)],
                          "nu_alone_calibrated_k10_20_30_40": [float(nuk_alone[k - 1]) for k in (10, 20, 30, 40)]},
        "dns_window_seconds": round(ex.t_dns, 3), "dns_spinup_seconds": round(ex.t_spin, 3),
        "training_dns_seconds": round(ex.t_train_dns, 3),
        "speedup_with_training_and_calibration": round(ex.t_dns / (with_c["result_seconds"] + setup), 3),
        "speedup_with_training_dns_training_and_calibration":
            round(ex.t_dns / (with_c["result_seconds"] + setup + ex.t_train_dns), 3),
        "spectrum_dns": [float(x) for x in np.mean(ex.ref, axis=0)[:LES_KC]],
        "spectrum_training_dns": [float(x) for x in ex.ref_runtime[:LES_KC]] if ex.ref_runtime is not None else None,
        "training": {"points": ex.n_train, "steps": len(ex.losses), "seconds": round(ex.t_train, 3),
                     "loss_first": float(ex.losses[0]), "loss_last": float(ex.losses[-1]),
                     "dissipation_scale": round(float(ex.scale), 4)},
        "calibration_seconds": round(t_cal, 3),
        "roofline": {"bound": "hbm", "kernel": "closure inference (k_mlp_fields)",
                     "achieved": round(mlp_bytes / t_mlp / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(mlp_bytes / t_mlp / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_ms": round(t_mlp * 1e3, 4),
                     "mfma_tflops": round(mlp_flops / t_mlp / 1e12, 2), "mfma_peak_tflops_bf16": 2500.0},
    }


def cpu_threads(args):
    """Threads for the CPU baseline: --cpu-threads, else 16, the host share of
    one GPU on the measurement machines (os.cpu_count() reports the whole
    host)."""
    return args.cpu_threads or min(16, os.cpu_count() or 1)


def cpu_baseline(n, threads, npart):
    """Time one whole RK2 step of the same workload (two substeps, each
    followed by a tracer stage) with the CPU oracle (test infrastructure)."""
    from oracle import oracle as O
    O.set_threads(threads)
    og = O.Grid(n)
    v, _ = og.initialv_k4exp(K0, U0, IC_SEED)
    f, _ = og.initialfv_random(KDN, KUP, F0, F_SEED)
    t0 = time.perf_counter()
    if npart:
        pos = tracer_positions(n, npart)
        og.hd_step_tracers(v, pos, DT, NU, f=f, ord=2, zcompat=True)
    else:
        og.hd_step(v, DT, NU, f=f, ord=2)
    t = time.perf_counter() - t0
    return {"value": 1.0 / t, "unit": "steps/s", "cores": int(threads), "kind": "port",
            "sample": f"one whole RK2 step of the {n}^3 workload ({npart} tracers) with the C/OpenMP oracle on "
                      f"{threads} host threads (os.cpu_count() = {os.cpu_count()}): {t:.2f} s"}


SLAB_SHAPES = ((512, 512, 64), (64, 512, 512))


def run_slab(args, torch):
    from ghost_amd.pseudospec import Grid
    res = {}
    for shape in SLAB_SHAPES:
        g = Grid(*shape)
        r = torch.randn(g.rshape, dtype=torch.float64, device="cuda")
        c = g.r2c(r)
        out_r = torch.empty_like(r)
        out_c = torch.empty_like(c)
        s_r, s_c = r.numel() * 8.0, c.numel() * 16.0
        times = {}
        for name, fn in (("r2c", lambda: g.r2c(r, out_c)), ("c2r", lambda: g.c2r(c, out_r))):
            for _ in range(args.warmup):
                fn()
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(args.steps):
                fn()
            b.record()
            torch.cuda.synchronize()
            ms = a.elapsed_time(b) / args.steps
            times[name] = {"ms": round(ms, 4), "GBps_on_Sr_plus_Sc": round((s_r + s_c) / ms / 1e6, 1)}
        key = "x".join(map(str, shape))
        res[key] = times
        log(f"slab {key}: {times}")
        del g, r, c, out_r, out_c
        torch.cuda.empty_cache()
    return {
        "metric": "per-GPU slab share of the 512^3 FP64 transform (8-way decomposition), ms per transform",
        "value": round(sum(t["ms"] for v in res.values() for t in v.values()) / 2, 4), "unit": "ms (r2c + c2r pair, both shapes, averaged)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (standard normal real fields)",
        "config": {"workload": "512x512x64 (y and x passes on one GPU's 64 z-planes) and 64x512x512 (z pass on "
                               "33 kx columns) FP64 r2c and c2r; the all-to-all is not built",
                   "shapes": [list(s) for s in SLAB_SHAPES]},
        "shapes": res,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--tracers", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", choices=("hd", "mhd", "ensemble", "les", "slab"), default="hd")
    ap.add_argument("--save-closure", default="",
                    help="with --workload les: write the trained closure (network parameters, input/output "
                         "normalisation, calibrated spectral part, filter width) to this .npz file")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (a kernel of known traffic for checking the "
                         "rocprofv3 read and write counters) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        log("bench.py runs on exactly one GPU: this build has no multi-GPU path (one MI355X per run); "
            "use --gpus 1")
        sys.exit(2)

    import torch

    import ghost_amd
    from ghost_amd import _lib
    from ghost_amd.lagpart import GPart
    from ghost_amd.pseudospec import Grid, randu
    from ghost_amd.solvers import HDSolver

    if not torch.cuda.is_available():
        log("no GPU visible")
        sys.exit(1)
    ghost_amd.load()
    torch.cuda.set_device(0)
    n = args.n

    if args.calibrate_copy:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("copied 1 GiB device to device")
        return
    if args.workload == "mhd":
        print(json.dumps(run_mhd(args, torch)), flush=True)
        return
    if args.workload == "ensemble":
        print(json.dumps(run_ensemble(args, torch)), flush=True)
        return
    if args.workload == "les":
        print(json.dumps(run_les(args, torch)), flush=True)
        return
    if args.workload == "slab":
        print(json.dumps(run_slab(args, torch)), flush=True)
        return

    g = Grid(n)
    s = HDSolver(g, dt=DT, nu=NU, ord=2)
    v, _ = g.initialv_k4exp(K0, U0, IC_SEED)
    f, fseed = g.initialfv_random(KDN, KUP, F0, F_SEED)
    s.set_velocity(v)
    s.set_forcing(f)
    e0 = g.energy(*v, 1)
    del v, f
    npart = max(0, args.tracers)
    parts = None
    if npart:
        parts = GPart(g, npart, zcompat=True)
        parts.set_positions(*[torch.from_numpy(p).cuda() for p in tracer_positions(n, npart)])
        s.set_tracers(parts)
    torch.cuda.empty_cache()
    log(f"grid {n}^3, <u^2> = {e0:.6f} (E = {0.5 * e0:.6f}), {npart} tracers")

    state = {"timef": FSTEP, "seed": fseed}

    def one_step():
        # main3D.fpp:2544-2556: phase shift of the forcing when timef == fstep
        if state["timef"] == FSTEP:
            state["timef"] = 0
            r, state["seed"] = randu(state["seed"])
            phase = 2 * math.pi * r
            s.phaseshift_forcing(complex(math.cos(phase), math.sin(phase)))
        s.step(1)
        state["timef"] += 1

    elapsed, prof = time_steps(one_step, args.steps, args.warmup, _lib, torch)
    steps_per_s = args.steps / elapsed
    v = s.velocity()
    e1 = g.energy(*v, 1)
    assert math.isfinite(e1), "energy is not finite"
    if parts is not None:
        for c, p in enumerate(parts.positions()):
            assert bool(((p >= 0) & (p < n)).all()), "tracer outside the box"
    log(f"{args.steps} steps in {elapsed:.4f} s -> {steps_per_s:.3f} steps/s; <u^2> = {e1:.6f}")

    passes, dom, achieved, members, avg_ms = pass_table(prof, elapsed)
    # free the solver's fields (about 23 GB at 512^3) before the yardstick runs
    del v, s
    torch.cuda.empty_cache()
    bw = stream_bandwidth(torch)
    copy_gbs = max(bw["copy"], bw["copy_nt"], bw["copy_chunk"], bw["copy_chunk_nt"])
    log(f"yardstick kernels: {bw}; dominant kernel {dom} at {achieved:.1f} GB/s")
    rates = transform_rates(torch, n)
    log(f"plain transforms at {n}^3: {rates}")
    # HBM traffic per launch of the dominant kernel, pooled over its profiler slots
    tr = [(read_traffic(k, n, npart), prof[k][0]) for k in members]
    traffic = (int(round(sum(t * c for t, c in tr) / sum(c for _, c in tr)))
               if all(t is not None for t, _ in tr) else None)

    # algorithmic bytes of a step (18 transforms read once + written once, and the
    # 30 spectral field reads/writes of the two updates), plus per tracer stage
    # 3 c2r and per particle 3 x 64 control points, position in/out and x0
    s_r, s_c = n ** 3 * 8.0, (n // 2 + 1) * n * n * 16.0
    b_step = 18 * (s_r + s_c) + 30 * s_c
    if npart:
        b_step += 2 * (3 * (s_r + s_c) + npart * (3 * 64 * 8 + 24 + 24 + 24))

    out = {
        "metric": METRIC,
        "value": round(steps_per_s, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random-phase k^4 exp IC, shell-2..3 forcing and uniform tracer positions generated on the host)",
        "config": {"workload": f"HD3D {n}^3 FP64 forced turbulence, RK2, dt={DT}, nu={NU}, k0={K0}, "
                               f"E0=0.5, forcing f0={F0} in 2<=|k|<=3 with phase shift every {FSTEP} steps; "
                               f"{npart} Lagrangian tracers, cubic-spline interpolation, GHOST z stencil",
                   "grid": [n, n, n], "ord": 2, "dt": DT, "nu": NU, "tracers": npart},
        "roofline": {"bound": "hbm", "kernel": kernel_name(dom, n),
                     "passes": members, "avg_ms": round(avg_ms, 4), "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_unit": "HBM bytes per launch (rocprofv3: 32-, 64- and 128-byte "
                                                         "TCC read requests, WRITE_SIZE)",
                     "algorithmic_bytes": int(round(achieved * 1e9 * avg_ms * 1e-3)),
                     "achievable_copy_GBps": round(copy_gbs, 1),
                     "frac_of_copy": round(achieved / copy_gbs, 4),
                     "yardstick_GBps": bw},
        "transforms": rates,
        "fft_passes": passes,
        "step_algorithmic_GB": round(b_step / 1e9, 3),
        "step_effective_GBps": round(b_step * steps_per_s / 1e9, 1),
    }
    if not args.no_cpu_baseline:
        threads = cpu_threads(args)
        log(f"cpu baseline: one oracle RK2 step at {n}^3 on {threads} threads ...")
        del parts
        torch.cuda.empty_cache()
        out["cpu_baseline"] = cpu_baseline(n, threads, npart)
        log(f"cpu baseline: {out['cpu_baseline']['value']:.5f} steps/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
