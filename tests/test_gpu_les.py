"""Learned-closure LES on the GPU: the training pairs (filtered state and
subgrid stress) and the velocity gradients against the CPU oracle's sharp
filter, truncation and transforms; the closure's forcing inside the HD step
against the oracle's step with that forcing; and an a-posteriori run where
the trained closure brings the LES spectrum closer to the filtered DNS than
the same LES without it."""
import math

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


def to_np(t):
    return t.detach().cpu().numpy()


def test_sgs_targets_and_gradients_match_oracle(gpu):
    from ghost_amd import les
    from ghost_amd.pseudospec import Grid
    from oracle import oracle as O
    n, m, kc = 32, 16, 5.0
    fine, small = Grid(n), Grid(m)
    of, os_ = O.Grid(n), O.Grid(m)
    v, _ = of.initialv_k4exp(3.0, 1.0, 61)
    tau, vs = les.sgs_targets(fine, small, [to_gpu(x) for x in v], kc)
    # oracle: ub = sharp(v), products of the physical fields, filtered and truncated
    ub = [of.sharp(x, math.pi / kc) for x in v]
    N3 = float(n ** 3)
    u = [of.c2r(x) / N3 for x in v]
    b = [of.c2r(x) / N3 for x in ub]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for c, (i, j) in enumerate(pairs):
        p = u[i] * u[j] - b[i] * b[j]
        P = of.sharp(of.r2c(p), math.pi / kc)
        want = os_.c2r(of.truncate(P, os_))
        assert rel(to_np(tau[c]), want) < 1e-12, c
    for k in range(3):
        want = of.truncate(ub[k], os_) * float(m ** 3)
        assert rel(to_np(vs[k]), want) < 1e-13
    grad = les.velocity_gradients(small, vs)
    vsn = [to_np(x) for x in vs]
    for i in range(3):
        for j in range(3):
            want = os_.c2r(os_.derivk3(vsn[i], j + 1)) / float(m ** 3)
            assert rel(to_np(grad[3 * i + j]), want) < 1e-13, (i, j)


def test_closure_forcing_in_the_step_matches_oracle(gpu):
    """ord 1: v' = v + dt (nu lap v + N(v) + f + P(-div tau)), tau from the
    network on grad v; the oracle step receives f + P(-div tau) as its forcing."""
    import torch
    from ghost_amd import les
    from ghost_amd.closure import MLP
    from ghost_amd.pseudospec import Grid
    from ghost_amd.solvers import HDSolver
    from oracle import oracle as O
    n, dt, nu, delta = 32, 2e-3, 5e-3, 0.3
    g, og = Grid(n), O.Grid(n)
    v, _ = og.initialv_k4exp(3.0, 1.0, 71)
    f, _ = og.initialfv_random(2.0, 3.0, 0.5, 72)
    m = MLP(5)
    grad = les.velocity_gradients(g, [to_gpu(x) for x in v])
    tau6 = [g.rzeros() for _ in range(6)]
    m.fit_norm(grad, delta, [x * 0.01 + 1e-3 * (c + 1) for c, x in enumerate(grad[:6])])
    m.stress(grad, delta, tau6)
    tau = [to_np(t) for t in tau6]
    s = HDSolver(g, dt=dt, nu=nu, ord=1)
    s.set_velocity([to_gpu(x) for x in v])
    s.set_forcing([to_gpu(x) for x in f])
    s.set_sgs(m, delta)
    s.step(1)
    got = [to_np(x) for x in s.velocity()]
    # the oracle's forcing: f + P(-i k_j tau_ij) on the kept modes
    T = [og.r2c(t) for t in tau]
    Tm = [[T[0], T[3], T[4]], [T[3], T[1], T[5]], [T[4], T[5], T[2]]]
    K = np.meshgrid(np.r_[0:n // 2 + 1] * 1.0, np.fft.fftfreq(n, 1.0 / n), np.fft.fftfreq(n, 1.0 / n), indexing="ij")
    K[0][n // 2] = -n // 2  # GHOST's Nyquist kx carries a negative sign
    kk2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    sgs = [-1j * (K[0] * Tm[i][0] + K[1] * Tm[i][1] + K[2] * Tm[i][2]) for i in range(3)]
    dot = K[0] * sgs[0] + K[1] * sgs[1] + K[2] * sgs[2]
    kept = (kk2 <= n * n / 9.0) & (kk2 > 0)
    ftot = []
    for i in range(3):
        proj = np.where(kept, sgs[i] - K[i] * dot / np.where(kk2 > 0, kk2, 1.0), 0.0)
        ftot.append(f[i] + proj)
    ref = [x.copy() for x in v]
    og.hd_step(ref, dt, nu, f=ftot, ord=1)
    for k in range(3):
        assert rel(got[k], ref[k]) < 1e-12
    # and the closure changed the step
    plain = [x.copy() for x in v]
    og.hd_step(plain, dt, nu, f=f, ord=1)
    assert rel(got[0], plain[0]) > 1e-8


def test_closure_steps_pruned_equal_full(gpu):
    """After the first step the LES state is dealiased, and the velocity
    gradients the network reads skip the zero modes (pruned passes); the
    steps must equal the unpruned ones (GHX_NO_PRUNE) to round-off."""
    import os
    from ghost_amd import les
    from ghost_amd.closure import MLP
    from ghost_amd.pseudospec import Grid
    from ghost_amd.solvers import HDSolver
    from oracle import oracle as O
    n, dt, nu, delta = 48, 2e-3, 5e-3, 0.3
    g, og = Grid(n), O.Grid(n)
    v, _ = og.initialv_k4exp(3.0, 1.0, 73)
    f, _ = og.initialfv_random(2.0, 3.0, 0.5, 74)
    m = MLP(6)
    grad = les.velocity_gradients(g, [to_gpu(x) for x in v])
    m.fit_norm(grad, delta, [x * 0.01 + 1e-3 * (c + 1) for c, x in enumerate(grad[:6])])
    outs = []
    for env in (None, "1"):
        if env:
            os.environ["GHX_NO_PRUNE"] = env
        try:
            s = HDSolver(g, dt=dt, nu=nu, ord=2)
            s.set_velocity([to_gpu(x) for x in v])
            s.set_forcing([to_gpu(x) for x in f])
            s.set_sgs(m, delta)
            s.step(3)
            outs.append([to_np(x) for x in s.velocity()])
        finally:
            os.environ.pop("GHX_NO_PRUNE", None)
    for k in range(3):
        assert rel(outs[0][k], outs[1][k]) < 1e-14, k


def test_spectral_part_in_the_step_matches_oracle(gpu):
    """The closure's spectral part: -nu(shell) kk2 v_in joins the forcing of
    each substep (ord 2, so v_in differs between the substeps); with a
    network attached both terms add.  The oracle takes the substeps one by
    one with that forcing computed from each substep's input."""
    from ghost_amd.pseudospec import Grid
    from ghost_amd.solvers import HDSolver
    from oracle import oracle as O
    n, dt, nu = 32, 2e-3, 5e-3
    g, og = Grid(n), O.Grid(n)
    v, _ = og.initialv_k4exp(3.0, 1.0, 81)
    f, _ = og.initialfv_random(2.0, 3.0, 0.5, 82)
    nb = n // 2 + 1
    nuk = np.linspace(-2e-3, 8e-3, nb)  # signed, as the fitted residual can be
    s = HDSolver(g, dt=dt, nu=nu, ord=2)
    s.set_velocity([to_gpu(x) for x in v])
    s.set_forcing([to_gpu(x) for x in f])
    s.set_sgs_spectral(nuk)
    s.step(2)
    got = [to_np(x) for x in s.velocity()]
    K = np.meshgrid(np.r_[0:n // 2 + 1] * 1.0, np.fft.fftfreq(n, 1.0 / n), np.fft.fftfreq(n, 1.0 / n), indexing="ij")
    K[0][n // 2] = -n // 2
    kk2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    shell = (np.sqrt(kk2) + 0.501).astype(int)
    kept = (kk2 <= n * n / 9.0) & (kk2 > 0)
    coef = np.where(kept & (shell >= 1) & (shell <= nb), -nuk[np.clip(shell - 1, 0, nb - 1)] * kk2, 0.0)
    ref = [x.copy() for x in v]
    for _ in range(2):
        c = [x.copy() for x in ref]
        for o in (2, 1):
            fo = [f[i] + coef * ref[i] for i in range(3)]
            og.hd_substep(ref, c, o, dt, nu, f=fo)
    for k in range(3):
        assert rel(got[k], ref[k]) < 1e-12
    # removing it restores the plain step
    s2 = HDSolver(g, dt=dt, nu=nu, ord=2)
    s2.set_velocity([to_gpu(x) for x in v])
    s2.set_forcing([to_gpu(x) for x in f])
    s2.set_sgs_spectral(nuk)
    s2.set_sgs_spectral(None)
    s2.step(1)
    plain = [x.copy() for x in v]
    og.hd_step(plain, dt, nu, f=f, ord=2)
    for k in range(3):
        assert rel(to_np(s2.velocity()[k]), plain[k]) < 1e-12


def test_trained_closure_beats_no_closure(gpu):
    """128^3 DNS (nu = 5e-3, resolved) filtered at kc = 10 onto a 32^3 LES:
    the closure trained on three filtered DNS states brings the time-averaged
    LES spectrum closer to the filtered DNS than the same LES without it."""
    from ghost_amd.les import Experiment
    ex = Experiment(nfine=128, nles=32, kc=10, nu=5e-3, dt=2e-3, spin=1000, snaps=[800, 900, 1000], window=480,
                    ratio=2, samples=8, log=lambda *a: None)
    ex.run_dns()
    ex.fit(steps=2000, batch=16384, lr=1e-3)
    with_c = ex.run_les(closure=True)
    without = ex.run_les(closure=False)
    assert with_c["finite"] and without["finite"]
    assert ex.losses[-1] < 0.9 * ex.losses[0]
    assert with_c["spectrum_error_mean"] < 0.8 * without["spectrum_error_mean"]


CLOSURE_FIXTURE = "golden/les_closure_128_kc42.npz"


def test_trained_closure_substep_at_128_kc42_matches_oracle(gpu):
    """The LES of the learned-closure workload at its size: 128^3, filter
    width pi/42, the network trained on the filtered 512^3 DNS and the
    calibrated spectral part both attached (tests/golden/les_closure_128_kc42.npz,
    written by bench.py --workload les --save-closure).  One ord-1 step with
    the workload's nu and LES time step against the oracle's step whose
    forcing is f + P(-div tau) - nu(shell) kk2 v, tau taken from the network
    on the GPU's velocity gradients.  Relative L2 error 1e-12."""
    import os
    from ghost_amd import les
    from ghost_amd.closure import MLP
    from ghost_amd.pseudospec import Grid
    from ghost_amd.solvers import HDSolver
    from oracle import oracle as O
    fx = np.load(os.path.join(os.path.dirname(__file__), CLOSURE_FIXTURE))
    n, kc = int(fx["nles"]), int(fx["kc"])
    delta, nu, dt, nuk = float(fx["delta"]), float(fx["nu"]), float(fx["dt"]), np.asarray(fx["nuk"])
    assert (n, kc) == (128, 42) and abs(delta - math.pi / 42) < 1e-15
    g, og = Grid(n), O.Grid(n)
    v, _ = og.initialv_k4exp(3.0, 1.0, 91)
    f, _ = og.initialfv_random(2.0, 3.0, 0.4, 92)
    m = MLP(0)
    m.set_params(fx["params"])
    m.set_norm(fx["norm"])
    grad = les.velocity_gradients(g, [to_gpu(x) for x in v])
    tau6 = [g.rzeros() for _ in range(6)]
    m.stress(grad, delta, tau6)
    tau = [to_np(t) for t in tau6]
    assert all(np.all(np.isfinite(t)) for t in tau) and max(np.abs(t).max() for t in tau) > 0
    s = HDSolver(g, dt=dt, nu=nu, ord=1)
    s.set_velocity([to_gpu(x) for x in v])
    s.set_forcing([to_gpu(x) for x in f])
    s.set_sgs(m, delta)
    s.set_sgs_spectral(nuk)
    s.step(1)
    got = [to_np(x) for x in s.velocity()]
    T = [og.r2c(t) for t in tau]
    Tm = [[T[0], T[3], T[4]], [T[3], T[1], T[5]], [T[4], T[5], T[2]]]
    K = np.meshgrid(np.r_[0:n // 2 + 1] * 1.0, np.fft.fftfreq(n, 1.0 / n), np.fft.fftfreq(n, 1.0 / n), indexing="ij")
    K[0][n // 2] = -n // 2  # GHOST's Nyquist kx carries a negative sign
    kk2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    kept = (kk2 <= n * n / 9.0) & (kk2 > 0)
    nb = nuk.size
    shell = (np.sqrt(kk2) + 0.501).astype(int)
    coef = np.where(kept & (shell >= 1) & (shell <= nb), -nuk[np.clip(shell - 1, 0, nb - 1)] * kk2, 0.0)
    sgs = [-1j * (K[0] * Tm[i][0] + K[1] * Tm[i][1] + K[2] * Tm[i][2]) for i in range(3)]
    dot = K[0] * sgs[0] + K[1] * sgs[1] + K[2] * sgs[2]
    ftot = []
    for i in range(3):
        proj = np.where(kept, sgs[i] - K[i] * dot / np.where(kk2 > 0, kk2, 1.0), 0.0)
        ftot.append(f[i] + proj + coef * v[i])
    ref = [x.copy() for x in v]
    og.hd_step(ref, dt, nu, f=ftot, ord=1)
    for k in range(3):
        assert rel(got[k], ref[k]) < 1e-12
    # both parts of the closure act: the step differs from the plain one
    plain = [x.copy() for x in v]
    og.hd_step(plain, dt, nu, f=f, ord=1)
    assert rel(got[0], plain[0]) > 1e-9
