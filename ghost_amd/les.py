"""Learned-closure LES on the gfx950 path (no counterpart in GHOST): training
pairs from sharp-spectrally filtered DNS states that the GPU DNS generates,
training of the closure network on the GPU, and the LES itself, an HD solver
on the coarse grid whose forcing gains the projected divergence of the
network's subgrid stress before every substep.

The filter is GHOST's sharp cut-off (pseudospec3D_filt.f90:200-241,
delta = pi/kc) followed by trunc3D's truncation onto the LES grid
(analysis/trunc3D.fpp:190-320).
"""
import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .closure import MLP
from .pseudospec import Grid


def _pp(ts):
    arr = (C.c_void_p * len(ts))(*[t.data_ptr() for t in ts])
    return C.cast(arr, _lib._PP), arr


def velocity_gradients(grid: Grid, v):
    """Nine real fields d u_i / d x_j (i major) of the spectral state v."""
    out = [grid.rzeros() for _ in range(9)]
    vp, ka = _pp([x.contiguous() for x in v])
    gp, kb = _pp(out)
    torch.cuda.synchronize()
    _lib.check(_lib.load().ghx_velocity_gradients(grid._h, vp, gp))
    torch.cuda.synchronize()
    return out


def sgs_targets(fine: Grid, les: Grid, v, kc):
    """(tau6, vsmall): the physical subgrid stress on the LES grid's points
    (xx, yy, zz, xy, xz, yz) and the filtered state in the LES grid's
    normalisation, from the fine-grid spectral state v."""
    tau = [les.rzeros() for _ in range(6)]
    vs = [les.czeros() for _ in range(3)]
    vp, ka = _pp([x.contiguous() for x in v])
    tp, kb = _pp(tau)
    sp, kc_ = _pp(vs)
    torch.cuda.synchronize()
    _lib.check(_lib.load().ghx_sgs_targets(fine._h, les._h, vp, float(kc), tp, sp))
    torch.cuda.synchronize()
    return tau, vs


def filter_forcing(fine: Grid, les: Grid, f, kc):
    """The fine-grid forcing carried to the LES grid: sharp(kc), truncated,
    in the LES grid's normalisation."""
    out = []
    scale = float(les.nx * les.ny * les.nz)
    for x in f:
        b = fine.czeros()
        fine.sharp(x, b, math.pi / kc)
        t = les.czeros()
        fine.truncate(b, les, out=t)
        out.append(t * scale)
    return out


def spectrum(grid: Grid, v):
    """E(k) of shells k = 1..nmax/2+1 at indices k - 1 (GHOST's spectrumc, kin = 1)"""
    e, _ = grid.spectrumc(*v, kin=1)
    return np.asarray(e)


def spectrum_error(e, eref, kmax):
    """Mean over the shells k = 1..kmax of |E(k) - E_ref(k)| / E_ref(k)."""
    i = np.arange(0, int(kmax))
    return float(np.mean(np.abs(e[i] - eref[i]) / eref[i]))


def spectrum_error_band(e, eref, kmin, kmax):
    """Mean over the shells k = kmin..kmax of |E(k) - E_ref(k)| / E_ref(k)."""
    i = np.arange(int(kmin) - 1, int(kmax))
    return float(np.mean(np.abs(e[i] - eref[i]) / eref[i]))


def dissipation(grad, tau):
    """Mean subgrid dissipation -<tau_ij S_ij> with S_ij = (d_j u_i + d_i u_j)/2
    (stress components xx, yy, zz, xy, xz, yz; gradients i major)."""
    s = lambda i, j: 0.5 * (grad[3 * i + j] + grad[3 * j + i])
    pairs = [(0, 0, 1.0), (1, 1, 1.0), (2, 2, 1.0), (0, 1, 2.0), (0, 2, 2.0), (1, 2, 2.0)]
    tot = 0.0
    for c, (i, j, w) in enumerate(pairs):
        tot += w * float(torch.sum(tau[c] * s(i, j)))
    return -tot / grad[0].numel()


class TrainingSet:
    """Normalised features and targets of the closure on the LES grid,
    accumulated from several filtered DNS states."""

    def __init__(self, les: Grid, kc):
        self.les, self.kc, self.delta = les, kc, math.pi / kc
        self.snaps = []  # (grad9, tau6) per state
        self.vels = []  # the filtered states themselves (LES grid, spectral)

    def add(self, fine: Grid, v):
        tau, vs = sgs_targets(fine, self.les, v, self.kc)
        grad = velocity_gradients(self.les, vs)
        self.snaps.append((grad, tau))
        self.vels.append(vs)
        return vs

    def force(self, tau):
        """P(-div tau) on the LES grid, in the solver's spectral units, from
        the six physical stress components (xx, yy, zz, xy, xz, yz)."""
        g = self.les
        th = [g.r2c(t) for t in tau]
        rows = ((0, 3, 4), (3, 1, 5), (4, 5, 2))  # tau_xj, tau_yj, tau_zj
        d = []
        for r in rows:
            acc = g.czeros()
            for dirn, c in enumerate(r, start=1):
                acc += g.derivk3(th[c], g.czeros(), dirn)
            d.append(acc)
        return [g.nonlhd3(*d, g.czeros(), dirn) for dirn in (1, 2, 3)]  # -P(div tau)

    def transfers(self, mlp=None):
        """Shell sums over the training states of the energy the exact subgrid
        force and the network's force put into each shell (entrans of the
        filtered velocity with each force), and of kk2 |v|^2."""
        g = self.les
        te = tm = dk = 0.0
        for (grad, tau), vs in zip(self.snaps, self.vels):
            te = te + np.array(g.entrans(*vs, *self.force(tau), kin=1))
            if mlp is not None:
                p = [torch.empty_like(x) for x in tau]
                mlp.stress(grad, self.delta, p)
                tm = tm + np.array(g.entrans(*vs, *self.force(p), kin=1))
            dk = dk + np.array(g.entrans(*vs, *vs, kin=0))
        return te, tm, dk

    def spectral_viscosity(self, mlp=None, floor=0.0):
        """Spectral part of the closure, fitted on the training states: the
        viscosity per shell nu(k) = (T_net(k) - T_exact(k)) / sum kk2 |v|^2
        that makes the network's shell transfer (zero without one) equal the
        exact subgrid transfer; clipped below at floor, and carried from the
        last shell with data to the shells past kc."""
        te, tm, dk = self.transfers(mlp)
        diff = (tm - te) if mlp is not None else -te
        nu = np.zeros_like(dk)
        has = dk > 0
        nu[has] = diff[has] / dk[has]
        nu = np.maximum(nu, floor)
        last = int(np.nonzero(has)[0][-1]) if np.any(has) else -1
        if 0 <= last < len(nu) - 1:
            nu[last + 1:] = nu[last]
        return nu

    def arrays(self, mlp: MLP):
        """Fit the network's normalisation on all states (the pooled moments of
        the per-state fits), then pack every state: (X, T) as n x 10 and n x 6
        FP32 device arrays."""
        mu, m2 = [], []
        for g, t in self.snaps:
            mlp.fit_norm(g, self.delta, t)
            nm = mlp.norm().astype(np.float64)
            mean = np.concatenate([nm[0:10], nm[20:26]])
            sd = np.concatenate([nm[10:20], nm[26:32]])
            mu.append(mean)
            m2.append(sd * sd + mean * mean)
        mean = np.mean(mu, axis=0)
        sd = np.sqrt(np.maximum(np.mean(m2, axis=0) - mean * mean, 0.0))
        sd[sd == 0] = 1.0
        sd[9] = 1.0  # the filter-width input: one value, centred
        mlp.set_norm(np.concatenate([mean[0:10], sd[0:10], mean[10:16], sd[10:16]]).astype(np.float32))
        xs, ts = [], []
        for g, t in self.snaps:
            X, T = mlp.pack(g, self.delta, t)
            xs.append(X)
            ts.append(T)
        return torch.cat(xs), torch.cat(ts)

    def match_dissipation(self, mlp: MLP):
        """Scale the network's outputs so that its mean subgrid dissipation on
        the training states equals the filtered DNS's: the regression towards
        the mean that a weakly correlated target imposes on the fit shrinks
        the predicted stress, and with it the energy drained at the cut-off.
        Returns the factor applied."""
        true_d = pred_d = 0.0
        for g, t in self.snaps:
            p = [torch.empty_like(x) for x in t]
            mlp.stress(g, self.delta, p)
            true_d += dissipation(g, t)
            pred_d += dissipation(g, p)
        if not (pred_d > 0.0 and true_d > 0.0):
            return 1.0
        a = true_d / pred_d
        nm = mlp.norm()
        nm[20:26] *= a
        nm[26:32] *= a
        mlp.set_norm(nm)
        return a


class Experiment:
    """The a-posteriori comparison.  Two forced DNS on the fine grid (GHOST's
    rand=1 forcing, phase shift every fstep steps) with different seeds: the
    training DNS spins up and leaves the filtered training states and the
    calibration spectra behind; the scored DNS spins up, then runs a window
    while its filtered spectrum is sampled.  LES run on the coarse grid over
    the scored window from its filtered state, with the trained closure and
    without it, with the scored DNS's forcing and phase schedule and ratio
    times the DNS time step.  Errors are the mean over shells 1..kc of the
    relative error of the time-averaged spectra (and of the spectra at the end
    of the window) against the scored DNS, which neither the training nor the
    calibration sees.  The run-time check of the fallback compares the LES
    with the training DNS's filtered spectrum instead (the scored DNS is not
    known to an LES that replaces it)."""

    def __init__(self, nfine, nles, kc, nu, dt, spin, snaps, window, ratio=2, f0=0.4, fstep=20, k0=3.0, u0=1.0,
                 ic_seed=1000, f_seed=2000, samples=8, cal_window=0, train_seeds=None, check_kmin=4, log=print):
        if train_seeds is None:
            train_seeds = (ic_seed + 1, f_seed + 1)
        self.__dict__.update(dict(nfine=nfine, nles=nles, kc=kc, nu=nu, dt=dt, spin=spin, snaps=set(snaps),
                                  window=window, ratio=ratio, f0=f0, fstep=fstep, k0=k0, u0=u0, ic_seed=ic_seed,
                                  f_seed=f_seed, samples=samples, cal_window=int(cal_window),
                                  train_seeds=tuple(train_seeds), check_kmin=int(check_kmin), log=log))
        assert window % (ratio * samples) == 0, "window must hold whole LES steps per sample"
        assert cal_window % (ratio * samples) == 0, "the calibration window must hold whole LES steps per sample"

    def _dns(self, ic_seed, f_seed):
        """A forced DNS on the fine grid and its step function; the phases it
        draws are appended to the returned list."""
        from .solvers import HDSolver
        from .pseudospec import randu
        s = HDSolver(self.fine, dt=self.dt, nu=self.nu, ord=2)
        v, _ = self.fine.initialv_k4exp(self.k0, self.u0, ic_seed)
        f, fseed = self.fine.initialfv_random(2.0, 3.0, self.f0, f_seed)
        s.set_velocity(v)
        s.set_forcing(f)
        del v, f
        st = {"timef": self.fstep, "seed": fseed}
        phases = []

        def step():
            if st["timef"] == self.fstep:
                st["timef"] = 0
                r, st["seed"] = randu(st["seed"])
                ph = 2 * math.pi * r
                phases.append(ph)
                s.phaseshift_forcing(complex(math.cos(ph), math.sin(ph)))
            s.step(1)
            st["timef"] += 1

        return s, step, phases, st

    def run_dns(self):
        """The training DNS, then the scored DNS."""
        self.run_training_dns()
        self.run_scored_dns()

    def run_training_dns(self):
        import time
        self.fine, self.small = Grid(self.nfine), Grid(self.nles)
        s, step, phases, st = self._dns(*self.train_seeds)
        self.phases_train = phases
        self.train = TrainingSet(self.small, self.kc)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for it in range(1, self.spin + 1):
            step()
            if it % 1000 == 0:
                self.log(f"training DNS step {it}/{self.spin}")
            if it in self.snaps:
                self.train.add(self.fine, s.velocity())
        torch.cuda.synchronize()
        self.t_spin_train = time.perf_counter() - t0
        # the calibration window: the filtered state and forcing at its start,
        # then its filtered spectra, sampled as the scored window samples its own
        self.cal_ref = []
        if self.cal_window:
            _, self.cal_v = sgs_targets(self.fine, self.small, s.velocity(), self.kc)
            self.cal_f = filter_forcing(self.fine, self.small, s.forcing(), self.kc)
            self.cal_phase, self.cal_timef = len(phases), st["timef"]
            every = self.cal_window // self.samples
            for it in range(1, self.cal_window + 1):
                step()
                if it % every == 0:
                    _, vf = sgs_targets(self.fine, self.small, s.velocity(), self.kc)
                    self.cal_ref.append(spectrum(self.small, vf))
                    self.log(f"training DNS calibration window step {it}/{self.cal_window}")
        torch.cuda.synchronize()
        self.t_train_dns = time.perf_counter() - t0
        # the reference an LES can be checked against at run time: the training
        # DNS's filtered spectrum, time-averaged over its calibration window
        self.ref_runtime = np.mean(self.cal_ref, axis=0) if self.cal_ref else None
        self.log(f"training DNS {self.nfine}^3 (seeds {self.train_seeds}): {self.spin} + {self.cal_window} steps "
                 f"{self.t_train_dns:.1f} s, {len(self.train.snaps)} training states")
        del s
        torch.cuda.empty_cache()

    def run_scored_dns(self):
        import time
        if not hasattr(self, "fine"):
            self.fine, self.small = Grid(self.nfine), Grid(self.nles)
        s, step, phases, st = self._dns(self.ic_seed, self.f_seed)
        self.phases = phases
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for it in range(1, self.spin + 1):
            step()
            if it % 1000 == 0:
                self.log(f"scored DNS spin-up step {it}/{self.spin}")
        torch.cuda.synchronize()
        self.t_spin = time.perf_counter() - t0
        _, self.v_start = sgs_targets(self.fine, self.small, s.velocity(), self.kc)
        self.f_start = filter_forcing(self.fine, self.small, s.forcing(), self.kc)
        self.phase0, self.timef0 = len(phases), st["timef"]
        # the fine state at the window's start, for the DNS fallback
        self.ckpt = ([x.clone() for x in s.velocity()], [x.clone() for x in s.forcing()], st["timef"], st["seed"],
                     len(phases))
        self.ref, self.t_dns = self._dns_window(s, step)
        self.log(f"scored DNS {self.nfine}^3 (seeds {(self.ic_seed, self.f_seed)}): spin-up {self.spin} steps "
                 f"{self.t_spin:.1f} s, window {self.window} steps {self.t_dns:.2f} s")
        del s
        torch.cuda.empty_cache()

    def _dns_window(self, s, step):
        """The window: filtered spectra every window/samples steps, and the GPU
        seconds of the DNS steps themselves."""
        import time
        every = self.window // self.samples
        ref = []
        t_sample = 0.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for it in range(1, self.window + 1):
            step()
            if it % every == 0:
                torch.cuda.synchronize()
                ta = time.perf_counter()
                _, vf = sgs_targets(self.fine, self.small, s.velocity(), self.kc)
                ref.append(spectrum(self.small, vf))
                torch.cuda.synchronize()
                t_sample += time.perf_counter() - ta
                self.log(f"DNS window step {it}/{self.window}")
        torch.cuda.synchronize()
        return ref, time.perf_counter() - t0 - t_sample

    def fallback_dns(self):
        """The plain DNS over the window from the checkpoint at its start, as
        the fallback runs it: returns (filtered spectra, GPU seconds)."""
        from .solvers import HDSolver
        from .pseudospec import randu
        v, f, timef, seed, nph = self.ckpt
        s = HDSolver(self.fine, dt=self.dt, nu=self.nu, ord=2)
        s.set_velocity(v)
        s.set_forcing(f)
        st = {"timef": timef, "seed": seed}

        def step():
            if st["timef"] == self.fstep:
                st["timef"] = 0
                r, st["seed"] = randu(st["seed"])
                ph = 2 * math.pi * r
                s.phaseshift_forcing(complex(math.cos(ph), math.sin(ph)))
            s.step(1)
            st["timef"] += 1

        out = self._dns_window(s, step)
        del s
        torch.cuda.empty_cache()
        return out

    def fit(self, steps=3000, batch=65536, lr=1e-3, seed=7, match=True):
        import time
        self.mlp = MLP(seed)
        X, T = self.train.arrays(self.mlp)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.losses = self.mlp.train(X, T, batch=min(batch, X.shape[0]), steps=steps, lr=lr, seed=seed)
        torch.cuda.synchronize()
        self.t_train = time.perf_counter() - t0
        self.n_train = int(X.shape[0])
        del X, T
        self.scale = self.train.match_dissipation(self.mlp) if match else 1.0
        self.log(f"closure trained on {self.n_train} points, {steps} Adam steps in {self.t_train:.2f} s: "
                 f"loss {self.losses[0]:.4f} -> {self.losses[-1]:.4f}; dissipation-matching scale {self.scale:.3f}")

    def fit_spectral(self, with_network=True):
        """The spectral part of the closure on the same training states (the
        network's residual transfer, or the whole subgrid transfer without a
        network)."""
        nu = self.train.spectral_viscosity(self.mlp if with_network else None)
        if with_network:
            self.nuk = nu
        else:
            self.nuk_alone = nu
        self.log(f"spectral part ({'residual of the network' if with_network else 'alone'}): nu(k) at k = 10, 20, "
                 f"30, 40: {nu[9]:.2e}, {nu[19]:.2e}, {nu[29]:.2e}, {nu[39]:.2e}")
        return nu

    def _les(self, v0, f0, phases, phase0, timef0, dns_steps, every_dns, closure, nuk, ratio):
        """LES from state v0 (forcing f0, the DNS's phase schedule from index
        phase0 with timef0) over dns_steps DNS steps: spectra of the sharp-
        filtered velocity every every_dns DNS steps and the GPU seconds of the
        steps themselves."""
        from .solvers import HDSolver
        import time
        assert dns_steps % ratio == 0 and every_dns % ratio == 0, "samples must fall on whole LES steps"
        L = HDSolver(self.small, dt=self.dt * ratio, nu=self.nu, ord=2)
        L.set_velocity(v0)
        L.set_forcing(f0)
        if closure:
            L.set_sgs(self.mlp, math.pi / self.kc)
        if nuk is not None:
            L.set_sgs_spectral(nuk)
        st = {"timef": timef0, "pi": phase0}
        every = every_dns // ratio
        spectra = []
        t_sample = 0.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for it in range(1, dns_steps // ratio + 1):
            for _ in range(ratio):  # the DNS's phase shifts that fall in this step
                if st["timef"] == self.fstep:
                    st["timef"] = 0
                    ph = phases[st["pi"]]
                    st["pi"] += 1
                    L.phaseshift_forcing(complex(math.cos(ph), math.sin(ph)))
                st["timef"] += 1
            L.step(1)
            if it % every == 0:
                torch.cuda.synchronize()
                ta = time.perf_counter()
                # the same sharp filter as the reference: the LES grid keeps
                # modes up to its own dealias bound, slightly above kc
                vel = [self.small.sharp(x, self.small.czeros(), math.pi / self.kc) for x in L.velocity()]
                spectra.append(spectrum(self.small, vel))
                t_sample += time.perf_counter() - ta
        torch.cuda.synchronize()
        return spectra, time.perf_counter() - t0 - t_sample

    def calibrate(self, nuk, closure=True, iters=4, gamma=1.0, ratio=None):
        """A-posteriori fit of the spectral part on the training DNS: LES from
        its filtered state at the first calibration step over the rest of its
        spin-up, against its filtered spectra sampled there; each pass scales the effective viscosity nu(k) + nu_mol by
        (E_les/E_dns)^gamma on the shells 1..kc (shells past kc take the value
        at kc)."""
        ratio = ratio or self.ratio
        span = self.cal_window
        every = self.cal_window // self.samples
        ref = np.mean(self.cal_ref, axis=0)
        nu = np.array(nuk, dtype=np.float64)
        kc = int(self.kc)
        for it in range(iters):
            spectra, _ = self._les(self.cal_v, self.cal_f, self.phases_train, self.cal_phase, self.cal_timef, span,
                                   every, closure, nu, ratio)
            if not all(np.all(np.isfinite(e)) for e in spectra):
                self.log(f"calibration pass {it + 1}: the LES blew up; stopping with the previous nu")
                break
            e = np.mean(spectra, axis=0)
            err = spectrum_error(e, ref, kc)
            r = np.ones_like(nu)
            r[:kc] = e[:kc] / ref[:kc]
            r[kc:] = r[kc - 1]
            # the effective viscosity nu + nu_mol (positive) is what is scaled
            nu = (nu + self.nu) * np.power(r, gamma) - self.nu
            self.log(f"calibration pass {it + 1}: spectrum error on the spin-up {err:.4f}; E_les/E_dns at k = 30, 40, "
                     f"{kc}: {r[29]:.3f}, {r[39]:.3f}, {r[kc - 1]:.3f}")
        return nu

    def run_les(self, closure=True, nuk=None, ratio=None, label=None, fallback_tol=None):
        """closure: attach the network; nuk: the spectral part (or None);
        ratio: LES steps per DNS step (default the experiment's).  With
        fallback_tol, the LES's time-averaged spectrum is checked against the
        training DNS's (the reference known at run time); above the tolerance
        the plain DNS runs the window from its checkpoint and its result is
        the run's result."""
        import time
        ratio = ratio or self.ratio
        assert self.window % (ratio * self.samples) == 0, "window must hold whole LES steps per sample"
        spectra, elapsed = self._les(self.v_start, self.f_start, self.phases, self.phase0, self.timef0, self.window,
                                     self.window // self.samples, closure, nuk, ratio)
        ok = all(np.all(np.isfinite(e)) for e in spectra)
        mean_err = spectrum_error(np.mean(spectra, axis=0), np.mean(self.ref, axis=0), self.kc) if ok else math.inf
        end_err = spectrum_error(spectra[-1], self.ref[-1], self.kc) if ok else math.inf
        name = label or ("closure" if closure else "no closure")
        self.log(f"LES {self.nles}^3 {name}: {self.window // ratio} steps {elapsed:.2f} s, spectrum error "
                 f"(time average) {mean_err:.4f}, at the end {end_err:.4f}")
        if ok:
            ea, er = np.mean(spectra, axis=0), np.mean(self.ref, axis=0)
            shells = [k for k in (1, 2, 3, 5, 8, 12, 16, 20, 25, 30, 35, 40, int(self.kc)) if k <= self.kc]
            self.log("  E_les/E_dns at k = " + ", ".join(f"{k}: {ea[k - 1] / er[k - 1]:.3f}" for k in shells))
        out = {"seconds": elapsed, "steps": self.window // ratio, "spectrum_error_mean": mean_err,
               "spectrum_error_end": end_err, "finite": bool(ok)}
        if ok:
            out["spectrum_les"] = [float(x) for x in np.mean(spectra, axis=0)[:int(self.kc)]]
        if fallback_tol is not None:
            # The LES against the training DNS: another realisation of the same
            # forced flow, so only the shells above the forcing band, whose
            # time averages do not depend on the realisation's large eddies,
            # are compared (check_kmin .. kc)
            check = (spectrum_error_band(np.mean(spectra, axis=0), self.ref_runtime, self.check_kmin, self.kc)
                     if ok and self.ref_runtime is not None else math.inf)
            out["runtime_check_error"] = check
            out["fallback"] = bool(not (check <= fallback_tol))
            if out["fallback"]:
                self.log(f"  run-time check {check:.4f} > {fallback_tol}: falling back to the {self.nfine}^3 DNS")
                t0 = time.perf_counter()
                dspec, dsec = self.fallback_dns()
                out["fallback_dns_seconds"] = dsec
                out["fallback_wall_seconds"] = time.perf_counter() - t0
                out["result_error_mean"] = spectrum_error(np.mean(dspec, axis=0), np.mean(self.ref, axis=0), self.kc)
                out["result_seconds"] = elapsed + dsec
            else:
                out["result_error_mean"] = mean_err
                out["result_seconds"] = elapsed
        return out
