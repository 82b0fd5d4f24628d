"""Resident-field HD and MHD solvers (the RK loop of main3D.fpp:2539-3418
without output and diagnostics), on the fused gfx950 substep."""
import ctypes as C

import torch

from . import _lib
from .pseudospec import Grid


class _Resident:
    def _copy_in(self, ptr, t):
        t = t.contiguous()
        _lib.check(_lib.load().ghx_memcpy_d2d(C.c_void_p(ptr), C.c_void_p(t.data_ptr()), t.numel() * 16))

    def _copy_out(self, ptr):
        t = self.grid.czeros()
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_memcpy_d2d(C.c_void_p(t.data_ptr()), C.c_void_p(ptr), t.numel() * 16))
        return t


class HDSolver(_Resident):
    """HD3D state v (3 fields) and forcing f (3 fields) resident in HBM."""

    def __init__(self, grid: Grid, dt, nu, ord=2):
        self.grid, self.dt, self.nu, self.ord = grid, dt, nu, ord
        h = C.c_void_p()
        _lib.check(_lib.load().ghx_hd_create(C.byref(h), grid._h, ord, float(dt), float(nu)))
        self._h = h
        v = (C.c_void_p * 3)()
        f = (C.c_void_p * 3)()
        _lib.check(_lib.load().ghx_hd_fields(h, C.cast(v, C.POINTER(C.c_void_p)), C.cast(f, C.POINTER(C.c_void_p))))
        self._v = [v[k] for k in range(3)]
        self._f = [f[k] for k in range(3)]

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ghx_hd_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def set_velocity(self, v):
        torch.cuda.synchronize()
        for k in range(3):
            self._copy_in(self._v[k], v[k])
        _lib.check(_lib.load().ghx_hd_touch(self._h))

    def set_forcing(self, f):
        torch.cuda.synchronize()
        if f is None:
            _lib.check(_lib.load().ghx_hd_set_forcing(self._h, 0))
            return
        for k in range(3):
            self._copy_in(self._f[k], f[k])
        _lib.check(_lib.load().ghx_hd_set_forcing(self._h, 1))

    def velocity(self):
        return [self._copy_out(p) for p in self._v]

    def forcing(self):
        return [self._copy_out(p) for p in self._f]

    def step(self, n=1):
        _lib.check(_lib.load().ghx_hd_step(self._h, int(n)))

    def set_tracers(self, parts):
        """Advance the GPart `parts` with the flow: one tracer stage after every
        substep (main3D.fpp:3202-3267); None detaches."""
        _lib.check(_lib.load().ghx_hd_set_tracers(self._h, parts._h if parts is not None else None))
        self._parts = parts  # keep the tracer handle alive while attached

    def phaseshift_forcing(self, cdump):
        """The rand=1 forcing update of main3D.fpp:2544-2556 on the resident f."""
        L = _lib.load()
        for p in self._f:
            _lib.check(L.ghx_phaseshift(self.grid._h, C.c_void_p(p), float(cdump.real), float(cdump.imag)))

    def set_rotation(self, omegaz):
        """The ROTH solver (include/roth_rkstep2.f90): a frame rotating about z
        at rate omegaz."""
        _lib.check(_lib.load().ghx_hd_set_rotation(self._h, float(omegaz)))

    def set_sgs(self, mlp, delta):
        """Attach the learned closure (ghost_amd.closure.MLP; None detaches): before
        each substep the forcing becomes f + P(-div tau), tau = mlp(grad u, delta)."""
        _lib.check(_lib.load().ghx_hd_set_sgs(self._h, mlp._h if mlp is not None else None, float(delta)))
        self._sgs = mlp  # keep the network alive while attached

    def set_sgs_spectral(self, nuk):
        """Spectral part of the closure: -nuk[q] kk2 v on shell q + 1 added to
        the force of every substep (None or empty removes it)."""
        import numpy as np
        a = np.ascontiguousarray(nuk if nuk is not None else [], dtype=np.float64)
        p = a.ctypes.data_as(C.POINTER(C.c_double)) if a.size else None
        _lib.check(_lib.load().ghx_hd_set_sgs_spectral(self._h, int(a.size), p))


class MHDSolver(_Resident):
    """MHD3D state: velocity v, vector potential a, forcings f (mechanical) and m (electromotive)."""

    def __init__(self, grid: Grid, dt, nu, mu, ord=2):
        self.grid, self.dt, self.nu, self.mu, self.ord = grid, dt, nu, mu, ord
        h = C.c_void_p()
        _lib.check(_lib.load().ghx_mhd_create(C.byref(h), grid._h, ord, float(dt), float(nu), float(mu)))
        self._h = h
        arrs = [(C.c_void_p * 3)() for _ in range(4)]
        _lib.check(_lib.load().ghx_mhd_fields(h, *[C.cast(a, C.POINTER(C.c_void_p)) for a in arrs]))
        self._v, self._a, self._f, self._m = ([a[k] for k in range(3)] for a in arrs)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ghx_mhd_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def set_state(self, v, a):
        torch.cuda.synchronize()
        for k in range(3):
            self._copy_in(self._v[k], v[k])
            self._copy_in(self._a[k], a[k])
        _lib.check(_lib.load().ghx_mhd_touch(self._h))

    def set_forcing(self, f=None, m=None):
        torch.cuda.synchronize()
        if f is None and m is None:
            _lib.check(_lib.load().ghx_mhd_set_forcing(self._h, 0))
            return
        zero = self.grid.czeros()
        for k in range(3):
            self._copy_in(self._f[k], f[k] if f is not None else zero)
            self._copy_in(self._m[k], m[k] if m is not None else zero)
        _lib.check(_lib.load().ghx_mhd_set_forcing(self._h, 1))

    def state(self):
        return [self._copy_out(p) for p in self._v], [self._copy_out(p) for p in self._a]

    def phaseshift_forcing(self, cdump):
        """The rand=1 update of the mechanical forcing (main3D.fpp:2544-2556)."""
        L = _lib.load()
        for p in self._f:
            _lib.check(L.ghx_phaseshift(self.grid._h, C.c_void_p(p), float(cdump.real), float(cdump.imag)))

    def step(self, n=1):
        _lib.check(_lib.load().ghx_mhd_step(self._h, int(n)))

    def set_tracers(self, parts):
        """As HDSolver.set_tracers."""
        _lib.check(_lib.load().ghx_mhd_set_tracers(self._h, parts._h if parts is not None else None))
        self._parts = parts


class BoussSolver(_Resident):
    """Boussinesq state: velocity v (3), temperature th, forcings f (3) and fs
    (include/bouss_rkstep{1,2}.f90).  xmom and xtemp are the buoyancy and
    heat-current factors already multiplied by bvfreq."""

    def __init__(self, grid: Grid, dt, nu, kappa, xmom, xtemp, ord=2):
        self.grid, self.dt, self.nu, self.kappa, self.ord = grid, dt, nu, kappa, ord
        h = C.c_void_p()
        _lib.check(_lib.load().ghx_bouss_create(C.byref(h), grid._h, ord, float(dt), float(nu), float(kappa),
                                                float(xmom), float(xtemp)))
        self._h = h
        v, f = (C.c_void_p * 3)(), (C.c_void_p * 3)()
        th, fs = C.c_void_p(), C.c_void_p()
        _lib.check(_lib.load().ghx_bouss_fields(h, C.cast(v, C.POINTER(C.c_void_p)), C.byref(th),
                                                C.cast(f, C.POINTER(C.c_void_p)), C.byref(fs)))
        self._v = [v[k] for k in range(3)]
        self._f = [f[k] for k in range(3)]
        self._th, self._fs = th.value, fs.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ghx_bouss_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def set_state(self, v, th):
        torch.cuda.synchronize()
        for k in range(3):
            self._copy_in(self._v[k], v[k])
        self._copy_in(self._th, th)
        _lib.check(_lib.load().ghx_bouss_touch(self._h))

    def set_forcing(self, f=None, fs=None):
        torch.cuda.synchronize()
        if f is None and fs is None:
            _lib.check(_lib.load().ghx_bouss_set_forcing(self._h, 0))
            return
        zero = self.grid.czeros()
        for k in range(3):
            self._copy_in(self._f[k], f[k] if f is not None else zero)
        self._copy_in(self._fs, fs if fs is not None else zero)
        _lib.check(_lib.load().ghx_bouss_set_forcing(self._h, 1))

    def state(self):
        return [self._copy_out(p) for p in self._v], self._copy_out(self._th)

    def step(self, n=1):
        _lib.check(_lib.load().ghx_bouss_step(self._h, int(n)))

    def set_tracers(self, parts):
        """As HDSolver.set_tracers."""
        _lib.check(_lib.load().ghx_bouss_set_tracers(self._h, parts._h if parts is not None else None))
        self._parts = parts

    def set_rotation(self, omegaz):
        """The ROTBOUSS solver (include/rotbouss_rkstep2.f90): Coriolis force of
        rotation omegaz about z; any RK order."""
        _lib.check(_lib.load().ghx_bouss_set_rotation(self._h, float(omegaz)))


class PHDSolver(BoussSolver):
    """Passive scalar th advected by the flow (include/phd_rkstep{1,2}.f90):
    the Boussinesq step without buoyancy and heat current."""

    def __init__(self, grid: Grid, dt, nu, kappa, ord=2):
        super().__init__(grid, dt, nu, kappa, 0.0, 0.0, ord=ord)


class Ensemble:
    """Independent unforced HD runs advanced together: members of one grid
    size share each pass launch, and after the first step the ensemble step
    is a replayed HIP graph."""

    def __init__(self, sizes, nus, dt, ord=2):
        n = len(sizes)
        self.sizes = [int(x) for x in sizes]
        arr_n = (C.c_int * n)(*self.sizes)
        arr_nu = (C.c_double * n)(*[float(x) for x in nus])
        h = C.c_void_p()
        _lib.check(_lib.load().ghx_ensemble_create(C.byref(h), n, arr_n, arr_nu, float(dt), int(ord)))
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ghx_ensemble_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _ptrs(self, m):
        v = (C.c_void_p * 3)()
        _lib.check(_lib.load().ghx_ensemble_fields(self._h, m, C.cast(v, C.POINTER(C.c_void_p))))
        return [v[k] for k in range(3)]

    def set_velocity(self, m, v):
        n = self.sizes[m]
        torch.cuda.synchronize()
        for p, t in zip(self._ptrs(m), v):
            t = t.contiguous()
            assert t.numel() == (n // 2 + 1) * n * n
            _lib.check(_lib.load().ghx_memcpy_d2d(C.c_void_p(p), C.c_void_p(t.data_ptr()), t.numel() * 16))
        _lib.check(_lib.load().ghx_ensemble_touch(self._h))

    def velocity(self, m):
        n = self.sizes[m]
        out = []
        _lib.check(_lib.load().ghx_ensemble_sync(self._h))
        for p in self._ptrs(m):
            t = torch.empty((n // 2 + 1, n, n), dtype=torch.complex128, device="cuda")
            _lib.check(_lib.load().ghx_memcpy_d2d(C.c_void_p(t.data_ptr()), C.c_void_p(p), t.numel() * 16))
            out.append(t)
        return out

    def step(self, n=1):
        _lib.check(_lib.load().ghx_ensemble_step(self._h, int(n)))

    def sync(self):
        _lib.check(_lib.load().ghx_ensemble_sync(self._h))

    def set_graph(self, enabled):
        _lib.check(_lib.load().ghx_ensemble_set_graph(self._h, int(bool(enabled))))

    def groups(self):
        ng = C.c_int()
        sizes, counts = (C.c_int * 64)(), (C.c_int * 64)()
        _lib.check(_lib.load().ghx_ensemble_groups(self._h, C.byref(ng), sizes, counts))
        return [(sizes[k], counts[k]) for k in range(ng.value)]

    def autotune(self, steps=3):
        """Time every radix schedule and tile-width pair per size group; returns
        the chosen (z, y) tile widths per group (schedules() has all three)."""
        ng = len(self.groups())
        tz, ty = (C.c_int * ng)(), (C.c_int * ng)()
        _lib.check(_lib.load().ghx_ensemble_autotune(self._h, int(steps), tz, ty))
        return [(tz[k], ty[k]) for k in range(ng)]

    def schedules(self):
        """Per size group: (schedule bits, z tile width, y tile width); bit 0
        puts the z passes, bit 1 the y passes on four points per lane."""
        ng = len(self.groups())
        sc, tz, ty = (C.c_int * ng)(), (C.c_int * ng)(), (C.c_int * ng)()
        _lib.check(_lib.load().ghx_ensemble_schedules(self._h, sc, tz, ty))
        return [(sc[k], tz[k], ty[k]) for k in range(ng)]

    def set_schedule(self, group, sched, tz=0, ty=0):
        _lib.check(_lib.load().ghx_ensemble_set_schedule(self._h, int(group), int(sched), int(tz), int(ty)))
