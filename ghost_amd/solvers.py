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
