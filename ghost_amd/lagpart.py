"""Lagrangian tracers on the gfx950 path: GHOST's GPart (lagpart/gparts_mod.f90)
with cubic-spline interpolation (lagpart/gpsplineint_mod.fpp) on one rank.

Positions are in grid units, [0, n) per axis, as GPart's px_/py_/pz_.
"""
import ctypes as C
import math

import torch

from . import _lib
from .pseudospec import Grid


class GPart:
    """n tracers on `grid`.  zcompat=True reproduces GHOST's z stencil (the
    field is sampled one grid cell above the particle, clamped in the top
    cell); zcompat=False interpolates at the particle."""

    def __init__(self, grid: Grid, n, zcompat=True):
        self.grid, self.n, self.zcompat = grid, int(n), bool(zcompat)
        h = C.c_void_p()
        _lib.check(_lib.load().ghx_tracers_create(C.byref(h), grid._h, self.n, int(self.zcompat)))
        self._h = h
        p = [C.c_void_p() for _ in range(3)]
        u = [C.c_void_p() for _ in range(3)]
        L = _lib.load()
        _lib.check(L.ghx_tracers_positions(h, *[C.byref(x) for x in p]))
        _lib.check(L.ghx_tracers_velocities(h, *[C.byref(x) for x in u]))
        self._p = [x.value for x in p]
        self._u = [x.value for x in u]

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ghx_tracers_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _copy_in(self, ptr, t):
        t = t.to(device="cuda", dtype=torch.float64).contiguous()
        assert t.numel() == self.n, "one value per tracer"
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_memcpy_d2d(C.c_void_p(ptr), C.c_void_p(t.data_ptr()), self.n * 8))

    def _copy_out(self, ptr):
        t = torch.empty(self.n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_memcpy_d2d(C.c_void_p(t.data_ptr()), C.c_void_p(ptr), self.n * 8))
        return t

    def set_positions(self, px, py, pz):
        """GPart_InitUserSeed (gparts_mod.f90:640-727) from arrays in grid units."""
        for ptr, t in zip(self._p, (px, py, pz)):
            self._copy_in(ptr, torch.as_tensor(t))

    def positions(self):
        return [self._copy_out(p) for p in self._p]

    def velocities(self):
        """Velocities interpolated by the last stage (GPart lvx_, lvy_, lvz_)."""
        return [self._copy_out(p) for p in self._u]

    def set_step(self):
        """GPart_SetStepRKK: x0 = x."""
        _lib.check(_lib.load().ghx_tracers_set_step(self._h))

    def step(self, v, dt, xk):
        """One stage (GPart_StepRKK): v = spectral velocity (3 fields)."""
        vs = [x.contiguous() for x in v]
        pv, _keep = _lib.ptr_array([x.data_ptr() for x in vs])
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_tracers_step(self._h, pv, float(dt), float(xk)))

    def euler_to_lag(self, field):
        """GPart_EulerToLag: spline interpolant of a real field at the tracers."""
        f = field.contiguous()
        out = torch.empty(self.n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_tracers_euler_to_lag(self._h, C.c_void_p(f.data_ptr()),
                                                        C.c_void_p(out.data_ptr())))
        return out

    def invdel(self):
        """n / (2 pi L) per axis (gparts_mod.f90:207-213)."""
        return [n / (2 * math.pi * L) for n, L in zip((self.grid.nx, self.grid.ny, self.grid.nz), self.grid.L)]
