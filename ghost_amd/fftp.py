"""Python mirror of GHOST's ``fftp`` module on one rank (3D/src/fftp-3/fftp3D.fpp).

Arrays are torch tensors on the GPU (torch is used only for device memory):
  real     float64    shape (nz, ny, nx)        == Fortran REAL(GP) R(nx,ny,nz)
  spectral complex128 shape (nx//2+1, ny, nz)   == Fortran COMPLEX(GP) C(nz,ny,nx/2+1)
"""
import ctypes as C

import torch

from . import _lib

FFTW_REAL_TO_COMPLEX = -1  # fftp-3/fftp_mod.fpp:27
FFTW_COMPLEX_TO_REAL = 1   # fftp-3/fftp_mod.fpp:28
FFTW_ESTIMATE = 64
FFTW_MEASURE = 0


def _dptr(t, dtype, shape):
    if not (t.is_cuda and t.dtype == dtype and t.is_contiguous() and tuple(t.shape) == tuple(shape)):
        raise ValueError(f"expected a contiguous {dtype} GPU tensor of shape {tuple(shape)}, got "
                         f"{t.dtype} {tuple(t.shape)} on {t.device}")
    return t.data_ptr()


class FFTPlan:
    """TYPE(FFTPLAN) (fftp-3/fftp_mod.fpp:34-41): nx, ny, nz and the direction."""

    def __init__(self, n, fftdir, flags=FFTW_ESTIMATE):
        nx, ny, nz = (int(v) for v in n)
        self.nx, self.ny, self.nz, self.fftdir = nx, ny, nz, fftdir
        L = _lib.load()
        h = C.c_void_p()
        _lib.check(L.ghx_plan_create(C.byref(h), nx, ny, nz, fftdir, flags))
        self._h = h
        # launches follow torch's current stream so ordering with torch ops holds
        _lib.check(L.ghx_plan_set_stream(h, C.c_void_p(torch.cuda.current_stream().cuda_stream)))

    @property
    def rshape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def cshape(self):
        return (self.nx // 2 + 1, self.ny, self.nz)

    def destroy(self):
        if getattr(self, "_h", None):
            _lib.load().ghx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def timers(self):
        f, t, c, tot = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        _lib.check(_lib.load().ghx_plan_timers(self._h, C.byref(f), C.byref(t), C.byref(c), C.byref(tot)))
        return {"ffttime": f.value, "tratime": t.value, "comtime": c.value, "tottime": tot.value}

    def reset_timers(self):
        _lib.check(_lib.load().ghx_plan_reset_timers(self._h))


def fftp3d_create_plan(n, fftdir, flags=FFTW_ESTIMATE):
    """fftp3d_create_plan (fftp-3/fftp3D.fpp:56-124)."""
    return FFTPlan(n, fftdir, flags)


def fftp3d_destroy_plan(plan):
    """fftp3d_destroy_plan (fftp-3/fftp3D.fpp:126-155)."""
    plan.destroy()


def fftp3d_real_to_complex(plan, inp, out, comm=None):
    """fftp3d_real_to_complex (fftp-3/fftp3D.fpp:204-315): out = FFT(in), sign -1, unnormalised."""
    a = _dptr(inp, torch.float64, plan.rshape)
    b = _dptr(out, torch.complex128, plan.cshape)
    _lib.check(_lib.load().ghx_r2c(plan._h, C.c_void_p(a), C.c_void_p(b)))
    return out


def fftp3d_complex_to_real(plan, inp, out, comm=None):
    """fftp3d_complex_to_real (fftp-3/fftp3D.fpp:318-431): sign +1, unnormalised.
    The input is preserved (GHOST's CPU path destroys it)."""
    a = _dptr(inp, torch.complex128, plan.cshape)
    b = _dptr(out, torch.float64, plan.rshape)
    _lib.check(_lib.load().ghx_c2r(plan._h, C.c_void_p(a), C.c_void_p(b)))
    return out
