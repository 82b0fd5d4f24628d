"""GHOST's binary field files (mpiio/binary_io.f90, modules iovar and iompi of
mpiio/binary_mod.f90) over the library's staged writer.

The file holds the global array in Fortran order as raw native FP64 (complex:
re, im pairs), no header, from byte offset ``disp``.  A real field of this
package, shape (nz, ny, nx), is exactly Fortran R(nx,ny,nz); a spectral field,
shape (nx/2+1, ny, nz), exactly C(nz,ny,nx/2+1).  ``var`` may be a CUDA tensor
(staged through pinned buffers) or a contiguous host array (numpy or CPU
tensor), which is read or written directly.
"""
import ctypes as C

from . import _lib


class IOPlan:
    """TYPE(IOPLAN) of io_init (real arrays) or io_initc (complex arrays):
    arrays var(n1, n2, ksta:kend) of the global (n1, n2, n3) array.  iswap,
    bmangle and disp are the module variables of iovar and iompi, held per
    plan here."""

    def __init__(self, myrank, n, ksta, kend, complex_=False, iswap=0, bmangle=1, disp=0):
        h = C.c_void_p()
        nn = (C.c_int * 3)(*[int(x) for x in n])
        init = _lib.load().ghx_io_initc if complex_ else _lib.load().ghx_io_init
        _lib.check(init(C.byref(h), int(myrank), nn, int(ksta), int(kend)))
        self._h, self.n, self.ksta, self.kend, self.complex = h, tuple(int(x) for x in n), ksta, kend, complex_
        self.set_options(iswap, bmangle, disp)

    def set_options(self, iswap=0, bmangle=1, disp=0):
        self.iswap, self.bmangle, self.disp = int(iswap), int(bmangle), int(disp)
        _lib.check(_lib.load().ghx_io_set_options(self._h, self.iswap, self.bmangle, self.disp))

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ghx_io_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def slab_elems(self):
        return self.n[0] * self.n[1] * (self.kend - self.ksta + 1)


def _ptr(var, plan):
    """Address of a contiguous array of the plan's slab size (host or device)."""
    esize = 16 if plan.complex else 8
    if hasattr(var, "data_ptr"):  # torch tensor
        assert var.is_contiguous(), "var must be contiguous"
        assert var.numel() * var.element_size() == plan.slab_elems * esize, "var does not match the plan"
        return var.data_ptr()
    assert var.flags["C_CONTIGUOUS"] or var.flags["F_CONTIGUOUS"], "var must be contiguous"
    assert var.nbytes == plan.slab_elems * esize, "var does not match the plan"
    return var.ctypes.data


def _b(s):
    return s.encode() if isinstance(s, str) else s


def io_read(unit, dir, fname, nmb, plan, var):
    """io_read(unit,dir,fname,nmb,plan,var) (mpiio/binary_io.f90:62-129); unit is unused."""
    _lib.check(_lib.load().ghx_io_read(plan._h, _b(dir), _b(fname), _b(nmb), C.c_void_p(_ptr(var, plan))))


def io_write(unit, dir, fname, nmb, plan, var):
    """io_write (mpiio/binary_io.f90:132-186); unit is unused."""
    _lib.check(_lib.load().ghx_io_write(plan._h, _b(dir), _b(fname), _b(nmb), C.c_void_p(_ptr(var, plan))))


def io_readc(unit, dir, fname, nmb, plan, var):
    """io_readc (mpiio/binary_io.f90:236-282)."""
    _lib.check(_lib.load().ghx_io_readc(plan._h, _b(dir), _b(fname), _b(nmb), C.c_void_p(_ptr(var, plan))))


def io_writec(unit, dir, fname, nmb, plan, var):
    """io_writec (mpiio/binary_io.f90:285-331)."""
    _lib.check(_lib.load().ghx_io_writec(plan._h, _b(dir), _b(fname), _b(nmb), C.c_void_p(_ptr(var, plan))))


def ext_label(tind):
    """main3D's time label: WRITE(ext, fmtext), fmtext = '(i4.4)' (pseudospec3D_mod.fpp:40)."""
    return "%04d" % int(tind)


def hd_output(grid, plan, odir, ext, v, outs=0):
    """Field output of main3D.fpp:2675-2710: vx, vy, vz = c2r(v/(nx ny nz)),
    and wx, wy, wz of the curl when outs >= 1."""
    _lib.check(_lib.load().ghx_hd_output(grid._h, plan._h, _b(odir), _b(ext), int(outs),
                                         *[grid._c(x) for x in v]))


def hd_restart(grid, plan, idir, ext, v):
    """Restart of main3D.fpp:2143-2169: vx, vy, vz read and transformed (r2c)."""
    _lib.check(_lib.load().ghx_hd_restart(grid._h, plan._h, _b(idir), _b(ext),
                                          *[grid._c(x) for x in v]))
