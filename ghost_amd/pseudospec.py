"""Python mirror of GHOST's pseudospectral subroutines on device-resident fields.

The routines keep the Fortran names, argument order and meaning of
3D/src/pseudo/pseudospec3D_hd.f90, pseudospec3D_mhd.f90, pseudospec3D_filt.f90
and module_dns.f90; the module state GHOST keeps in ``grid``, ``kes``,
``ali``, ``boxsize`` and ``fft`` (pseudospec3D_mod.fpp:13-120) lives in
:class:`Grid`.  Spectral fields are complex128 GPU tensors of shape
(nx//2+1, ny, nz); real fields are float64 of shape (nz, ny, nx).
"""
import ctypes as C

import torch

from . import _lib
from .fftp import _dptr


class Grid:
    """Grid, wavenumbers and dealias bounds (main3D.fpp:737-749, 881-909, 1727-1800)."""

    def __init__(self, nx, ny=None, nz=None, L=(1.0, 1.0, 1.0), arbsize=False, Dkk=0.0):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.nxh = self.nx // 2 + 1
        self.L = tuple(float(x) for x in L) if arbsize else (1.0, 1.0, 1.0)
        L_ = _lib.load()
        h = C.c_void_p()
        _lib.check(L_.ghx_grid_create(C.byref(h), self.nx, self.ny, self.nz, float(L[0]), float(L[1]),
                                      float(L[2]), int(bool(arbsize)), float(Dkk)))
        self._h = h
        _lib.check(L_.ghx_grid_set_stream(h, C.c_void_p(torch.cuda.current_stream().cuda_stream)))
        kmax, tiny, nmax, dkk, anis = C.c_double(), C.c_double(), C.c_int(), C.c_double(), C.c_int()
        _lib.check(L_.ghx_grid_info(h, C.byref(kmax), C.byref(tiny), C.byref(nmax), C.byref(dkk), C.byref(anis)))
        self.kmax, self.tiny, self.nmax, self.Dkk, self.anis = kmax.value, tiny.value, nmax.value, dkk.value, anis.value
        self.device = torch.device("cuda", torch.cuda.current_device())

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ghx_grid_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # ---- shapes and allocation -------------------------------------------
    @property
    def cshape(self):
        return (self.nxh, self.ny, self.nz)

    @property
    def rshape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def nbins(self):
        return self.nmax // 2 + 1

    def czeros(self):
        return torch.zeros(self.cshape, dtype=torch.complex128, device=self.device)

    def rzeros(self):
        return torch.zeros(self.rshape, dtype=torch.float64, device=self.device)

    def _c(self, t):
        return C.c_void_p(_dptr(t, torch.complex128, self.cshape))

    def _r(self, t):
        return C.c_void_p(_dptr(t, torch.float64, self.rshape))

    # ---- fftp -------------------------------------------------------------
    def r2c(self, r, out=None):
        out = self.czeros() if out is None else out
        _lib.check(_lib.load().ghx_grid_r2c(self._h, self._r(r), self._c(out)))
        return out

    def c2r(self, c, out=None):
        out = self.rzeros() if out is None else out
        _lib.check(_lib.load().ghx_grid_c2r(self._h, self._c(c), self._r(out)))
        return out

    # ---- pseudospec3D_hd.f90 ----------------------------------------------
    def derivk3(self, a, b, dir):
        _lib.check(_lib.load().ghx_derivk3(self._h, self._c(a), self._c(b), dir))
        return b

    def laplak3(self, a, b):
        _lib.check(_lib.load().ghx_laplak3(self._h, self._c(a), self._c(b)))
        return b

    def rotor3(self, a, b, c, dir):
        _lib.check(_lib.load().ghx_rotor3(self._h, self._c(a), self._c(b), self._c(c), dir))
        return c

    def prodre3(self, a, b, c, d, e, f):
        _lib.check(_lib.load().ghx_prodre3(self._h, *map(self._c, (a, b, c, d, e, f))))
        return d, e, f

    def nonlhd3(self, a, b, c, g, dir):
        _lib.check(_lib.load().ghx_nonlhd3(self._h, *map(self._c, (a, b, c, g)), dir))
        return g

    def energy(self, a, b, c, kin):
        d = C.c_double()
        _lib.check(_lib.load().ghx_energy(self._h, *map(self._c, (a, b, c)), kin, C.byref(d)))
        return d.value

    def helicity(self, a, b, c):
        d = C.c_double()
        _lib.check(_lib.load().ghx_helicity(self._h, *map(self._c, (a, b, c)), C.byref(d)))
        return d.value

    def cross(self, a, b, c, d, e, f, kin):
        g = C.c_double()
        _lib.check(_lib.load().ghx_cross(self._h, *map(self._c, (a, b, c, d, e, f)), kin, C.byref(g)))
        return g.value

    def maxabs(self, a, b, c, kin):
        d = C.c_double()
        _lib.check(_lib.load().ghx_maxabs(self._h, *map(self._c, (a, b, c)), kin, C.byref(d)))
        return d.value

    def divergence(self, a, b, c):
        d = C.c_double()
        _lib.check(_lib.load().ghx_divergence(self._h, *map(self._c, (a, b, c)), C.byref(d)))
        return d.value

    def spectrumc(self, a, b, c, kin=1, hel=0):
        """Returns (Ektot, Hktot) as lists of nmax/2+1 doubles (spectrumc :1190)."""
        Ek = (C.c_double * self.nbins)()
        Hk = (C.c_double * self.nbins)()
        _lib.check(_lib.load().ghx_spectrumc(self._h, *map(self._c, (a, b, c)), kin, hel,
                                             C.cast(Ek, C.POINTER(C.c_double)), C.cast(Hk, C.POINTER(C.c_double))))
        return list(Ek), list(Hk)

    def _bins(self):
        return (C.c_double * self.nbins)()

    def entrans(self, a, b, c, d, e, f, kin=1):
        """Energy transfer per shell (entrans, pseudospec3D_hd.f90:1629-1820), the
        reduced totals; kin = 0 weights each mode by kk2."""
        out = self._bins()
        _lib.check(_lib.load().ghx_entrans(self._h, *map(self._c, (a, b, c, d, e, f)), int(kin),
                                           C.cast(out, C.POINTER(C.c_double))))
        return list(out)

    def heltrans(self, a, b, c, d, e, f):
        """Helicity transfer per shell (heltrans, :1823-1949)."""
        out = self._bins()
        _lib.check(_lib.load().ghx_heltrans(self._h, *map(self._c, (a, b, c, d, e, f)),
                                            C.cast(out, C.POINTER(C.c_double))))
        return list(out)

    def crosspecc(self, a, b, c, d, e, f, shift=0.0):
        """Cross-helicity spectrum of v and curl A (crosspecc, pseudospec3D_mhd.f90:603-712)."""
        out = self._bins()
        _lib.check(_lib.load().ghx_crosspecc(self._h, *map(self._c, (a, b, c, d, e, f)), float(shift),
                                             C.cast(out, C.POINTER(C.c_double))))
        return list(out)

    def spectr1d(self, a, dir):
        """1D spectrum of one component along k_dir (spectr1d, :1418-1626)."""
        out = self._bins()
        _lib.check(_lib.load().ghx_spectr1d(self._h, self._c(a), int(dir), C.cast(out, C.POINTER(C.c_double))))
        return list(out)

    # ---- scalar diagnostics (pseudospec3D_phd.f90) -------------------------
    def _dp(self, n):
        buf = (C.c_double * n)()
        return buf, C.cast(buf, C.POINTER(C.c_double))

    def variance(self, a, kin=1):
        out = C.c_double()
        _lib.check(_lib.load().ghx_variance(self._h, self._c(a), int(kin), C.byref(out)))
        return out.value

    def product(self, a, b):
        out = C.c_double()
        _lib.check(_lib.load().ghx_product(self._h, self._c(a), self._c(b), C.byref(out)))
        return out.value

    def pscheck(self, a, b):
        buf, p = self._dp(3)
        _lib.check(_lib.load().ghx_pscheck(self._h, self._c(a), self._c(b), p))
        return list(buf)

    def spectrscc(self, a, shift=0.0):
        buf, p = self._dp(self.nbins)
        _lib.check(_lib.load().ghx_spectrscc(self._h, self._c(a), float(shift), p))
        return list(buf)

    def sctrans(self, a, b):
        buf, p = self._dp(self.nbins)
        _lib.check(_lib.load().ghx_sctrans(self._h, self._c(a), self._c(b), p))
        return list(buf)

    def difuc(self, a, b, dir):
        """difucx (dir 1) / difucz (dir 3): [<v>, <th>, <v th>] along x or z."""
        n = self.nx if dir == 1 else self.nz
        buf, p = self._dp(3 * n)
        fn = _lib.load().ghx_difucx if dir == 1 else _lib.load().ghx_difucz
        _lib.check(fn(self._h, self._c(a), self._c(b), p))
        return [list(buf[q * n:(q + 1) * n]) for q in range(3)]

    # ---- Boussinesq diagnostics (pseudospec3D_bouss.f90) ---------------------
    def havgcomp(self, u, v, w, s, fo, bv, itype):
        """havgcomp (pseudospec3D_bouss.f90:19-555): horizontal average of quantity itype, per z."""
        buf, p = self._dp(self.nz)
        _lib.check(_lib.load().ghx_havgcomp(self._h, *map(self._c, (u, v, w, s)), float(fo), float(bv), int(itype), p))
        return list(buf)

    def tbouss(self, u, v, w, s, fo, bv):
        """tbouss (:612-709): (avg, max, min) over z of havgcomp itype 0..10."""
        (a, pa), (b, pb), (c, pc) = self._dp(11), self._dp(11), self._dp(11)
        _lib.check(_lib.load().ghx_tbouss(self._h, *map(self._c, (u, v, w, s)), float(fo), float(bv), pa, pb, pc))
        return list(a), list(b), list(c)

    def spectpv(self, u, v, w, s):
        buf, p = self._dp(self.nz // 2 + 1)
        _lib.check(_lib.load().ghx_spectpv(self._h, *map(self._c, (u, v, w, s)), p))
        return list(buf)

    # ---- axisymmetric spectra (pseudospec3D_rot.f90) -------------------------
    def axis_bins(self):
        a, b = C.c_int(), C.c_int()
        _lib.check(_lib.load().ghx_axis_bins(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def _axis(self, fn, nb, a, b, c, kin, hel):
        E, pe = self._dp(3 * nb) if kin <= 1 else (None, None)
        H, ph = self._dp(3 * nb) if hel == 1 else (None, None)
        _lib.check(fn(self._h, *map(self._c, (a, b, c)), int(kin), int(hel), pe, ph))
        split = (lambda x: [list(x[q * nb:(q + 1) * nb]) for q in range(3)])
        return (split(E) if E is not None else None), (split(H) if H is not None else None)

    def specpara(self, a, b, c, kin=1, hel=0):
        """specpara (pseudospec3D_rot.f90:25-335): ([tot, xy, z], [tot, xy, z]) per |kz| plane."""
        return self._axis(_lib.load().ghx_specpara, self.axis_bins()[0], a, b, c, kin, hel)

    def specperp(self, a, b, c, kin=1, hel=0):
        """specperp (:338-686): per k_perp cylinder; xy and z from the kz = 0 plane."""
        return self._axis(_lib.load().ghx_specperp, self.axis_bins()[1], a, b, c, kin, hel)

    def _axis1(self, fn, nb, args, kin=None):
        buf, p = self._dp(nb)
        extra = (int(kin),) if kin is not None else ()
        _lib.check(fn(self._h, *map(self._c, args), *extra, p))
        return list(buf)

    def entpara(self, a, b, c, d, e, f, kin=1):
        return self._axis1(_lib.load().ghx_entpara, self.axis_bins()[0], (a, b, c, d, e, f), kin)

    def entperp(self, a, b, c, d, e, f, kin=1):
        return self._axis1(_lib.load().ghx_entperp, self.axis_bins()[1], (a, b, c, d, e, f), kin)

    def heltpara(self, a, b, c, d, e, f):
        return self._axis1(_lib.load().ghx_heltpara, self.axis_bins()[0], (a, b, c, d, e, f))

    def heltperp(self, a, b, c, d, e, f):
        return self._axis1(_lib.load().ghx_heltperp, self.axis_bins()[1], (a, b, c, d, e, f))

    def spec2d(self, a, b, c, kin=1, hel=0):
        """spec2D (:1331-1629): arrays [kpara][kperp] of E and H (None where not asked)."""
        npara, nperp = self.axis_bins()
        n = npara * nperp
        E, pe = self._dp(n) if kin <= 1 else (None, None)
        H, ph = self._dp(n) if hel == 1 else (None, None)
        _lib.check(_lib.load().ghx_spec2d(self._h, *map(self._c, (a, b, c)), int(kin), int(hel), pe, ph))
        shape = lambda x: [list(x[q * nperp:(q + 1) * nperp]) for q in range(npara)]
        return (shape(E) if E is not None else None), (shape(H) if H is not None else None)

    def hdcheck(self, a, b, c, d, e, f, hel=1, chk=1):
        """hdcheck (pseudospec3D_hd.f90:984-1105): {eng, ens, pot, khe, div}."""
        out = (C.c_double * 5)()
        _lib.check(_lib.load().ghx_hdcheck(self._h, *map(self._c, (a, b, c, d, e, f)), int(hel), int(chk),
                                           C.cast(out, C.POINTER(C.c_double))))
        return list(out)

    def mhdcheck(self, a, b, c, ma, mb, mc, hel=1, crs=1, chk=1):
        """mhdcheck (pseudospec3D_mhd.f90:339-547): {eng, ens, cur, engk, engm, helk,
        helm, crh, asq, divk, divm}."""
        out = (C.c_double * 11)()
        _lib.check(_lib.load().ghx_mhdcheck(self._h, *map(self._c, (a, b, c, ma, mb, mc)), int(hel), int(crs),
                                            int(chk), C.cast(out, C.POINTER(C.c_double))))
        return list(out)

    def phaseshift(self, a, cdump):
        _lib.check(_lib.load().ghx_phaseshift(self._h, self._c(a), float(cdump.real), float(cdump.imag)))
        return a

    def normalize(self, fx, fy, fz, f0, kin=1):
        _lib.check(_lib.load().ghx_normalize(self._h, *map(self._c, (fx, fy, fz)), float(f0), kin))

    def sharp(self, a, b, delta):
        _lib.check(_lib.load().ghx_sharp(self._h, self._c(a), self._c(b), float(delta)))
        return b

    # ---- analysis/trunc3D.fpp ---------------------------------------------
    def truncate(self, c, small, out=None):
        """Spectral truncation of c onto the smaller Grid `small` (trunc3D)."""
        out = small.czeros() if out is None else out
        _lib.check(_lib.load().ghx_truncate(self._h, self._c(c), small._h, small._c(out)))
        return out

    # ---- pseudospec3D_phd.f90 ---------------------------------------------
    def advect3(self, a, b, c, d, e):
        """e = FFT[-(A . grad d)] / N^6 (pseudospec3D_phd.f90:23-115)."""
        _lib.check(_lib.load().ghx_advect3(self._h, *map(self._c, (a, b, c, d, e))))
        return e

    # ---- pseudospec3D_mhd.f90 ---------------------------------------------
    def vector3(self, a, b, c, d, e, f, x, y, z):
        _lib.check(_lib.load().ghx_vector3(self._h, *map(self._c, (a, b, c, d, e, f, x, y, z))))
        return x, y, z

    def nonlin3(self, a, b, c, d, e, f, g, dir):
        _lib.check(_lib.load().ghx_nonlin3(self._h, *map(self._c, (a, b, c, d, e, f, g)), dir))
        return g

    def gauge3(self, a, b, c, g, dir):
        _lib.check(_lib.load().ghx_gauge3(self._h, *map(self._c, (a, b, c, g)), dir))
        return g

    # ---- IC / forcing -----------------------------------------------------
    def initialv_tg(self, kdn=1.0, kup=1.0, u0=0.5):
        v = [self.czeros() for _ in range(3)]
        _lib.check(_lib.load().ghx_initialv_tg(self._h, kdn, kup, u0, *map(self._c, v)))
        return v

    def initialfv_random(self, kdn, kup, f0, seed, out=None):
        f = [self.czeros() for _ in range(3)] if out is None else out
        s = C.c_int32(seed)
        _lib.check(_lib.load().ghx_initialfv_random(self._h, kdn, kup, f0, C.byref(s), *map(self._c, f)))
        return f, s.value

    def initialv_k4exp(self, k0, u0, seed, out=None):
        v = [self.czeros() for _ in range(3)] if out is None else out
        s = C.c_int32(seed)
        _lib.check(_lib.load().ghx_initialv_k4exp(self._h, k0, u0, C.byref(s), *map(self._c, v)))
        return v, s.value

    # ---- RK ----------------------------------------------------------------
    def hd_rkstep2(self, o, dt, nu, v_in, C_, f, v_out, omegaz=None):
        """include/hd_rkstep2.f90 for one value of o, fused on the GPU; with
        omegaz, include/roth_rkstep2.f90 (a frame rotating about z)."""
        L_ = _lib.load()
        pv, _a = _lib.ptr_array([t.data_ptr() for t in map(lambda t: t, v_in)])
        pc, _b = _lib.ptr_array([t.data_ptr() for t in C_])
        po, _d = _lib.ptr_array([t.data_ptr() for t in v_out])
        for t in list(v_in) + list(C_) + list(v_out) + (list(f) if f is not None else []):
            _dptr(t, torch.complex128, self.cshape)
        if f is not None:
            pf, _c = _lib.ptr_array([t.data_ptr() for t in f])
        else:
            pf = None
        if omegaz is None:
            _lib.check(L_.ghx_hd_rkstep2(self._h, int(o), float(dt), float(nu), pv, pc, pf, po))
        else:
            _lib.check(L_.ghx_roth_rkstep2(self._h, int(o), float(dt), float(nu), float(omegaz), pv, pc, pf, po))


    def _ptrs(self, fields):
        if fields is None:
            return None, None
        for t in fields:
            _dptr(t, torch.complex128, self.cshape)
        return _lib.ptr_array([t.data_ptr() for t in fields])

    def mhd_rkstep2(self, o, dt, nu, mu, v_in, C_, f, m, v_out):
        """include/mhd_rkstep2.f90 for one value of o: fields (v_x, v_y, v_z,
        a_x, a_y, a_z); f and m the forcings (3 each, or both None)."""
        pv, _a = self._ptrs(v_in)
        pc, _b = self._ptrs(C_)
        pf, _c = self._ptrs(f)
        pm, _d = self._ptrs(m)
        po, _e = self._ptrs(v_out)
        _lib.check(_lib.load().ghx_mhd_rkstep2(self._h, int(o), float(dt), float(nu), float(mu), pv, pc, pf, pm, po))

    def bouss_rkstep2(self, o, dt, nu, kappa, xmom, xtemp, v_in, C_, f, v_out, omegaz=None, ord=None):
        """include/bouss_rkstep2.f90 (rotbouss_rkstep2.f90 with omegaz) for one
        value of o: fields (v_x, v_y, v_z, th); f the 4 forcings or None.
        With ord = 3, o = 1 and no rotation, the correction of :73-145 follows
        (v_out's velocity may then not alias C_)."""
        pv, _a = self._ptrs(v_in)
        pc, _b = self._ptrs(C_)
        pf, _c = self._ptrs(f)
        po, _e = self._ptrs(v_out)
        L_ = _lib.load()
        if omegaz is None:
            _lib.check(L_.ghx_bouss_rkstep2(self._h, int(o), float(dt), float(nu), float(kappa), float(xmom),
                                            float(xtemp), pv, pc, pf, po))
            if ord == 3 and o <= 1:
                pu, _g = self._ptrs(list(v_out)[:3])
                _lib.check(L_.ghx_bouss_ord3(self._h, float(dt), float(xmom), pc, pu))
        else:
            _lib.check(L_.ghx_rotbouss_rkstep2(self._h, int(o), float(dt), float(nu), float(kappa), float(xmom),
                                               float(xtemp), float(omegaz), pv, pc, pf, po))


def randu(seed):
    """randu (pseudospec3D_mod.fpp:130-148); returns (value, new seed)."""
    s = C.c_int32(seed)
    v = _lib.load().ghx_randu(C.byref(s))
    return v, s.value
