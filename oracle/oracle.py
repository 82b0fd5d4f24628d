"""ctypes binding of the CPU oracle (``libghost_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by the test-suite, the smoke check and the
CPU-baseline leg of the benchmark.  The product package never imports it.

Array conventions (GHOST's, reference 3D/src/fftp-3/fftp3D.fpp:204-431):
  real fields     numpy float64    shape (nz, ny, nx)        == Fortran R(nx,ny,nz)
  spectral fields numpy complex128 shape (nx//2+1, ny, nz)   == Fortran C(nz,ny,nx/2+1)
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libghost_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class _Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nxh", C.c_int), ("anis", C.c_int),
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("Dkx", C.c_double), ("Dky", C.c_double), ("Dkz", C.c_double), ("Dkk", C.c_double),
        ("kmax", C.c_double), ("tiny", C.c_double), ("nmax", C.c_int),
        ("kx", C.POINTER(C.c_double)), ("ky", C.POINTER(C.c_double)), ("kz", C.POINTER(C.c_double)),
        ("kn2", C.POINTER(C.c_double)), ("kk2", C.POINTER(C.c_double)), ("fft", C.c_void_p),
    ]


_P = C.c_void_p
_G = C.POINTER(_Grid)


def _declare(L):
    L.og_grid_new.restype = _G
    L.og_grid_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                              C.c_int, C.c_double]
    L.og_grid_free.argtypes = [_G]
    L.og_set_threads.argtypes = [C.c_int]
    L.og_get_threads.restype = C.c_int
    L.og_r2c.argtypes = [_G, _P, _P]
    L.og_c2r.argtypes = [_G, _P, _P]
    L.og_fft1d.argtypes = [C.c_int, _P, C.c_int]
    L.og_derivk3.argtypes = [_G, _P, _P, C.c_int]
    L.og_laplak3.argtypes = [_G, _P, _P]
    L.og_rotor3.argtypes = [_G, _P, _P, _P, C.c_int]
    L.og_prodre3.argtypes = [_G, _P, _P, _P, _P, _P, _P]
    L.og_nonlhd3.argtypes = [_G, _P, _P, _P, _P, C.c_int]
    L.og_nonlin3.argtypes = [_G, _P, _P, _P, _P, _P, _P, _P, C.c_int]
    L.og_gauge3.argtypes = [_G, _P, _P, _P, _P, C.c_int]
    L.og_vector3.argtypes = [_G, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    for name in ("og_energy", "og_maxabs"):
        getattr(L, name).restype = C.c_double
        getattr(L, name).argtypes = [_G, _P, _P, _P, C.c_int]
    L.og_helicity.restype = C.c_double
    L.og_helicity.argtypes = [_G, _P, _P, _P]
    L.og_divergence.restype = C.c_double
    L.og_divergence.argtypes = [_G, _P, _P, _P]
    L.og_cross.restype = C.c_double
    L.og_cross.argtypes = [_G, _P, _P, _P, _P, _P, _P, C.c_int]
    L.og_spectrumc.argtypes = [_G, _P, _P, _P, C.c_int, C.c_int, _P, _P]
    L.og_phaseshift.argtypes = [_G, _P, C.c_double, C.c_double]
    L.og_normalize.argtypes = [_G, _P, _P, _P, C.c_double, C.c_int]
    L.og_randu.restype = C.c_double
    L.og_randu.argtypes = [C.POINTER(C.c_int32)]
    L.og_initialv_tg.argtypes = [_G, C.c_double, C.c_double, C.c_double, _P, _P, _P]
    L.og_initialfv_random.argtypes = [_G, C.c_double, C.c_double, C.c_double,
                                      C.POINTER(C.c_int32), _P, _P, _P]
    L.og_initialv_k4exp.argtypes = [_G, C.c_double, C.c_double, C.POINTER(C.c_int32), _P, _P, _P]
    L.og_hd_step.argtypes = [_G, C.c_int, C.c_double, C.c_double, _P, _P, _P, _P, _P, _P]
    L.og_hd_substep.argtypes = [_G, C.c_int, C.c_double, C.c_double, _P, _P, _P, _P, _P, _P,
                                _P, _P, _P]
    L.og_mhd_step.argtypes = [_G, C.c_int, C.c_double, C.c_double, C.c_double] + [_P] * 12
    L.og_sharp.argtypes = [_G, _P, _P, C.c_double]
    L.og_advect3.argtypes = [_G, _P, _P, _P, _P, _P]
    L.og_truncate.argtypes = [_G, _P, _G, _P]
    L.og_bouss_step.argtypes = [_G, C.c_int] + [C.c_double] * 5 + [_P] * 8
    L.og_rotbouss_step.argtypes = [_G, C.c_int] + [C.c_double] * 6 + [_P] * 8
    L.og_roth_step.argtypes = [_G, C.c_int] + [C.c_double] * 3 + [_P] * 6
    L.og_roth_substep.argtypes = [_G, C.c_int] + [C.c_double] * 3 + [_P] * 9
    L.og_spline_coef.argtypes = [_G, _P, _P, C.c_int]
    L.og_spline_interp.argtypes = [_G, _P, C.c_long, _P, _P, _P, C.c_int, _P]
    L.og_tracer_stage.argtypes = [_G, _P, _P, _P, C.c_long, _P, _P, _P, _P, _P, _P, C.c_double,
                                  C.c_double, C.c_int, C.c_int, _P]
    L.og_hd_step_tracers.argtypes = [_G, C.c_int, C.c_double, C.c_double] + [_P] * 6 + \
        [C.c_long, _P, _P, _P, C.c_int, C.c_int]
    L.og_entrans.argtypes = [_G] + [_P] * 6 + [C.c_int, _P]
    L.og_heltrans.argtypes = [_G] + [_P] * 6 + [_P]
    L.og_crosspecc.argtypes = [_G] + [_P] * 6 + [C.c_double, _P]
    L.og_spectr1d.argtypes = [_G, _P, C.c_int, _P]
    L.og_mhdcheck.argtypes = [_G] + [_P] * 6 + [C.c_int, C.c_int, C.c_int, _P]
    L.og_variance.restype = C.c_double
    L.og_variance.argtypes = [_G, _P, C.c_int]
    L.og_product.restype = C.c_double
    L.og_product.argtypes = [_G, _P, _P]
    L.og_spectrscc.argtypes = [_G, _P, C.c_double, _P]
    L.og_sctrans.argtypes = [_G, _P, _P, _P]
    L.og_difuc.argtypes = [_G, _P, _P, C.c_int, _P]
    L.og_nmaxperp.restype = C.c_int
    L.og_nmaxperp.argtypes = [_G]
    L.og_axispec.argtypes = [_G, C.c_int, C.c_int] + [_P] * 9
    L.og_spec2d.argtypes = [_G, C.c_int] + [_P] * 7
    L.og_havgcomp.argtypes = [_G] + [_P] * 4 + [C.c_double, C.c_double, C.c_int, _P]
    L.og_spectpv.argtypes = [_G] + [_P] * 5
    L.og_mlp_forward.argtypes = [_P, C.c_long, _P, C.c_int, _P]
    L.og_mlp_grad.restype = C.c_double
    L.og_mlp_grad.argtypes = [_P, C.c_long, _P, _P, C.c_int, _P]


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def set_threads(n):
    lib().og_set_threads(int(n))


def get_threads():
    return lib().og_get_threads()


def randu(seed):
    """Draw one value; returns (value, new_seed)."""
    s = C.c_int32(seed)
    v = lib().og_randu(C.byref(s))
    return v, s.value


def fft1d(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().og_fft1d(x.size, _p(x), int(sign))
    return x


class Grid:
    """GHOST grid: wavenumbers, dealias bounds and FFT plans (main3D.fpp:1727-1800)."""

    def __init__(self, nx, ny=None, nz=None, L=(1.0, 1.0, 1.0), arbsize=False, Dkk=0.0):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        self._g = lib().og_grid_new(nx, ny, nz, L[0], L[1], L[2], int(bool(arbsize)), Dkk)
        g = self._g.contents
        self.nx, self.ny, self.nz, self.nxh = g.nx, g.ny, g.nz, g.nxh
        self.kmax, self.tiny, self.nmax, self.Dkk = g.kmax, g.tiny, g.nmax, g.Dkk
        self.anis = g.anis
        self.kx = np.ctypeslib.as_array(g.kx, shape=(nx,)).copy()
        self.ky = np.ctypeslib.as_array(g.ky, shape=(ny,)).copy()
        self.kz = np.ctypeslib.as_array(g.kz, shape=(nz,)).copy()
        self.kn2 = np.ctypeslib.as_array(g.kn2, shape=(self.nxh, ny, nz)).copy()
        self.kk2 = np.ctypeslib.as_array(g.kk2, shape=(self.nxh, ny, nz)).copy()

    def __del__(self):
        try:
            lib().og_grid_free(self._g)
        except Exception:
            pass

    # shapes / allocation
    @property
    def cshape(self):
        return (self.nxh, self.ny, self.nz)

    @property
    def rshape(self):
        return (self.nz, self.ny, self.nx)

    def czeros(self, n=None):
        return np.zeros(self.cshape if n is None else (n,) + self.cshape, np.complex128)

    def rzeros(self):
        return np.zeros(self.rshape, np.float64)

    @property
    def nbins(self):
        return self.nmax // 2 + 1

    # FFT
    def r2c(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = self.czeros()
        lib().og_r2c(self._g, _p(r), _p(out))
        return out

    def c2r(self, c):
        c = np.ascontiguousarray(c, dtype=np.complex128)
        out = self.rzeros()
        lib().og_c2r(self._g, _p(c), _p(out))
        return out

    # operators
    def derivk3(self, a, d):
        b = self.czeros(); lib().og_derivk3(self._g, _p(a), _p(b), d); return b

    def laplak3(self, a):
        b = self.czeros(); lib().og_laplak3(self._g, _p(a), _p(b)); return b

    def rotor3(self, a, b, d):
        c = self.czeros(); lib().og_rotor3(self._g, _p(a), _p(b), _p(c), d); return c

    def curl(self, vx, vy, vz):
        return self.rotor3(vy, vz, 1), self.rotor3(vx, vz, 2), self.rotor3(vx, vy, 3)

    def prodre3(self, a, b, c):
        d, e, f = self.czeros(), self.czeros(), self.czeros()
        lib().og_prodre3(self._g, _p(a), _p(b), _p(c), _p(d), _p(e), _p(f))
        return d, e, f

    def nonlhd3(self, a, b, c, d):
        g = self.czeros(); lib().og_nonlhd3(self._g, _p(a), _p(b), _p(c), _p(g), d); return g

    def nonlin3(self, a, b, c, d, e, f, dr):
        g = self.czeros()
        lib().og_nonlin3(self._g, _p(a), _p(b), _p(c), _p(d), _p(e), _p(f), _p(g), dr)
        return g

    def gauge3(self, a, b, c, d):
        g = self.czeros(); lib().og_gauge3(self._g, _p(a), _p(b), _p(c), _p(g), d); return g

    def vector3(self, a, b, c, d, e, f):
        x, y, z = self.czeros(), self.czeros(), self.czeros()
        lib().og_vector3(self._g, *map(_p, (a, b, c, d, e, f, x, y, z)))
        return x, y, z

    def energy(self, a, b, c, kin=1):
        return lib().og_energy(self._g, _p(a), _p(b), _p(c), kin)

    def helicity(self, a, b, c):
        return lib().og_helicity(self._g, _p(a), _p(b), _p(c))

    def cross(self, a, b, c, d, e, f, kin=1):
        return lib().og_cross(self._g, *map(_p, (a, b, c, d, e, f)), kin)

    def divergence(self, a, b, c):
        return lib().og_divergence(self._g, _p(a), _p(b), _p(c))

    def maxabs(self, a, b, c, kin):
        return lib().og_maxabs(self._g, _p(a), _p(b), _p(c), kin)

    def spectrumc(self, a, b, c, kin=1, hel=0):
        Ek = np.zeros(self.nbins)
        Hk = np.zeros(self.nbins)
        lib().og_spectrumc(self._g, _p(a), _p(b), _p(c), kin, hel, _p(Ek), _p(Hk))
        return Ek, Hk

    def entrans(self, a, b, c, d, e, f, kin=1):
        out = np.zeros(self.nbins)
        lib().og_entrans(self._g, *map(_p, (a, b, c, d, e, f)), int(kin), _p(out))
        return out

    def heltrans(self, a, b, c, d, e, f):
        out = np.zeros(self.nbins)
        lib().og_heltrans(self._g, *map(_p, (a, b, c, d, e, f)), _p(out))
        return out

    def crosspecc(self, a, b, c, d, e, f, shift=0.0):
        out = np.zeros(self.nbins)
        lib().og_crosspecc(self._g, *map(_p, (a, b, c, d, e, f)), float(shift), _p(out))
        return out

    def spectr1d(self, a, dir):
        out = np.zeros(self.nbins)
        lib().og_spectr1d(self._g, _p(a), int(dir), _p(out))
        return out

    # -- scalar diagnostics (pseudospec3D_phd.f90) ----------------------------
    def variance(self, a, kin=1):
        return lib().og_variance(self._g, _p(a), int(kin))

    def product(self, a, b):
        return lib().og_product(self._g, _p(a), _p(b))

    def pscheck(self, a, b):
        """pscheck (pseudospec3D_phd.f90:286-333): <th^2>, <|grad th|^2>, <th s>."""
        return np.array([self.variance(a, 1), self.variance(a, 0), self.product(a, b)])

    def spectrscc(self, a, shift=0.0):
        out = np.zeros(self.nbins)
        lib().og_spectrscc(self._g, _p(a), float(shift), _p(out))
        return out

    def sctrans(self, a, b):
        out = np.zeros(self.nbins)
        lib().og_sctrans(self._g, _p(a), _p(b), _p(out))
        return out

    def difuc(self, a, b, dir):
        n = self.nx if dir == 1 else self.nz
        out = np.zeros(3 * n)
        lib().og_difuc(self._g, _p(a), _p(b), int(dir), _p(out))
        return out.reshape(3, n)

    # -- axisymmetric spectra (pseudospec3D_rot.f90) --------------------------
    @property
    def nbpara(self):
        return self.nz // 2 + 1

    @property
    def nbperp(self):
        return lib().og_nmaxperp(self._g) // 2 + 1

    def axispec(self, bins, kind, a, b, c, d=None, e=None, f=None):
        nb = self.nbpara if bins == 0 else self.nbperp
        out = np.zeros((3, nb))
        z = [None] * 3 if d is None else [d, e, f]
        lib().og_axispec(self._g, int(bins), int(kind), _p(a), _p(b), _p(c), *[_p(x) if x is not None else None
                                                                              for x in z],
                         _p(out[0]), _p(out[1]), _p(out[2]))
        return out

    def _axis_ek(self, bins, a, b, c, kin, hel):
        """specpara (bins 0) / specperp (bins 1): energy (kin 1), energy of
        curl (a,b,c) (kin 0), skipped (kin 2); helicity a . curl a (hel 1)."""
        cu = self.curl(a, b, c) if (kin == 0 or hel == 1) else None
        E = None
        if kin == 1:
            E = self.axispec(bins, 0, a, b, c)
        elif kin == 0:
            E = self.axispec(bins, 0, *cu)
        H = self.axispec(bins, 1, a, b, c, *cu) if hel == 1 else None
        return E, H

    def specpara(self, a, b, c, kin=1, hel=0):
        return self._axis_ek(0, a, b, c, kin, hel)

    def specperp(self, a, b, c, kin=1, hel=0):
        return self._axis_ek(1, a, b, c, kin, hel)

    def entpara(self, a, b, c, d, e, f, kin=1):
        return self.axispec(0, 2 if kin == 0 else 1, a, b, c, d, e, f)[0]

    def entperp(self, a, b, c, d, e, f, kin=1):
        return self.axispec(1, 2 if kin == 0 else 1, a, b, c, d, e, f)[0]

    def heltpara(self, a, b, c, d, e, f):
        return self.axispec(0, 1, *self.curl(a, b, c), d, e, f)[0]

    def heltperp(self, a, b, c, d, e, f):
        return self.axispec(1, 1, *self.curl(a, b, c), d, e, f)[0]

    def spec2d(self, a, b, c, kin=1, hel=0):
        """spec2D: E(kperp, kpara) as an array [kpara][kperp] (Fortran order of Ek(kperp, kpara))."""
        cu = self.curl(a, b, c) if (kin == 0 or hel == 1) else None
        shape = (self.nbpara, self.nbperp)
        E = H = None
        if kin <= 1:
            E = np.zeros(shape)
            src = (a, b, c) if kin == 1 else cu
            lib().og_spec2d(self._g, 0, *map(_p, src), None, None, None, _p(E))
        if hel == 1:
            H = np.zeros(shape)
            lib().og_spec2d(self._g, 1, _p(a), _p(b), _p(c), *map(_p, cu), _p(H))
        return E, H

    # -- Boussinesq diagnostics (pseudospec3D_bouss.f90) ----------------------
    def havgcomp(self, u, v, w, s, fo, bv, itype):
        out = np.zeros(self.nz)
        lib().og_havgcomp(self._g, _p(u), _p(v), _p(w), _p(s), float(fo), float(bv), int(itype), _p(out))
        return out

    def tbouss(self, u, v, w, s, fo, bv):
        """tbouss (:612-709): (avg, max, min) over z of havgcomp itype 0..10."""
        prof = [self.havgcomp(u, v, w, s, fo, bv, i) for i in range(11)]
        return (np.array([np.sum(p) / self.nz for p in prof]), np.array([np.max(p) for p in prof]),
                np.array([np.min(p) for p in prof]))

    def spectpv(self, u, v, w, s):
        out = np.zeros(self.nz // 2 + 1)
        lib().og_spectpv(self._g, _p(u), _p(v), _p(w), _p(s), _p(out))
        return out

    def hdcheck(self, a, b, c, d, e, f, hel=1, chk=1):
        """hdcheck (pseudospec3D_hd.f90:984-1105) from its parts: {eng, ens, pot, khe, div}."""
        return np.array([self.energy(a, b, c, 1), self.energy(a, b, c, 0), self.cross(a, b, c, d, e, f, 1),
                         self.helicity(a, b, c) if hel == 1 else 0.0, self.divergence(a, b, c) if chk == 1 else 0.0])

    def mhdcheck(self, a, b, c, ma, mb, mc, hel=1, crs=1, chk=1):
        out = np.zeros(11)
        lib().og_mhdcheck(self._g, *map(_p, (a, b, c, ma, mb, mc)), int(hel), int(crs), int(chk), _p(out))
        return out

    def phaseshift(self, a, phase_c):
        a = a.copy()
        lib().og_phaseshift(self._g, _p(a), float(np.real(phase_c)), float(np.imag(phase_c)))
        return a

    def normalize(self, fx, fy, fz, f0, kin=1):
        fx, fy, fz = fx.copy(), fy.copy(), fz.copy()
        lib().og_normalize(self._g, _p(fx), _p(fy), _p(fz), f0, kin)
        return fx, fy, fz

    def sharp(self, a, delta):
        b = self.czeros(); lib().og_sharp(self._g, _p(a), _p(b), delta); return b

    # initial conditions / forcing
    def initialv_tg(self, kdn=1.0, kup=1.0, u0=0.5):
        v = [self.czeros() for _ in range(3)]
        lib().og_initialv_tg(self._g, kdn, kup, u0, *map(_p, v))
        return v

    def initialfv_random(self, kdn, kup, f0, seed):
        s = C.c_int32(seed)
        f = [self.czeros() for _ in range(3)]
        lib().og_initialfv_random(self._g, kdn, kup, f0, C.byref(s), *map(_p, f))
        return f, s.value

    def initialv_k4exp(self, k0, u0, seed):
        s = C.c_int32(seed)
        v = [self.czeros() for _ in range(3)]
        lib().og_initialv_k4exp(self._g, k0, u0, C.byref(s), *map(_p, v))
        return v, s.value

    # time stepping
    def hd_step(self, v, dt, nu, f=None, ord=2):
        """One RK step in place on the list v=[vx,vy,vz]."""
        fp = [None] * 3 if f is None else list(map(_p, f))
        lib().og_hd_step(self._g, ord, dt, nu, *map(_p, v), *fp)

    def roth_step(self, v, dt, nu, omegaz, f=None, ord=2):
        """One RK step of HD in a frame rotating about z, in place."""
        fp = [None] * 3 if f is None else list(map(_p, f))
        lib().og_roth_step(self._g, ord, dt, nu, omegaz, *map(_p, v), *fp)

    def roth_substep(self, v, c, o, dt, nu, omegaz, f=None):
        fp = [None] * 3 if f is None else list(map(_p, f))
        lib().og_roth_substep(self._g, o, dt, nu, omegaz, *map(_p, v), *map(_p, c), *fp)

    def hd_substep(self, v, c, o, dt, nu, f=None):
        fp = [None] * 3 if f is None else list(map(_p, f))
        lib().og_hd_substep(self._g, o, dt, nu, *map(_p, v), *map(_p, c), *fp)

    def truncate(self, c, small):
        """trunc3D's spectral truncation of c onto the smaller Grid `small`."""
        out = small.czeros()
        lib().og_truncate(self._g, _p(np.ascontiguousarray(c)), small._g, _p(out))
        return out

    # -- Boussinesq -----------------------------------------------------------
    def advect3(self, a, b, c, d):
        """FFT[-(A . grad d)] / N^6 (pseudospec3D_phd.f90:23-115)."""
        e = self.czeros()
        lib().og_advect3(self._g, _p(a), _p(b), _p(c), _p(d), _p(e))
        return e

    def bouss_step(self, v, th, dt, nu, kappa, xmom, xtemp, f=None, fs=None, ord=2):
        """One Boussinesq RK step in place on v = [vx, vy, vz] and th."""
        fp = [None] * 3 if f is None else list(map(_p, f))
        lib().og_bouss_step(self._g, ord, dt, nu, kappa, xmom, xtemp, *map(_p, v), _p(th), *fp, _p(fs))

    def rotbouss_step(self, v, th, dt, nu, kappa, xmom, xtemp, omegaz, f=None, fs=None, ord=2):
        """One rotating-Boussinesq RK step in place on v = [vx, vy, vz] and th."""
        fp = [None] * 3 if f is None else list(map(_p, f))
        lib().og_rotbouss_step(self._g, ord, dt, nu, kappa, xmom, xtemp, omegaz, *map(_p, v), _p(th), *fp, _p(fs))

    # -- Lagrangian tracers (positions in grid units, float64 arrays) --------
    def spline_coef(self, field, ghost=True):
        """Spline control points of a real field (GPSplineInt_CompSpline3D)."""
        out = np.empty(self.rshape)
        lib().og_spline_coef(self._g, _p(np.ascontiguousarray(field, dtype=np.float64)), _p(out), int(ghost))
        return out

    def spline_interp(self, coef, px, py, pz, zcompat=True):
        """GPSplineInt_Interp3D of control points coef at (px, py, pz)."""
        px, py, pz = (np.ascontiguousarray(a, dtype=np.float64) for a in (px, py, pz))
        out = np.empty(len(px))
        lib().og_spline_interp(self._g, _p(np.ascontiguousarray(coef)), len(px), _p(px), _p(py), _p(pz),
                               int(zcompat), _p(out))
        return out

    def tracer_stage(self, v, pos, x0, dt, xk, zcompat=True, ghost_solve=True):
        """One tracer RK stage in place on pos = [px, py, pz]; returns the
        interpolated velocities (3, np)."""
        np_ = len(pos[0])
        lv = np.empty((3, np_))
        lib().og_tracer_stage(self._g, *map(_p, v), np_, *map(_p, pos), *map(_p, x0), dt, xk, int(zcompat),
                              int(ghost_solve), _p(lv))
        return lv

    def hd_step_tracers(self, v, pos, dt, nu, f=None, ord=2, zcompat=True, ghost_solve=True):
        """One RK step of the fluid (in place on v) and the tracers (in place on pos)."""
        fp = [None] * 3 if f is None else list(map(_p, f))
        lib().og_hd_step_tracers(self._g, ord, dt, nu, *map(_p, v), *fp, len(pos[0]), *map(_p, pos),
                                 int(zcompat), int(ghost_solve))

    def mhd_step(self, v, a, dt, nu, mu, f=None, m=None, ord=2):
        fp = [None] * 3 if f is None else list(map(_p, f))
        mp = [None] * 3 if m is None else list(map(_p, m))
        lib().og_mhd_step(self._g, ord, dt, nu, mu, *map(_p, v), *map(_p, a), *fp, *mp)


# ---- the closure network of the learned LES (no GHOST counterpart) ----------
MLP_NPAR = 5254


def mlp_forward(params, X, bf16=True):
    """Outputs (n x 6, double) of the 10-64-64-6 GELU network for inputs X (n x 10)."""
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros((X.shape[0], 6))
    lib().og_mlp_forward(_p(P), X.shape[0], _p(X), int(bool(bf16)), _p(Y))
    return Y


def mlp_grad(params, X, T, bf16=True):
    """(loss, gradient) of the mean squared error over the n x 6 outputs."""
    P = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    g = np.zeros(MLP_NPAR)
    loss = lib().og_mlp_grad(_p(P), X.shape[0], _p(X), _p(T), int(bool(bf16)), _p(g))
    return loss, g
