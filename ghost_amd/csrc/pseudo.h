// pseudo.h -- pointwise spectral operators and reductions (host interface).
#pragma once
#include "ghx_common.h"

namespace ghx {

// Elementwise operators on spectral fields s[kx][ky][kz] (GHOST C(nz,ny,nx/2+1)).
hipError_t op_derivk3(const GridK &g, hipStream_t st, const cplx *a, cplx *b, int dir);
hipError_t op_laplak3(const GridK &g, hipStream_t st, const cplx *a, cplx *b);
hipError_t op_rotor3(const GridK &g, hipStream_t st, const cplx *a, const cplx *b, cplx *c, int dir);
hipError_t op_nonlhd3(const GridK &g, hipStream_t st, const cplx *a, const cplx *b, const cplx *c, cplx *o, int dir);
hipError_t op_nonlin3(const GridK &g, hipStream_t st, const cplx *a, const cplx *b, const cplx *c,
                      const cplx *d, const cplx *e, const cplx *f, cplx *o, int dir);
hipError_t op_gauge3(const GridK &g, hipStream_t st, const cplx *a, const cplx *b, const cplx *c, cplx *o, int dir);
hipError_t op_phaseshift(const GridK &g, hipStream_t st, cplx *a, double cre, double cim);
hipError_t op_scale(const GridK &g, hipStream_t st, cplx *a, const cplx *src, double s);  // a = src*s
// a = a*f0/sqrt(e) (module_dns.f90 normalize) for three fields
hipError_t op_normalize_apply(const GridK &g, hipStream_t st, cplx *x, cplx *y, cplx *z, double f0, double e);
hipError_t op_sharp(const GridK &g, hipStream_t st, const cplx *a, cplx *b, double kc2);
// trunc3D's spectral truncation of c (grid g) onto the smaller grid t
// learned LES: subgrid stress products on the fine grid and the projected
// subgrid force (f0 and tau may be null; v and nuk, when given, add the
// spectral viscosity -nuk[shell - 1] kk2 v for shells 1..nb of width dkk)
hipError_t op_sgs_products(const GridK &g, hipStream_t st, const double *const *u, const double *const *ub, double s2,
                           double *const *out);
hipError_t op_sgs_force(const GridK &g, hipStream_t st, const cplx *const *tau, const cplx *const *f0,
                        cplx *const *out, const cplx *const *v = nullptr, const double *nuk = nullptr, int nb = 0,
                        double dkk = 1.0);
hipError_t op_truncate(const GridK &g, const GridK &t, hipStream_t st, const cplx *c, cplx *out);
hipError_t op_zero(hipStream_t st, void *p, size_t bytes);
// real-space cross product r7=(r2 r6-r5 r3)tmp ... in place (prodre3's loop)
hipError_t op_cross_real(const GridK &g, hipStream_t st, double *r1, double *r2, double *r3,
                         const double *r4, const double *r5, const double *r6, double *r7);
// RK update of include/hd_rkstep2.f90 for one component over all modes:
// v = C + dt*(nu*lap + N + f)*rmp where kept, else 0; lap given (already -kk2 v).
hipError_t op_rk_update(const GridK &g, hipStream_t st, cplx *v, const cplx *c, const cplx *lap,
                        const cplx *nl, const cplx *f, double nu, double dt, double rmp);
// ord = 3 correction of include/bouss_rkstep2.f90:73-145, pointwise steps.
// n = nonlhd3 of (w0, w1, w2 + xmom th); mode 0: out = lap(c) + n on kept
// modes, else 0; mode 1: out = n (c unused, may be null).
hipError_t op_bouss3_nonl(const GridK &g, hipStream_t st, int mode, const cplx *const *w, const cplx *th,
                          double xmom, const cplx *const *c, cplx *const *out);
// v = v + 2 x rmp on every mode (three fields)
hipError_t op_bouss3_add(const GridK &g, hipStream_t st, cplx *const *v, const cplx *const *x, double rmp);
// TG velocity in real space (examples/initialv.f90_tg:37-56)
hipError_t op_tg_real(const GridK &g, hipStream_t st, double *r1, double *r2, int kdn, int kup);

// Reductions.  kind: 0 = |a|^2+|b|^2+|c|^2, 1 = Re(a d* + b e* + c f*),
// 2 = that * kk2, 3 = |i kx a + i ky b + i kz c|^2 (divergence),
// 4 = number of modes outside tiny <= kn2 <= kmax where (a,b,c) is non-zero.
// Each mode is weighted 1 on the kx=0 plane, 2 elsewhere, times 1/N^6.
// `work` must hold reduce_work_doubles() doubles.  Result to *out (host).
size_t reduce_work_doubles();
hipError_t op_reduce(const GridK &g, hipStream_t st, int kind, const cplx *a, const cplx *b, const cplx *c,
                     const cplx *d, const cplx *e, const cplx *f, double *work, double *out);
// spectrumc binning (pseudospec3D_hd.f90:1190-1415); kind 0: energy of (a,b,c),
// kind 1: helicity Re(a d* + b e* + c f*).  nb bins; out[nb] host.
size_t spectrum_work_doubles(const GridK &g, int nb);
hipError_t op_spectrum(const GridK &g, hipStream_t st, int kind, double dkk, int nb, const cplx *a, const cplx *b,
                       const cplx *c, const cplx *d, const cplx *e, const cplx *f, double *work, double *out, double off = (double)0.501f);
// Axisymmetric spectra (pseudospec3D_rot.f90): bins 0 = planes of |kz|,
// 1 = cylinders of k_perp; kind 0 energy, 1 Re dot, 2 kk2-weighted dot.
// out: 3 nb host doubles (total, xy part, z part); work: axispec_work_doubles.
size_t axispec_work_doubles(const GridK &g, int nb);
hipError_t op_axispec(const GridK &g, hipStream_t st, int bins, int kind, int nb, double dkk, double lz,
                      const cplx *a, const cplx *b, const cplx *c, const cplx *d, const cplx *e, const cplx *f,
                      double *work, double *out);
// spec2D: np * nzb host doubles (kperp fastest); work holds np * nzb doubles.
hipError_t op_spec2d(const GridK &g, hipStream_t st, int kind, int np, int nzb, double dkk, double lz, const cplx *a,
                     const cplx *b, const cplx *c, const cplx *d, const cplx *e, const cplx *f, double *work,
                     double *out);
// plane means of difucx (dir 1) / difucz (dir 3) of two real fields: 3 n host
// doubles; work holds 3 n doubles.
hipError_t op_plane_means(const GridK &g, hipStream_t st, const double *r1, const double *r2, int dir, double tmp,
                          double tmq, double *work, double *out);
// havgcomp's per-z-plane sums of real fields (kinds in pseudo.hip): nz host
// doubles, each times scale; work holds nz doubles.
hipError_t op_zplane(const GridK &g, hipStream_t st, int kind, const double *r1, const double *r2, const double *r3,
                     const double *r4, double fo, double bv, double scale, double *work, double *out);
hipError_t op_spectr1d(const GridK &g, hipStream_t st, int dir, double ldir, int nb, const cplx *a, double *work,
                       double *out);
// max over real points of sqrt(r1^2+r2^2+r3^2) (maxabs)
hipError_t op_maxnorm(const GridK &g, hipStream_t st, const double *r1, const double *r2, const double *r3,
                      double *work, double *out);

}  // namespace ghx
