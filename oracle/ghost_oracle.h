/*
 * ghost_oracle.h -- CPU restatement of GHOST's pseudospectral hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against, and the timed CPU baseline.  Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; the product library never links or calls it.
 *
 * Layouts follow GHOST exactly (reference 3D/src/fftp-3/fftp3D.fpp:204-431):
 *   real     r[nz][ny][nx]          (Fortran R(nx,ny,nz), x fastest)
 *   spectral c[nx/2+1][ny][nz]      (Fortran C(nz,ny,nx/2+1), kz fastest)
 * Transforms are unnormalised, forward sign -1, inverse sign +1.
 */
#ifndef GHOST_ORACLE_H
#define GHOST_ORACLE_H

#include <complex.h>
#include <stdint.h>

typedef double _Complex ocplx;

typedef struct og_grid {
  int nx, ny, nz;          /* grid points */
  int nxh;                 /* nx/2+1 */
  int anis;                /* 0: cubic, 1: anisotropic/arbitrary box */
  double Lx, Ly, Lz;       /* box lengths in units of 2*pi */
  double Dkx, Dky, Dkz, Dkk;
  double kmax, tiny;       /* dealias bounds on kn2 */
  int nmax;                /* spectrum uses bins 1..nmax/2+1 */
  double *kx, *ky, *kz;    /* wavenumbers (with box scaling), sizes nx,ny,nz */
  double *kn2;             /* [nxh][ny][nz] dimensionless, for dealiasing */
  double *kk2;             /* [nxh][ny][nz] physical k^2 (== kn2 when cubic) */
  void *fft;               /* private FFT plans */
} og_grid;

/* grid / wavenumbers: main3D.fpp:737-749, 881-909, 1727-1800 */
og_grid *og_grid_new(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                     int arbsize, double Dkk);
void og_grid_free(og_grid *g);
void og_set_threads(int n);
int og_get_threads(void);

/* fftp3d_real_to_complex / fftp3d_complex_to_real (fftp3D.fpp:204-431).
 * og_c2r preserves its input (GHOST's destroys it). */
void og_r2c(const og_grid *g, const double *in, ocplx *out);
void og_c2r(const og_grid *g, const ocplx *in, double *out);
/* 1D complex transform of length n (sign -1 forward, +1 inverse), unnormalised */
void og_fft1d(int n, ocplx *x, int sign);

/* pseudospec3D_hd.f90 */
void og_derivk3(const og_grid *g, const ocplx *a, ocplx *b, int dir);          /* :22  */
void og_laplak3(const og_grid *g, const ocplx *a, ocplx *b);                   /* :91  */
void og_rotor3(const og_grid *g, const ocplx *a, const ocplx *b, ocplx *c, int dir); /* :124 */
void og_prodre3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c,
                ocplx *d, ocplx *e, ocplx *f);                                 /* :316 */
void og_nonlhd3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c,
                ocplx *gout, int dir);                                         /* :399 */
double og_energy(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, int kin); /* :512 */
double og_helicity(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c);        /* :665 */
double og_cross(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c,
                const ocplx *d, const ocplx *e, const ocplx *f, int kin);      /* :787 */
double og_maxabs(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, int kin); /* :910 */
/* spectrumc (:1190): Ek,Hk have nmax/2+1 entries; kin in {0,1,2}, hel in {0,1} */
void og_spectrumc(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c,
                  int kin, int hel, double *Ek, double *Hk);
/* divergence term of hdcheck (:1036-1071) */
double og_divergence(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c);
void og_phaseshift(const og_grid *g, ocplx *a, double cre, double cim);       /* :2084 */

/* module_dns.f90:13-41 */
void og_normalize(const og_grid *g, ocplx *fx, ocplx *fy, ocplx *fz, double f0, int kin);

/* pseudospec3D_mod.fpp:130-148 */
double og_randu(int32_t *idum);

/* IC / forcing generators (examples/initialv.f90_tg, examples/initialfv.f90_random,
 * and the k^4 exp(-2(k/k0)^2) random-phase field in the same fill order) */
void og_initialv_tg(const og_grid *g, double kdn, double kup, double u0,
                    ocplx *vx, ocplx *vy, ocplx *vz);
void og_initialfv_random(const og_grid *g, double kdn, double kup, double f0,
                         int32_t *seed, ocplx *fx, ocplx *fy, ocplx *fz);
void og_initialv_k4exp(const og_grid *g, double k0, double u0, int32_t *seed,
                       ocplx *vx, ocplx *vy, ocplx *vz);

/* HD step (main3D.fpp:3202-3231, include/hd_rkstep1.f90, hd_rkstep2.f90).
 * fx..fz may be NULL (no forcing). ord = RK order. */
void og_hd_step(const og_grid *g, int ord, double dt, double nu,
                ocplx *vx, ocplx *vy, ocplx *vz,
                const ocplx *fx, const ocplx *fy, const ocplx *fz);
/* one substep o of that loop, given C = state at step start */
void og_hd_substep(const og_grid *g, int o, double dt, double nu,
                   ocplx *vx, ocplx *vy, ocplx *vz,
                   const ocplx *c1, const ocplx *c2, const ocplx *c3,
                   const ocplx *fx, const ocplx *fy, const ocplx *fz);

/* HD in a frame rotating about z (include/roth_rkstep{1,2}.f90): the
 * Coriolis force enters the products before the projection */
void og_roth_step(const og_grid *g, int ord, double dt, double nu, double omegaz,
                  ocplx *vx, ocplx *vy, ocplx *vz,
                  const ocplx *fx, const ocplx *fy, const ocplx *fz);
void og_roth_substep(const og_grid *g, int o, double dt, double nu, double omegaz,
                     ocplx *vx, ocplx *vy, ocplx *vz,
                     const ocplx *c1, const ocplx *c2, const ocplx *c3,
                     const ocplx *fx, const ocplx *fy, const ocplx *fz);

/* MHD (pseudospec3D_mhd.f90, include/mhd_rkstep2.f90) */
void og_vector3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c,
                const ocplx *d, const ocplx *e, const ocplx *f,
                ocplx *x, ocplx *y, ocplx *z);                                  /* :22  */
void og_nonlin3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c,
                const ocplx *d, const ocplx *e, const ocplx *f, ocplx *gout, int dir); /* :108 */
void og_gauge3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c,
               ocplx *gout, int dir);                                          /* :230 */
void og_mhd_step(const og_grid *g, int ord, double dt, double nu, double mu,
                 ocplx *vx, ocplx *vy, ocplx *vz, ocplx *ax, ocplx *ay, ocplx *az,
                 const ocplx *fx, const ocplx *fy, const ocplx *fz,
                 const ocplx *mx, const ocplx *my, const ocplx *mz);

/* pseudospec3D_filt.f90:200-241 */
void og_sharp(const og_grid *g, const ocplx *a, ocplx *b, double delta);

/* Lagrangian tracers (lagpart/): spline control points (GPSplineInt_CompSpline3D,
 * ghost != 0 keeps GHOST's row-n update, see the source), interpolation at np
 * points (GPSplineInt_Interp3D, zcompat != 0 reproduces GHOST's z stencil),
 * one RK stage of the tracers (main3D.fpp:3233-3267 + GPart_StepRKK) and an HD
 * step with tracers. Positions in grid units. */
void og_spline_coef(const og_grid *g, const double *field, double *coef, int ghost);
void og_spline_interp(const og_grid *g, const double *coef, long np, const double *px, const double *py,
                      const double *pz, int zcompat, double *out);
void og_tracer_stage(const og_grid *g, const ocplx *vx, const ocplx *vy, const ocplx *vz, long np,
                     double *px, double *py, double *pz, const double *x0, const double *y0,
                     const double *z0, double dt, double xk, int zcompat, int ghost_solve, double *lv);
void og_hd_step_tracers(const og_grid *g, int ord, double dt, double nu, ocplx *vx, ocplx *vy, ocplx *vz,
                        const ocplx *fx, const ocplx *fy, const ocplx *fz, long np, double *px, double *py,
                        double *pz, int zcompat, int ghost_solve);

/* Boussinesq: advect3 (pseudospec3D_phd.f90:23-115) and the RK step of
 * include/bouss_rkstep{1,2}.f90 (state v, th; kappa diffusivity, xmom and
 * xtemp already multiplied by bvfreq as main3D.fpp:973-974 does). */
void og_advect3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                ocplx *e);
void og_bouss_substep(const og_grid *g, int o, double dt, double nu, double kappa, double xmom, double xtemp,
                      ocplx *vx, ocplx *vy, ocplx *vz, ocplx *th, const ocplx *c1, const ocplx *c2,
                      const ocplx *c3, const ocplx *c20, const ocplx *fx, const ocplx *fy, const ocplx *fz,
                      const ocplx *fs);
void og_bouss_step(const og_grid *g, int ord, double dt, double nu, double kappa, double xmom, double xtemp,
                   ocplx *vx, ocplx *vy, ocplx *vz, ocplx *th, const ocplx *fx, const ocplx *fy,
                   const ocplx *fz, const ocplx *fs);

/* scalar diagnostics (pseudospec3D_phd.f90): variance, product, spectrscc
 * (offset shift + .501 in working precision), sctrans, and the plane means
 * of difucx (dir 1) / difucz (dir 3): out = {<v>, <th>, <v th>} */
double og_variance(const og_grid *g, const ocplx *a, int kin);
double og_product(const og_grid *g, const ocplx *a, const ocplx *b);
void og_spectrscc(const og_grid *g, const ocplx *a, double shift, double *Ek);
void og_sctrans(const og_grid *g, const ocplx *a, const ocplx *b, double *Ek);
void og_difuc(const og_grid *g, const ocplx *a, const ocplx *b, int dir, double *out);

/* Boussinesq diagnostics (pseudospec3D_bouss.f90): havgcomp's horizontal
 * averages gsh[nz] for itype 0..10, and spectpv (nz/2+1 bins) */
void og_havgcomp(const og_grid *g, const ocplx *u, const ocplx *v, const ocplx *w, const ocplx *s, double fo,
                 double bv, int itype, double *gsh);
void og_spectpv(const og_grid *g, const ocplx *u, const ocplx *v, const ocplx *w, const ocplx *s, double *Ek);

/* axisymmetric spectra (pseudospec3D_rot.f90): bins 0 = planes |kz|,
 * 1 = cylinders k_perp; kind 0 = energy, 1 = Re dot, 2 = kk2-weighted dot */
int og_nmaxperp(const og_grid *g);
void og_axispec(const og_grid *g, int bins, int kind, const ocplx *a, const ocplx *b, const ocplx *c,
                const ocplx *d, const ocplx *e, const ocplx *f, double *tot, double *xy, double *zz);
void og_spec2d(const og_grid *g, int kind, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
               const ocplx *e, const ocplx *f, double *E);

/* rotating Boussinesq (include/rotbouss_rkstep{1,2}.f90) */
void og_rotbouss_step(const og_grid *g, int ord, double dt, double nu, double kappa, double xmom, double xtemp,
                      double omegaz, ocplx *vx, ocplx *vy, ocplx *vz, ocplx *th, const ocplx *fx,
                      const ocplx *fy, const ocplx *fz, const ocplx *fs);

/* analysis/trunc3D.fpp: spectral truncation of c (grid g) onto the smaller grid t */
void og_truncate(const og_grid *g, const ocplx *c, const og_grid *t, ocplx *out);

/* The closure network of the learned LES (no counterpart in GHOST):
 * 10 -> 64 -> 64 -> 6, GELU, parameters W1[64][10], b1[64], W2[64][64],
 * b2[64], W3[6][64], b3[6] (5254 floats).  bf16 != 0 rounds every matrix
 * product operand to bfloat16 as the matrix cores do; sums in double. */
void og_mlp_forward(const float *P, long n, const float *X, int bf16, double *Y);
double og_mlp_grad(const float *P, long n, const float *X, const float *T, int bf16, double *grad);

/* transfer and 1D spectra, MHD checks (pseudospec3D_hd.f90:1418-1949,
 * pseudospec3D_mhd.f90:339-712); arrays of nmax/2+1 shells, the reduced
 * totals before the division GHOST applies when writing its files */
void og_entrans(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                const ocplx *e, const ocplx *f, int kin, double *Ek);
void og_heltrans(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                 const ocplx *e, const ocplx *f, double *Hk);
void og_crosspecc(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                  const ocplx *e, const ocplx *f, double shift, double *Ck);
void og_spectr1d(const og_grid *g, const ocplx *a, int dir, double *Ek);
void og_mhdcheck(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *ma,
                 const ocplx *mb, const ocplx *mc, int hel, int crs, int chk, double *out);

#endif
