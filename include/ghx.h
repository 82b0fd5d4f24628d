/*
 * ghx.h -- C ABI of the MI355X (gfx950) implementation of GHOST's
 * pseudospectral hot path.
 *
 * Reference paths below are relative to the GHOST repository root.
 *
 * Conventions (GHOST's, 3D/src/fftp-3/fftp3D.fpp:204-431):
 *   real field      double r[nz][ny][nx]              == Fortran REAL(GP) R(nx,ny,nz)
 *   spectral field  complex double c[nx/2+1][ny][nz]  == Fortran COMPLEX(GP) C(nz,ny,nx/2+1)
 *                   (16-byte interleaved re,im; kz is the contiguous index)
 * Transforms are unnormalised: forward sign -1, inverse +1, r2c->c2r round
 * trip multiplies by nx*ny*nz.  Nonlinear products carry 1/(nx ny nz)^2.
 *
 * Unless a name ends in _host, every field pointer is a DEVICE pointer
 * (from ghx_malloc, hipMalloc or a torch tensor), and calls are
 * asynchronous on the handle's stream; functions returning host scalars
 * synchronise that stream.
 *
 * Every function returns 0 on success or a negative GHX_ERR_* code;
 * ghx_last_error() describes the most recent failure on the calling thread.
 * Handles are not re-entrant: one host thread per handle at a time.
 */
#ifndef GHX_H
#define GHX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GHX_OK 0
#define GHX_ERR_ARG (-1)      /* bad argument (null handle, bad dir, ...) */
#define GHX_ERR_SIZE (-2)     /* grid size without a compiled FFT schedule */
#define GHX_ERR_HIP (-3)      /* HIP runtime error */
#define GHX_ERR_NOMEM (-4)    /* device allocation failed */
#define GHX_ERR_NODEVICE (-5) /* no GPU visible */

/* Direction constants of fftp-3/fftp_mod.fpp:27-28 */
#define GHX_FFTW_REAL_TO_COMPLEX (-1)
#define GHX_FFTW_COMPLEX_TO_REAL (1)
/* Planner flags of fftp-3/fftp_mod.fpp:24-26 (accepted, ignored on the GPU) */
#define GHX_FFTW_ESTIMATE 64
#define GHX_FFTW_MEASURE 0

const char *ghx_last_error(void);
/* 1 if the library found a usable GPU, else 0 */
int ghx_device_ok(void);

/* ---------------------------------------------------------------------- */
/* Device memory and streams                                              */
/* ---------------------------------------------------------------------- */
int ghx_malloc(void **ptr, size_t bytes);
int ghx_free(void *ptr);
int ghx_memcpy_h2d(void *dst, const void *src, size_t bytes);
int ghx_memcpy_d2h(void *dst, const void *src, size_t bytes);
int ghx_memcpy_d2d(void *dst, const void *src, size_t bytes);
int ghx_memset_zero(void *ptr, size_t bytes);
/* hipDeviceSynchronize */
int ghx_sync(void);

/* ---------------------------------------------------------------------- */
/* fftp: GHOST's parallel FFT module on one rank                          */
/* ---------------------------------------------------------------------- */
typedef struct ghx_plan ghx_plan;

/* fftp3d_create_plan(plan,n,fftdir,flags)  fftp-3/fftp3D.fpp:56-124 */
int ghx_plan_create(ghx_plan **plan, int nx, int ny, int nz, int fftdir, int flags);
/* fftp3d_destroy_plan(plan)  fftp-3/fftp3D.fpp:126-155 */
int ghx_plan_destroy(ghx_plan *plan);
/* stream for this plan's launches (hipStream_t cast to void*, NULL = default) */
int ghx_plan_set_stream(ghx_plan *plan, void *stream);
/* fftp3d_real_to_complex(plan,in,out,comm)  fftp-3/fftp3D.fpp:204-315
 * device in[nz][ny][nx] -> device out[nx/2+1][ny][nz] */
int ghx_r2c(ghx_plan *plan, const double *in, void *out);
/* fftp3d_complex_to_real(plan,in,out,comm)  fftp-3/fftp3D.fpp:318-431
 * device in -> device out.  Unlike GHOST, the input is preserved. */
int ghx_c2r(ghx_plan *plan, const void *in, double *out);
/* Batched forms: nf fields (nf <= 12) in one launch sequence. */
int ghx_r2c_many(ghx_plan *plan, int nf, const double *const *in, void *const *out);
int ghx_c2r_many(ghx_plan *plan, int nf, const void *const *in, double *const *out);
/* Host-array forms of fftp3d_real_to_complex / fftp3d_complex_to_real, for a
 * caller whose fields live on the host (copies H->D, transforms, D->H). */
int ghx_r2c_host(ghx_plan *plan, const double *in, void *out);
int ghx_c2r_host(ghx_plan *plan, const void *in, double *out);
/* fftp module timers ffttime/tratime/comtime/tottime (fftp-3/fftp_mod.fpp:29-33),
 * accumulated seconds of GPU time of this plan's transforms.  On one GPU the
 * transposes are fused into the FFT passes and there is no communication, so
 * tra and com stay 0. */
int ghx_plan_timers(ghx_plan *plan, double *ffttime, double *tratime, double *comtime, double *tottime);
int ghx_plan_reset_timers(ghx_plan *plan);

/* ---------------------------------------------------------------------- */
/* Grid: GHOST's grid/kes/ali/boxsize/fft module state                    */
/* main3D.fpp:737-749 (anis), 881-909 (box), 1727-1800 (kx,ky,kz,kn2,kk2,  */
/* kmax, tiny, nmax); pseudospec3D_mod.fpp:13-120                          */
/* ---------------------------------------------------------------------- */
typedef struct ghx_grid ghx_grid;

/* arbsize=0: 2*pi-periodic box (L ignored); arbsize=1: box of Lx,Ly,Lz
 * (units of 2*pi) as with GHOST's ARBSIZE build; Dkk<1e-5 selects the default
 * shell width min(1/Lx,1/Ly,1/Lz). */
int ghx_grid_create(ghx_grid **grid, int nx, int ny, int nz, double Lx, double Ly, double Lz, int arbsize,
                    double Dkk);
int ghx_grid_destroy(ghx_grid *grid);
int ghx_grid_set_stream(ghx_grid *grid, void *stream);
/* kmax, tiny (dealias bounds on kn2), nmax (spectra have nmax/2+1 bins), Dkk, anis */
int ghx_grid_info(const ghx_grid *grid, double *kmax, double *tiny, int *nmax, double *Dkk, int *anis);
/* fftp plans of the grid (planrc / plancr of pseudospec3D_mod.fpp:45-50) */
int ghx_grid_r2c(ghx_grid *grid, const double *in, void *out);
int ghx_grid_c2r(ghx_grid *grid, const void *in, double *out);

/* ---------------------------------------------------------------------- */
/* pseudospec operators (pseudospec3D_hd.f90 / _mhd.f90 / _filt.f90)      */
/* Same argument order and meaning as the Fortran routines.               */
/* ---------------------------------------------------------------------- */
int ghx_derivk3(ghx_grid *g, const void *a, void *b, int dir);                                /* hd :22   */
int ghx_laplak3(ghx_grid *g, const void *a, void *b);                                         /* hd :91   */
int ghx_rotor3(ghx_grid *g, const void *a, const void *b, void *c, int dir);                  /* hd :124  */
int ghx_prodre3(ghx_grid *g, const void *a, const void *b, const void *c, void *d, void *e, void *f); /* hd :316 */
int ghx_nonlhd3(ghx_grid *g, const void *a, const void *b, const void *c, void *out, int dir); /* hd :399 */
int ghx_vector3(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                const void *f, void *x, void *y, void *z);                                    /* mhd :22  */
int ghx_nonlin3(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                const void *f, void *out, int dir);                                           /* mhd :108 */
int ghx_gauge3(ghx_grid *g, const void *a, const void *b, const void *c, void *out, int dir);  /* mhd :230 */
int ghx_energy(ghx_grid *g, const void *a, const void *b, const void *c, int kin, double *out); /* hd :512 */
int ghx_helicity(ghx_grid *g, const void *a, const void *b, const void *c, double *out);       /* hd :665  */
int ghx_cross(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
              const void *f, int kin, double *out);                                           /* hd :787  */
int ghx_maxabs(ghx_grid *g, const void *a, const void *b, const void *c, int kin, double *out); /* hd :910 */
/* divergence term of hdcheck (hd :1036-1071): <|div v|^2> */
int ghx_divergence(ghx_grid *g, const void *a, const void *b, const void *c, double *out);
/* spectrumc (hd :1190-1415): Ek, Hk host arrays of nmax/2+1 doubles (either may be NULL) */
int ghx_spectrumc(ghx_grid *g, const void *a, const void *b, const void *c, int kin, int hel, double *Ek,
                  double *Hk);
/* No counterpart in GHOST: for callers that cannot allocate device memory
 * (code under stream capture, fixed-memory loops).  The energy spectrum of
 * (a, b, c) as ghx_spectrumc with kin = 1, hel = 0 (Ek: nmax/2+1 host
 * doubles), computed without any device allocation: its shell partials live
 * in a fixed 4 MiB device array per GPU (grids up to about 1020^3; larger ones
 * are refused with GHX_ERR_SIZE).  Calls that overlap on one GPU race. */
int ghx_spectrumc_noalloc(ghx_grid *g, const void *a, const void *b, const void *c, double *Ek);
/* Transfer and 1D spectra (reduced shell totals; GHOST's files hold them
 * divided by Dkk, or 0.5/Dk_dir for spectr1d).  Arrays of nmax/2+1 host doubles. */
/* entrans (hd :1629-1820): Re(a d* + b e* + c f*)/N^6 per shell, times kk2 when kin = 0 */
int ghx_entrans(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e, const void *f,
                int kin, double *Ek);
/* heltrans (hd :1823-1949): curl(a,b,c) against (d,e,f) */
int ghx_heltrans(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                 const void *f, double *Hk);
/* crosspecc (mhd :603-712): Re(v . conj(curl A)) per shell, offset shift + .501 */
int ghx_crosspecc(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                  const void *f, double shift, double *Ck);
/* spectr1d (hd :1418-1626): 1D spectrum of one component along k_dir, bins 1..n_dir/2+1 */
int ghx_spectr1d(ghx_grid *g, const void *a, int dir, double *Ek);
/* Scalar diagnostics (pseudo/pseudospec3D_phd.f90), values only.
 * variance (:116-217): <th^2> (kin 1) or <kk2 |th|^2> (kin 0); product
 * (:219-284): <a b>; pscheck (:286-333): out[3] = {<th^2>, <|grad th|^2>,
 * <th s>}; spectrscc (:516-610): |th|^2 per shell, offset shift + .501 in
 * working precision; sctrans (:612-727): Re(a b*) per shell; difucx/difucz
 * (:729-940): out[3 n] = plane means {<v>, <th>, <v th>} along x (n = nx) or
 * z (n = nz).  Shell arrays: nmax/2+1 host doubles. */
int ghx_variance(ghx_grid *g, const void *a, int kin, double *out);
int ghx_product(ghx_grid *g, const void *a, const void *b, double *out);
int ghx_pscheck(ghx_grid *g, const void *a, const void *b, double *out);
int ghx_spectrscc(ghx_grid *g, const void *a, double shift, double *Ek);
int ghx_sctrans(ghx_grid *g, const void *a, const void *b, double *Ek);
int ghx_difucx(ghx_grid *g, const void *a, const void *b, double *out);
int ghx_difucz(ghx_grid *g, const void *a, const void *b, double *out);

/* Boussinesq diagnostics (pseudo/pseudospec3D_bouss.f90), values only.
 * havgcomp (:19-555): horizontal averages gsh[nz] of quantity itype 0..10
 * (fo: Coriolis parameter, bv: Brunt-Vaisala frequency); itype 9 returns the
 * shear of itype 0 as GHOST's branch does, itype 2 the documented <u_z
 * dth/dz>.  tbouss (:612-709): avg, max, min over z of itype 0..10, 11 each.
 * spectpv (:712-812): |omega . grad th|^2 over |kz| planes, nz/2+1 doubles. */
int ghx_havgcomp(ghx_grid *g, const void *u, const void *v, const void *w, const void *s, double fo, double bv,
                 int itype, double *gsh);
int ghx_tbouss(ghx_grid *g, const void *u, const void *v, const void *w, const void *s, double fo, double bv,
               double *avg, double *mx, double *mn);
int ghx_spectpv(ghx_grid *g, const void *u, const void *v, const void *w, const void *s, double *Ek);

/* Axisymmetric spectra about z (pseudo/pseudospec3D_rot.f90), values only.
 * Plane bins int(|kz| Lz + 1), nbpara = nz/2+1; cylinder bins
 * int(sqrt(kx^2+ky^2)/Dkk + 1), nbperp = nmaxperp/2+1 (main3D.fpp:1729).
 * specpara (:25-335), specperp (:338-686): Ek (kin 1: of (a,b,c), 0: of its
 * curl, 2: skipped), Hk (hel 1: Re(a . conj(curl a))), each 3 nb doubles
 * {total, xy part, z part}, the parts over the kz = 0 plane only for
 * cylinders; either array may be NULL.  entpara (:689-875), entperp
 * (:878-1064): Re(a d* + b e* + c f*) per bin (times kk2 for kin 0);
 * heltpara (:1067-1196), heltperp (:1199-1328): curl(a,b,c) against (d,e,f);
 * spec2D (:1331-1629): nbperp * nbpara doubles, kperp fastest. */
int ghx_axis_bins(const ghx_grid *g, int *nbpara, int *nbperp);
int ghx_specpara(ghx_grid *g, const void *a, const void *b, const void *c, int kin, int hel, double *Ek,
                 double *Hk);
int ghx_specperp(ghx_grid *g, const void *a, const void *b, const void *c, int kin, int hel, double *Ek,
                 double *Hk);
int ghx_entpara(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                const void *f, int kin, double *Ek);
int ghx_entperp(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                const void *f, int kin, double *Ek);
int ghx_heltpara(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                 const void *f, double *Hk);
int ghx_heltperp(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                 const void *f, double *Hk);
int ghx_spec2d(ghx_grid *g, const void *a, const void *b, const void *c, int kin, int hel, double *Ek, double *Hk);
/* hdcheck (hd :984-1105) without file output: out[5] = {eng, ens, pot, khe, div}
 * (<v^2>, <omega^2>, <v.f>, kinetic helicity if hel = 1, <(div v)^2> if chk = 1) */
int ghx_hdcheck(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                const void *f, int hel, int chk, double *out);
/* mhdcheck (mhd :339-547) without file output: out[11] = {eng, ens, cur, engk,
 * engm, helk, helm, crh, asq, divk, divm}; crs = 2 (Hall) is refused */
int ghx_mhdcheck(ghx_grid *g, const void *a, const void *b, const void *c, const void *ma, const void *mb,
                 const void *mc, int hel, int crs, int chk, double *out);
int ghx_phaseshift(ghx_grid *g, void *a, double cdump_re, double cdump_im);                   /* hd :2084 */
int ghx_normalize(ghx_grid *g, void *fx, void *fy, void *fz, double f0, int kin);             /* module_dns.f90:13 */
int ghx_sharp(ghx_grid *g, const void *a, void *b, double delta);                             /* filt :200 */
/* trunc3D (analysis/trunc3D.fpp:190-320) spectral core: out (grid t, no
 * larger than g, even sizes) = the modes of c (spectrum on g) with
 * kx^2/nxt^2 + ky^2/nyt^2 + kz^2/nzt^2 <= 1/9, scaled by 1/(nx ny nz); the c2r
 * of out on t is the truncated field sampled on t's points. */
int ghx_truncate(ghx_grid *g, const void *c, ghx_grid *t, void *out);
/* advect3 (pseudo/pseudospec3D_phd.f90:23-115): e = FFT[-(A . grad d)] / N^6
 * with A = (a, b, c); the products of GHOST's three c2r pairs are fused into
 * one inverse chain, the x pass and one forward chain. */
int ghx_advect3(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, void *e);

/* ---------------------------------------------------------------------- */
/* Initial conditions / forcing (examples/initialv.f90_tg,                 */
/* examples/initialfv.f90_random; randu pseudospec3D_mod.fpp:130-148)      */
/* ---------------------------------------------------------------------- */
double ghx_randu(int32_t *seed);
int ghx_initialv_tg(ghx_grid *g, double kdn, double kup, double u0, void *vx, void *vy, void *vz);
int ghx_initialfv_random(ghx_grid *g, double kdn, double kup, double f0, int32_t *seed, void *fx, void *fy,
                         void *fz);
/* random-phase solenoidal field with shell spectrum ~ k^4 exp(-2 (k/k0)^2),
 * drawn in the fill order of initialfv.f90_random, normalised to <u^2> = u0^2 */
int ghx_initialv_k4exp(ghx_grid *g, double k0, double u0, int32_t *seed, void *vx, void *vy, void *vz);

/* ---------------------------------------------------------------------- */
/* Runge-Kutta substeps with fields resident in HBM                        */
/* ---------------------------------------------------------------------- */
/* One pass of include/hd_rkstep2.f90 for order index o (main3D.fpp:3229):
 *   v = C + dt*(nu*lap v + P[-(curl v) x v] + f)/o   where tiny <= kn2 <= kmax, else 0,
 * computed by the fused path (curl in the first inverse pass, cross product
 * between the x passes, projection and update in the last forward pass).
 * v_in/C/f: 3 fields each (f may be NULL); v_out may equal v_in or C. */
int ghx_hd_rkstep2(ghx_grid *g, int o, double dt, double nu, const void *const *v_in, const void *const *C,
                   const void *const *f, void *const *v_out);

/* include/roth_rkstep2.f90:1-49: the same pass in a frame rotating about z
 * at rate omegaz; the Coriolis force enters the products before the
 * projection (x: - 2 omegaz v_y, y: + 2 omegaz v_x). */
int ghx_roth_rkstep2(ghx_grid *g, int o, double dt, double nu, double omegaz, const void *const *v_in,
                     const void *const *C, const void *const *f, void *const *v_out);

typedef struct ghx_hd ghx_hd;
/* HD solver state: v (3), forcing f (3), ping-pong and pass scratch, all on device. */
int ghx_hd_create(ghx_hd **s, ghx_grid *g, int ord, double dt, double nu);
int ghx_hd_destroy(ghx_hd *s);
/* device pointers of the resident fields: v[3], f[3] */
int ghx_hd_fields(ghx_hd *s, void **v, void **f);
/* call after writing v through those pointers.  The next step checks whether
 * v is zero outside the dealias set (tiny <= kn2 <= kmax); while it is, the
 * passes skip every line that carries no kept mode (exact zeros in, discarded
 * values out), otherwise the first step runs the full transforms. */
int ghx_hd_touch(ghx_hd *s);
/* call after writing f: records which (kx,ky) lines carry forcing, or 0 for none */
int ghx_hd_set_forcing(ghx_hd *s, int enabled);
/* nsteps full RK steps (main3D.fpp:3202-3231 without output/diagnostics) */
int ghx_hd_step(ghx_hd *s, int nsteps);
/* the ROTH solver (include/roth_rkstep{1,2}.f90): rotation rate omegaz about z */
int ghx_hd_set_rotation(ghx_hd *s, double omegaz);

/* include/mhd_rkstep2.f90:4-65 for order index o on caller fields (the fused
 * substep of the MHD solver): vin, C, vout hold (v_x, v_y, v_z, a_x, a_y,
 * a_z); f (mechanical) and m (electromotive) forcing, 3 fields each, both
 * given or both NULL; vout may equal vin or C. */
int ghx_mhd_rkstep2(ghx_grid *g, int o, double dt, double nu, double mu, const void *const *vin, const void *const *C,
                    const void *const *f, const void *const *m, void *const *vout);
/* include/bouss_rkstep2.f90:1-71 (the ord = 3 correction of :73-145 is
 * ghx_bouss_ord3) and include/rotbouss_rkstep2.f90 (Coriolis force of
 * rotation omegaz about z) for order index o on caller fields: vin, C, vout
 * hold (v_x, v_y, v_z, th); f holds (f_x, f_y, f_z, f_s) or is NULL; vout may
 * equal vin or C. */
int ghx_bouss_rkstep2(ghx_grid *g, int o, double dt, double nu, double kappa, double xmom, double xtemp,
                      const void *const *vin, const void *const *C, const void *const *f, void *const *vout);
int ghx_rotbouss_rkstep2(ghx_grid *g, int o, double dt, double nu, double kappa, double xmom, double xtemp,
                         double omegaz, const void *const *vin, const void *const *C, const void *const *f,
                         void *const *vout);
/* include/bouss_rkstep2.f90:73-145, the ord = 3 correction that GHOST applies
 * after ghx_bouss_rkstep2 with o = 1: v (v_x, v_y, v_z) += 2 N(L(u) + N(u))
 * dt^3/24, from C (v_x, v_y, v_z, th at the step start), with N the projected
 * nonlinear term and buoyancy, L the Laplacian.  C is left unchanged (GHOST
 * overwrites C1-C3); v may not alias C. */
int ghx_bouss_ord3(ghx_grid *g, double dt, double xmom, const void *const *C, void *const *v);

typedef struct ghx_mhd ghx_mhd;
/* MHD solver state: v (3), a (3, vector potential), f (3), m (3). */
int ghx_mhd_create(ghx_mhd **s, ghx_grid *g, int ord, double dt, double nu, double mu);
int ghx_mhd_destroy(ghx_mhd *s);
int ghx_mhd_fields(ghx_mhd *s, void **v, void **a, void **f, void **m);
/* as ghx_hd_touch, for v and a */
int ghx_mhd_touch(ghx_mhd *s);
int ghx_mhd_set_forcing(ghx_mhd *s, int enabled);
int ghx_mhd_step(ghx_mhd *s, int nsteps);

typedef struct ghx_bouss ghx_bouss;
/* Boussinesq solver (include/bouss_rkstep1.f90, bouss_rkstep2.f90:1-71 in the
 * loop of main3D.fpp:3202-3231): state v (3) and th, forcing f (3) and fs.
 * kappa: diffusivity; xmom, xtemp: the buoyancy and heat-current factors
 * already multiplied by bvfreq (main3D.fpp:973-974).  ord 1 to 3; ord = 3
 * without rotation adds the correction of bouss_rkstep2.f90:73-145 after the
 * last substep (the state then keeps modes outside the dealias set, as
 * GHOST's does, and the next step runs unpruned).  With xmom = xtemp = 0
 * th is a passive scalar: the PHD solver (include/phd_rkstep{1,2}.f90). */
int ghx_bouss_create(ghx_bouss **s, ghx_grid *g, int ord, double dt, double nu, double kappa, double xmom,
                     double xtemp);
int ghx_bouss_destroy(ghx_bouss *s);
/* device pointers: v[3], th, f[3], fs */
int ghx_bouss_fields(ghx_bouss *s, void **v, void **th, void **f, void **fs);
/* as ghx_hd_touch; the mean of th (kn2 < tiny) is carried and does not count */
int ghx_bouss_touch(ghx_bouss *s);
int ghx_bouss_set_forcing(ghx_bouss *s, int enabled);
int ghx_bouss_step(ghx_bouss *s, int nsteps);
/* the ROTBOUSS solver (include/rotbouss_rkstep{1,2}.f90): Coriolis force of
 * rotation omegaz about z, and RK of any order without a correction */
int ghx_bouss_set_rotation(ghx_bouss *s, double omegaz);

/* ---------------------------------------------------------------------- */
/* Ensembles: many independent unforced HD runs (cubic 2 pi boxes of any     */
/* supported sizes, each with its own viscosity) advanced together.  Members */
/* of one size share every pass launch of the fused substep (include/        */
/* hd_rkstep2.f90 in the loop of main3D.fpp:3202-3231, once per member);     */
/* after the first step the ensemble step is a captured HIP graph.  No        */
/* counterpart in GHOST, which runs one solver per process.                  */
/* ---------------------------------------------------------------------- */
typedef struct ghx_ensemble ghx_ensemble;
/* n[m]: grid size of member m (n^3), nu[m]: its viscosity; RK order ord, step dt */
int ghx_ensemble_create(ghx_ensemble **e, int nmembers, const int *n, const double *nu, double dt, int ord);
int ghx_ensemble_destroy(ghx_ensemble *e);
/* device pointers of member m's velocity v[3] (spectral, on its grid) */
int ghx_ensemble_fields(ghx_ensemble *e, int member, void **v);
/* the grid of member m (for initial conditions and diagnostics) */
int ghx_ensemble_grid(ghx_ensemble *e, int member, ghx_grid **g);
/* call after writing member states through ghx_ensemble_fields */
int ghx_ensemble_touch(ghx_ensemble *e);
int ghx_ensemble_step(ghx_ensemble *e, int nsteps);
/* 0: launch every pass eagerly each step; 1 (default): replay a captured graph */
int ghx_ensemble_set_graph(ghx_ensemble *e, int enabled);
/* choose the radix schedules and the z and y tile widths of each size group
 * by timing `steps` steps per candidate (member states are restored): the z
 * and y passes each on the default schedule or on four points per lane where
 * that differs (sizes 32..256), times twelve tile-width pairs; tz/ty receive
 * one value per group, ghx_ensemble_schedules all three */
int ghx_ensemble_autotune(ghx_ensemble *e, int steps, int *tz, int *ty);
/* per group: schedule bits (1: z passes, 2: y passes on four points per
 * lane), z and y tile widths (0: built-in); any pointer may be NULL */
int ghx_ensemble_schedules(ghx_ensemble *e, int *sched, int *tz, int *ty);
/* set one group's schedule bits and tile widths (a choice the autotuner made
 * earlier, or a test's) */
int ghx_ensemble_set_schedule(ghx_ensemble *e, int group, int sched, int tz, int ty);
/* size groups: ngroups, and per group its size and member count */
int ghx_ensemble_groups(ghx_ensemble *e, int *ngroups, int *sizes, int *counts);
int ghx_ensemble_sync(ghx_ensemble *e);

/* ---------------------------------------------------------------------- */
/* Lagrangian tracers: GHOST's GPart with cubic-spline interpolation on one */
/* rank (lagpart/gparts_mod.f90, lagpart/gpsplineint_mod.fpp).  Positions   */
/* are in grid units, [0, n) along each axis, as GPart's px_/py_/pz_.        */
/* ---------------------------------------------------------------------- */
typedef struct ghx_tracers ghx_tracers;
/* GPart_ctor (gparts_mod.f90:148-327) for n tracers on grid g (at least 4
 * points per axis).  zcompat = 1 reproduces GHOST's z stencil: the spline is
 * set up for 3 ghost planes (gparts_mod.f90:291) while the extended field
 * holds 2 (gpartcomm_mod.f90:1005-1057), so the field is sampled one grid
 * cell above the particle and the top cell is clamped
 * (gpsplineint_mod.fpp:653-661).  zcompat = 0 interpolates at the particle.
 * Positions start at 0; write them through ghx_tracers_positions. */
int ghx_tracers_create(ghx_tracers **t, ghx_grid *g, long long n, int zcompat);
int ghx_tracers_destroy(ghx_tracers *t);
/* device pointers of the positions (GPart px_, py_, pz_; n doubles each) */
int ghx_tracers_positions(ghx_tracers *t, double **px, double **py, double **pz);
/* device pointers of the velocities interpolated by the last stage (lvx_, lvy_, lvz_) */
int ghx_tracers_velocities(ghx_tracers *t, double **ux, double **uy, double **uz);
/* GPart_SetStepRKK (gparts_mod.f90:1839-1908): x0 = x at the start of an RK step */
int ghx_tracers_set_step(ghx_tracers *t);
/* One RK stage, main3D.fpp:3233-3267 + GPart_StepRKK (gparts_mod.f90:2046-2145)
 * + GPart_EndStageRKK (:1911-2043): u = spline interpolant of c2r(v/(nx ny nz))
 * at x, then x = x0 + dt xk n/(2 pi L) u, periodised.  v: 3 spectral fields. */
int ghx_tracers_step(ghx_tracers *t, const void *const *v, double dt, double xk);
/* GPart_EulerToLag (gparts_mod.f90:2216-2269): out[q] = spline interpolant of
 * the real field r[z][y][x] at tracer q (n doubles, device) */
int ghx_tracers_euler_to_lag(ghx_tracers *t, const double *field, double *out);
/* Advance the tracers inside ghx_hd_step / ghx_mhd_step: x0 = x at the start
 * of each step and one tracer stage after every substep, on the velocity that
 * substep produced (main3D.fpp:3202-3267).  t = NULL detaches. */
int ghx_hd_set_tracers(ghx_hd *s, ghx_tracers *t);
int ghx_mhd_set_tracers(ghx_mhd *s, ghx_tracers *t);
int ghx_bouss_set_tracers(ghx_bouss *s, ghx_tracers *t);

/* ---------------------------------------------------------------------- */
/* Binary field files (mpiio/binary_io.f90, module iovar/iompi in           */
/* mpiio/binary_mod.f90:11-32): the global array in Fortran order, raw native     */
/* FP64 (complex: re, im pairs), no header, at byte offset disp; a slab     */
/* (n1, n2, ksta:kend) sits at disp + (ksta-1) n1 n2 esize.  The file is    */
/* dir/fname.nmb.out when bmangle = 1 (the default); otherwise dir/fname,   */
/* except io_write, which opens fname alone.  var may be device memory      */
/* (staged through pinned buffers; all streams are synchronised first) or   */
/* host memory.  Reading a file shorter than the slab fails.                */
/* ---------------------------------------------------------------------- */
typedef struct ghx_ioplan ghx_ioplan;
/* io_init (mpiio/binary_io.f90:13-59): real arrays var(n1, n2, ksta:kend) */
int ghx_io_init(ghx_ioplan **plan, int myrank, const int n[3], int ksta, int kend);
/* io_initc (mpiio/binary_io.f90:189-233): complex arrays */
int ghx_io_initc(ghx_ioplan **plan, int myrank, const int n[3], int ksta, int kend);
int ghx_io_destroy(ghx_ioplan *plan);
/* iovar's iswap (byte-reverse every real after io_read, gutils_mod.fpp:114),
 * iompi's bmangle and disp (mpiio/binary_mod.f90:11-30) */
int ghx_io_set_options(ghx_ioplan *plan, int iswap, int bmangle, long long disp);
int ghx_io_read(ghx_ioplan *plan, const char *dir, const char *fname, const char *nmb, double *var);       /* :62-129 */
int ghx_io_write(ghx_ioplan *plan, const char *dir, const char *fname, const char *nmb, const double *var); /* :132-186 */
int ghx_io_readc(ghx_ioplan *plan, const char *dir, const char *fname, const char *nmb, void *var);         /* :236-282 */
int ghx_io_writec(ghx_ioplan *plan, const char *dir, const char *fname, const char *nmb, const void *var);  /* :285-331 */
/* Field output of main3D.fpp:2675-2710: vx, vy, vz = c2r(v/(nx ny nz)) and,
 * with outs >= 1, wx, wy, wz = c2r(rotor3(v/(nx ny nz))) written first;
 * ext is the time label (fmtext, i4.4). */
int ghx_hd_output(ghx_grid *g, ghx_ioplan *io, const char *odir, const char *ext, int outs, const void *vx,
                  const void *vy, const void *vz);
/* Restart of main3D.fpp:2143-2169: vx, vy, vz read and transformed with r2c */
int ghx_hd_restart(ghx_grid *g, ghx_ioplan *io, const char *idir, const char *ext, void *vx, void *vy, void *vz);

/* ---------------------------------------------------------------------- */
/* Learned subgrid closure (no counterpart in GHOST; GHOST's own closures  */
/* are the alpha/LES/EDQNM models of pseudospec3D_{lles,lhd,lahd,edqnm}.f90,*/
/* which this build does not provide).  A 10 -> 64 -> 64 -> 6 perceptron    */
/* with GELU maps the resolved velocity-gradient tensor (d u_i/d x_j, i     */
/* major) and the filter width to the stress components tau_xx, tau_yy,     */
/* tau_zz, tau_xy, tau_xz, tau_yz.  FP32 master parameters (5254 floats:     */
/* W1[64][10], b1[64], W2[64][64], b2[64], W3[6][64], b3[6]) enter the     */
/* matrix cores as BF16 with FP32 accumulation.  Inputs and outputs are     */
/* normalised: norm = mean_x[10], sd_x[10], mean_y[6], sd_y[6].  Arrays X   */
/* (n x 10) and T, Y (n x 6) are row-major FP32 device arrays.             */
/* ---------------------------------------------------------------------- */
typedef struct ghx_mlp ghx_mlp;
int ghx_mlp_nparams(void);
/* Glorot-uniform weights from seed, zero biases, identity normalisation */
int ghx_mlp_create(ghx_mlp **m, unsigned long long seed);
int ghx_mlp_destroy(ghx_mlp *m);
int ghx_mlp_set_stream(ghx_mlp *m, void *stream);
/* device pointer to the FP32 master parameters (read or overwrite in place) */
int ghx_mlp_params(ghx_mlp *m, float **params);
int ghx_mlp_set_norm(ghx_mlp *m, const float *norm);  /* host array of 32 */
int ghx_mlp_get_norm(ghx_mlp *m, float *norm);
/* normalisation from the moments of n points of the fields (the filter-width
 * input is centred on delta with unit scale) */
int ghx_mlp_fit_norm(ghx_mlp *m, long long n, const double *const *grad9, double delta, const double *const *tau6);
/* X (and T when tau6 and T are given) from fields, normalised */
int ghx_mlp_pack(ghx_mlp *m, long long n, const double *const *grad9, double delta, const double *const *tau6,
                 float *X, float *T);
/* Y = network(X), normalised units */
int ghx_mlp_forward(ghx_mlp *m, long long n, const float *X, float *Y);
/* loss = mean over the n x 6 outputs of (Y - T)^2 and its gradient (grad:
 * device array of ghx_mlp_nparams() floats, may be NULL) */
int ghx_mlp_grad(ghx_mlp *m, long long n, const float *X, const float *T, float *grad, double *loss);
/* steps of Adam (beta 0.9, 0.999, eps 1e-8) on batches of the data set taken
 * in a seeded permutation; losses (host, steps floats, may be NULL) */
int ghx_mlp_train(ghx_mlp *m, long long n, const float *X, const float *T, long long batch, int steps, float lr,
                  unsigned long long seed, float *losses);
/* stress fields from gradient fields (n points, FP64 device fields) */
int ghx_mlp_stress(ghx_mlp *m, long long n, const double *const *grad9, double delta, double *const *tau6);
/* physical velocity gradients d u_i / d x_j (i major, nine real fields) of a
 * spectral state in GHOST's normalisation */
int ghx_velocity_gradients(ghx_grid *g, const void *const *v, double *const *grad9);
/* training pairs from a fine-grid state v: ub = sharp(v, kc) (filt.f90:200
 * with delta = pi/kc); vsmall = the LES state on the small grid t (trunc3D's
 * truncation of ub, in t's normalisation) and tau6 = the physical subgrid
 * stress F(u_i u_j) - F(ub_i ub_j) on t's points (F = sharp then truncate),
 * order xx, yy, zz, xy, xz, yz */
int ghx_sgs_targets(ghx_grid *g, ghx_grid *t, const void *const *v, double kc, double *const *tau6,
                    void *const *vsmall);
/* Attach a closure to an HD solver (m = NULL detaches): before each substep
 * the forcing becomes f + P(-div tau), tau = m(grad u, delta), P the
 * projection of nonlhd3 (pseudospec3D_hd.f90:399-509). */
int ghx_hd_set_sgs(ghx_hd *s, ghx_mlp *m, double delta);
/* Spectral part of the closure (no counterpart in GHOST): -nuk[q] kk2 v on
 * the shells q + 1 = int(sqrt(kk2)/Dkk + .501) = 1..nb is added to the force
 * of every substep, with or without a network; nb = 0 removes it. */
int ghx_hd_set_sgs_spectral(ghx_hd *s, int nb, const double *nuk);

/* ---------------------------------------------------------------------- */
/* Pass profiler (replaces the GTStart/GTStop handles around the fftp      */
/* phases, utils/gtimer_mod.f90:37-185 and fftp-3/fftp3D.fpp:244-312):      */
/* HIP events recorded around every pass launch on its stream.  Slots name  */
/* the pass kernels; total_bytes sums each launch's algorithmic traffic.    */
/* ---------------------------------------------------------------------- */
int ghx_prof_enable(int on);
int ghx_prof_reset(void);
int ghx_prof_read(int slot, const char **name, int *count, double *total_ms, double *total_bytes);

/* Bandwidth yardstick (no counterpart in GHOST): plain kernels over `bytes`
 * (a multiple of 16) on `blocks` blocks of 256 lanes, asynchronous on the
 * stream.  kind 0: copy src -> dst; 1: read src (dst: an 8-byte device sink that
 * the kernel writes only for a sum no input produces); 2: write dst;
 * 3: copy with streaming stores; 4: write with streaming stores.  Kinds 5-9
 * are kinds 0, 3, 1, 2, 4 with each block moving one contiguous chunk of
 * bytes / blocks instead of a grid-stride loop. */
int ghx_stream_probe(int kind, void *dst, const void *src, size_t bytes, int blocks, void *stream);

/* Launch shape of an FFT pass (no counterpart in GHOST; host-only, no GPU
 * needed): kind 0 z/y tile passes, 1 Boussinesq z forward, 2 wide y inverse,
 * 3 two-for-one x passes (n = nx), 4 half-length x passes (n = nx/2); for a
 * line length n, nlines lines along the tiled axis, a requested tile width
 * and an LDS limit in bytes.  Returns 0 when the shape passes the check every
 * launcher applies before it launches (lanes within the kernel's launch
 * bounds, whole lines, tile widths a power of two up to 32, LDS within the
 * limit), -2 when it would be refused, -1 for a size or kind without a
 * kernel. */
int ghx_pass_shape(int kind, int n, int nlines, int want, long long ldsmax, int *lines, int *block,
                   long long *lds, int *lanes);

#ifdef __cplusplus
}
#endif
#endif /* GHX_H */
