/*
 * ghost_oracle.c -- CPU restatement of GHOST's pseudospectral hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
 * baseline.  Never linked into, or called by, the product library.
 *
 * Every routine cites the GHOST source it restates (paths relative to the
 * reference repository root, 3D/src/...).  Arithmetic is FP64 throughout and
 * the file is compiled with -ffp-contract=off so that each expression rounds
 * the way it is written, in GHOST's operand order.
 *
 * The FFT is this file's own mixed-radix Stockham transform (radix 8/4/3/2
 * plus a generic odd-prime DFT), so the oracle needs no FFT library; its
 * conventions are FFTW's (sign -1 forward, +1 inverse, unnormalised; the
 * 1D c2r ignores the imaginary part of the DC and Nyquist inputs).
 */
#include "ghost_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OMP_FOR _Pragma("omp parallel for schedule(static)")

static int g_threads = 0;
void og_set_threads(int n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}
int og_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* 1D mixed-radix Stockham FFT                                              */
/* ------------------------------------------------------------------------ */

typedef struct fft1 {
  int n, nf;
  int fac[64];
  ocplx *tw; /* tw[k] = exp(-2 pi i k / n), k = 0..n-1 */
} fft1;

static fft1 *fft1_new(int n) {
  fft1 *p = (fft1 *)calloc(1, sizeof(fft1));
  p->n = n;
  int m = n;
  while (m % 8 == 0) { p->fac[p->nf++] = 8; m /= 8; }
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int q = 3; m > 1; q += 2)
    while (m % q == 0) { p->fac[p->nf++] = q; m /= q; }
  p->tw = (ocplx *)malloc(sizeof(ocplx) * (size_t)n);
  const long double twopi = 6.283185307179586476925286766559L;
  for (int k = 0; k < n; ++k) {
    long double a = twopi * (long double)k / (long double)n;
    p->tw[k] = (double)cosl(a) - I * (double)sinl(a);
  }
  return p;
}

static void fft1_free(fft1 *p) {
  if (!p) return;
  free(p->tw);
  free(p);
}

static inline ocplx twid(const fft1 *p, int e, int sign) {
  ocplx w = p->tw[e];
  return sign < 0 ? w : conj(w);
}

/* multiply by sign*i */
static inline ocplx mul_si(ocplx z, int sign) {
  return sign < 0 ? cimag(z) - I * creal(z) : -cimag(z) + I * creal(z);
}

static void dft_small(ocplx *v, int R, const fft1 *p, int sign) {
  const double s3 = 0.86602540378443864676; /* sqrt(3)/2 */
  const double r2 = 0.70710678118654752440; /* 1/sqrt(2) */
  switch (R) {
  case 2: {
    ocplx a = v[0] + v[1], b = v[0] - v[1];
    v[0] = a; v[1] = b;
    return;
  }
  case 3: {
    ocplx t = v[1] + v[2], d = v[1] - v[2];
    ocplx m = v[0] - 0.5 * t;
    ocplx e = mul_si(s3 * d, sign);
    v[0] = v[0] + t;
    v[1] = m + e;
    v[2] = m - e;
    return;
  }
  case 4: {
    ocplx a0 = v[0] + v[2], a1 = v[0] - v[2];
    ocplx a2 = v[1] + v[3], a3 = mul_si(v[1] - v[3], sign);
    v[0] = a0 + a2; v[2] = a0 - a2;
    v[1] = a1 + a3; v[3] = a1 - a3;
    return;
  }
  case 8: {
    ocplx e[4] = {v[0], v[2], v[4], v[6]};
    ocplx o[4] = {v[1], v[3], v[5], v[7]};
    dft_small(e, 4, p, sign);
    dft_small(o, 4, p, sign);
    /* W8^q for q=1,2,3 with W8 = exp(sign*2 pi i/8) */
    ocplx w1 = sign < 0 ? r2 - I * r2 : r2 + I * r2;
    ocplx o1 = o[1] * w1;
    ocplx o2 = mul_si(o[2], sign);
    ocplx o3 = mul_si(o[3] * w1, sign);
    v[0] = e[0] + o[0]; v[4] = e[0] - o[0];
    v[1] = e[1] + o1;   v[5] = e[1] - o1;
    v[2] = e[2] + o2;   v[6] = e[2] - o2;
    v[3] = e[3] + o3;   v[7] = e[3] - o3;
    return;
  }
  default: {
    ocplx out[64];
    int step = p->n / R;
    for (int q = 0; q < R; ++q) {
      ocplx acc = 0;
      for (int r = 0; r < R; ++r) acc += v[r] * twid(p, ((q * r) % R) * step, sign);
      out[q] = acc;
    }
    memcpy(v, out, sizeof(ocplx) * (size_t)R);
  }
  }
}

/* In place on x, using work w (length n). */
static void fft1_exec(const fft1 *p, ocplx *x, ocplx *w, int sign) {
  const int n = p->n;
  if (n == 1) return;
  ocplx *src = x, *dst = w;
  int Ns = 1;
  for (int f = 0; f < p->nf; ++f) {
    const int R = p->fac[f];
    const int m = n / R;
    const int tws = n / (Ns * R);
    for (int j = 0; j < m; ++j) {
      const int k = j % Ns;
      ocplx v[64];
      for (int r = 0; r < R; ++r) {
        ocplx t = src[j + r * m];
        if (k) t *= twid(p, (int)(((long)k * r * tws) % n), sign);
        v[r] = t;
      }
      dft_small(v, R, p, sign);
      const int base = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
    }
    ocplx *t = src; src = dst; dst = t;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(ocplx) * (size_t)n);
}

void og_fft1d(int n, ocplx *x, int sign) {
  fft1 *p = fft1_new(n);
  ocplx *w = (ocplx *)malloc(sizeof(ocplx) * (size_t)n);
  fft1_exec(p, x, w, sign);
  free(w);
  fft1_free(p);
}

/* Real row (length nx) -> half spectrum (nx/2+1), sign -1.  Even nx uses the
 * packed half-length transform, odd nx the full complex one. */
static void rfft_row(const fft1 *ph, const fft1 *pf, int nx, const double *x, ocplx *X,
                     ocplx *wa, ocplx *wb) {
  if (nx % 2 == 0) {
    const int M = nx / 2;
    for (int n = 0; n < M; ++n) wa[n] = x[2 * n] + I * x[2 * n + 1];
    fft1_exec(ph, wa, wb, -1);
    for (int k = 0; k <= M; ++k) {
      ocplx A = wa[k % M];
      ocplx B = conj(wa[(M - k) % M]);
      ocplx Xe = 0.5 * (A + B);
      ocplx Xo = -0.5 * I * (A - B);
      X[k] = Xe + pf->tw[k] * Xo;
    }
  } else {
    for (int n = 0; n < nx; ++n) wa[n] = x[n];
    fft1_exec(pf, wa, wb, -1);
    for (int k = 0; k <= nx / 2; ++k) X[k] = wa[k];
  }
}

/* Half spectrum (nx/2+1) -> real row, sign +1, unnormalised. */
static void irfft_row(const fft1 *ph, const fft1 *pf, int nx, const ocplx *Xin, double *x,
                      ocplx *wa, ocplx *wb) {
  if (nx % 2 == 0) {
    const int M = nx / 2;
    for (int k = 0; k < M; ++k) {
      ocplx A = (k == 0) ? creal(Xin[0]) : Xin[k];
      ocplx Bs = (k == 0) ? creal(Xin[M]) : Xin[M - k];
      ocplx B = conj(Bs);
      ocplx Xe = A + B;
      ocplx Xo = (A - B) * conj(pf->tw[k]);
      wa[k] = Xe + I * Xo;
    }
    fft1_exec(ph, wa, wb, +1);
    for (int n = 0; n < M; ++n) {
      x[2 * n] = creal(wa[n]);
      x[2 * n + 1] = cimag(wa[n]);
    }
  } else {
    wa[0] = creal(Xin[0]);
    for (int k = 1; k <= nx / 2; ++k) {
      wa[k] = Xin[k];
      wa[nx - k] = conj(Xin[k]);
    }
    fft1_exec(pf, wa, wb, +1);
    for (int n = 0; n < nx; ++n) x[n] = creal(wa[n]);
  }
}

typedef struct fftset {
  fft1 *px, *pxh, *py, *pz;
} fftset;

/* ------------------------------------------------------------------------ */
/* Grid / wavenumbers: main3D.fpp:737-749 (anis), 881-909 (box), 1727-1800  */
/* ------------------------------------------------------------------------ */

#define IDX(g, i, j, k) (((size_t)(i) * (g)->ny + (size_t)(j)) * (g)->nz + (size_t)(k))
#define RIDX(g, i, j, k) (((size_t)(k) * (g)->ny + (size_t)(j)) * (g)->nx + (size_t)(i))

og_grid *og_grid_new(int nx, int ny, int nz, double Lx, double Ly, double Lz, int arbsize,
                     double Dkk) {
  og_grid *g = (og_grid *)calloc(1, sizeof(og_grid));
  g->nx = nx; g->ny = ny; g->nz = nz; g->nxh = nx / 2 + 1;
  g->Lx = arbsize ? Lx : 1.0;
  g->Ly = arbsize ? Ly : 1.0;
  g->Lz = arbsize ? Lz : 1.0;
  g->Dkx = arbsize ? 1.0 / g->Lx : 1.0;
  g->Dky = arbsize ? 1.0 / g->Ly : 1.0;
  g->Dkz = arbsize ? 1.0 / g->Lz : 1.0;
  g->Dkk = arbsize ? Dkk : 0.0;
  if (g->Dkk < 1e-5) g->Dkk = fmin(g->Dkx, fmin(g->Dky, g->Dkz));
  g->anis = (arbsize || nx != ny || ny != nz) ? 1 : 0;

  g->kmax = 1.0 / 9.0;
  double mx = fmax(nx * g->Dkx, fmax(ny * g->Dky, nz * g->Dkz));
  g->nmax = (int)(mx / g->Dkk);
  if (!arbsize && g->anis == 0) g->kmax = g->kmax * ((double)nx * (double)nx);
  g->tiny = fmin(1e-5, 0.1 / ((double)g->nmax * (double)g->nmax));

  g->kx = (double *)malloc(sizeof(double) * nx);
  g->ky = (double *)malloc(sizeof(double) * ny);
  g->kz = (double *)malloc(sizeof(double) * nz);
  for (int i = 1; i <= nx / 2; ++i) { g->kx[i - 1] = i - 1; g->kx[i + nx / 2 - 1] = i - nx / 2 - 1; }
  if (nx == 1) g->kx[0] = 0;
  for (int j = 1; j <= ny / 2; ++j) { g->ky[j - 1] = j - 1; g->ky[j + ny / 2 - 1] = j - ny / 2 - 1; }
  if (ny == 1) g->ky[0] = 0;
  for (int k = 1; k <= nz / 2; ++k) { g->kz[k - 1] = k - 1; g->kz[k + nz / 2 - 1] = k - nz / 2 - 1; }
  double rmp = 1, rmq = 1, rms = 1;
  if (g->anis) {
    rmp = 1.0 / ((double)nx * (double)nx);
    rmq = 1.0 / ((double)ny * (double)ny);
    rms = 1.0 / ((double)nz * (double)nz);
  }
  size_t nc = (size_t)g->nxh * ny * nz;
  g->kn2 = (double *)malloc(sizeof(double) * nc);
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k)
        g->kn2[IDX(g, i, j, k)] = rmp * (g->kx[i] * g->kx[i]) + rmq * (g->ky[j] * g->ky[j]) +
                                  rms * (g->kz[k] * g->kz[k]);
  if (arbsize) {
    for (int i = 0; i < nx; ++i) g->kx[i] *= g->Dkx;
    for (int j = 0; j < ny; ++j) g->ky[j] *= g->Dky;
    for (int k = 0; k < nz; ++k) g->kz[k] *= g->Dkz;
  }
  if (g->anis) {
    g->kk2 = (double *)malloc(sizeof(double) * nc);
    for (int i = 0; i < g->nxh; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k)
          g->kk2[IDX(g, i, j, k)] = g->kx[i] * g->kx[i] + g->ky[j] * g->ky[j] + g->kz[k] * g->kz[k];
  } else {
    g->kk2 = g->kn2;
  }
  fftset *fs = (fftset *)calloc(1, sizeof(fftset));
  fs->px = fft1_new(nx);
  fs->pxh = fft1_new(nx % 2 == 0 ? nx / 2 : 1);
  fs->py = fft1_new(ny);
  fs->pz = fft1_new(nz);
  g->fft = fs;
  return g;
}

void og_grid_free(og_grid *g) {
  if (!g) return;
  fftset *fs = (fftset *)g->fft;
  fft1_free(fs->px); fft1_free(fs->pxh); fft1_free(fs->py); fft1_free(fs->pz);
  free(fs);
  if (g->kk2 != g->kn2) free(g->kk2);
  free(g->kn2);
  free(g->kx); free(g->ky); free(g->kz);
  free(g);
}

/* ------------------------------------------------------------------------ */
/* fftp3d_real_to_complex (fftp-3/fftp3D.fpp:204-315): 2D r2c per z-plane,   */
/* transpose to out(nz,ny,nx/2+1), then c2c along z.                         */
/* ------------------------------------------------------------------------ */

void og_r2c(const og_grid *g, const double *in, ocplx *out) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, nxh = g->nxh;
  const fftset *fs = (const fftset *)g->fft;
  int nmaxl = nx > ny ? nx : ny;
  if (nz > nmaxl) nmaxl = nz;
  ocplx *carr = (ocplx *)malloc(sizeof(ocplx) * (size_t)nz * ny * nxh); /* [z][y][kx] */
#pragma omp parallel
  {
    ocplx *wa = (ocplx *)malloc(sizeof(ocplx) * (size_t)nmaxl);
    ocplx *wb = (ocplx *)malloc(sizeof(ocplx) * (size_t)nmaxl);
#pragma omp for schedule(static)
    for (int z = 0; z < nz; ++z) {
      for (int y = 0; y < ny; ++y)
        rfft_row(fs->pxh, fs->px, nx, in + ((size_t)z * ny + y) * nx,
                 carr + ((size_t)z * ny + y) * nxh, wa, wb);
      ocplx *line = (ocplx *)malloc(sizeof(ocplx) * (size_t)ny);
      for (int kx = 0; kx < nxh; ++kx) {
        for (int y = 0; y < ny; ++y) line[y] = carr[((size_t)z * ny + y) * nxh + kx];
        fft1_exec(fs->py, line, wa, -1);
        for (int y = 0; y < ny; ++y) carr[((size_t)z * ny + y) * nxh + kx] = line[y];
      }
      free(line);
    }
#pragma omp for schedule(static)
    for (int kx = 0; kx < nxh; ++kx)
      for (int y = 0; y < ny; ++y) {
        ocplx *o = out + ((size_t)kx * ny + y) * nz;
        for (int z = 0; z < nz; ++z) o[z] = carr[((size_t)z * ny + y) * nxh + kx];
        fft1_exec(fs->pz, o, wa, -1);
      }
    free(wa);
    free(wb);
  }
  free(carr);
}

/* fftp3d_complex_to_real (fftp-3/fftp3D.fpp:318-431): c2c along z, transpose,
 * 2D c2r per z-plane.  The input is copied first, so it is preserved. */
void og_c2r(const og_grid *g, const ocplx *in, double *out) {
  const int nx = g->nx, ny = g->ny, nz = g->nz, nxh = g->nxh;
  const fftset *fs = (const fftset *)g->fft;
  int nmaxl = nx > ny ? nx : ny;
  if (nz > nmaxl) nmaxl = nz;
  ocplx *tmp = (ocplx *)malloc(sizeof(ocplx) * (size_t)nxh * ny * nz);
  ocplx *carr = (ocplx *)malloc(sizeof(ocplx) * (size_t)nz * ny * nxh);
#pragma omp parallel
  {
    ocplx *wa = (ocplx *)malloc(sizeof(ocplx) * (size_t)nmaxl);
    ocplx *wb = (ocplx *)malloc(sizeof(ocplx) * (size_t)nmaxl);
#pragma omp for schedule(static)
    for (int kx = 0; kx < nxh; ++kx)
      for (int y = 0; y < ny; ++y) {
        ocplx *t = tmp + ((size_t)kx * ny + y) * nz;
        memcpy(t, in + ((size_t)kx * ny + y) * nz, sizeof(ocplx) * (size_t)nz);
        fft1_exec(fs->pz, t, wa, +1);
      }
#pragma omp for schedule(static)
    for (int z = 0; z < nz; ++z) {
      for (int y = 0; y < ny; ++y)
        for (int kx = 0; kx < nxh; ++kx)
          carr[((size_t)z * ny + y) * nxh + kx] = tmp[((size_t)kx * ny + y) * nz + z];
      ocplx *line = (ocplx *)malloc(sizeof(ocplx) * (size_t)ny);
      for (int kx = 0; kx < nxh; ++kx) {
        for (int y = 0; y < ny; ++y) line[y] = carr[((size_t)z * ny + y) * nxh + kx];
        fft1_exec(fs->py, line, wa, +1);
        for (int y = 0; y < ny; ++y) carr[((size_t)z * ny + y) * nxh + kx] = line[y];
      }
      free(line);
      for (int y = 0; y < ny; ++y)
        irfft_row(fs->pxh, fs->px, nx, carr + ((size_t)z * ny + y) * nxh,
                  out + ((size_t)z * ny + y) * nx, wa, wb);
    }
    free(wa);
    free(wb);
  }
  free(carr);
  free(tmp);
}

/* ------------------------------------------------------------------------ */
/* pseudospec3D_hd.f90 pointwise operators                                  */
/* ------------------------------------------------------------------------ */

static inline size_t ncplx(const og_grid *g) { return (size_t)g->nxh * g->ny * g->nz; }
static inline size_t nreal(const og_grid *g) { return (size_t)g->nx * g->ny * g->nz; }

/* i*k*a evaluated as GHOST's im*k*a: (-k*Im a, k*Re a) */
static inline ocplx ika(double k, ocplx a) { return -k * cimag(a) + I * (k * creal(a)); }

/* derivk3 (:22-88) */
void og_derivk3(const og_grid *g, const ocplx *a, ocplx *b, int dir) {
  OMP_FOR
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        double kk = dir == 1 ? g->kx[i] : (dir == 2 ? g->ky[j] : g->kz[k]);
        b[IDX(g, i, j, k)] = ika(kk, a[IDX(g, i, j, k)]);
      }
}

/* laplak3 (:91-121) */
void og_laplak3(const og_grid *g, const ocplx *a, ocplx *b) {
  const size_t n = ncplx(g);
  OMP_FOR
  for (size_t m = 0; m < n; ++m) b[m] = -g->kk2[m] * a[m];
}

/* rotor3 (:124-200): dir 1: d_y b - d_z a; dir 2: d_z a - d_x b; dir 3: d_x b - d_y a */
void og_rotor3(const og_grid *g, const ocplx *a, const ocplx *b, ocplx *c, int dir) {
  OMP_FOR
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        size_t m = IDX(g, i, j, k);
        ocplx c1, c2;
        if (dir == 1) { c1 = ika(g->kz[k], a[m]); c2 = ika(g->ky[j], b[m]); c[m] = c2 - c1; }
        else if (dir == 2) { c1 = ika(g->kz[k], a[m]); c2 = ika(g->kx[i], b[m]); c[m] = c1 - c2; }
        else { c1 = ika(g->ky[j], a[m]); c2 = ika(g->kx[i], b[m]); c[m] = c2 - c1; }
      }
}

static inline double nsix(const og_grid *g) {
  double t = (double)g->nx * (double)g->ny * (double)g->nz;
  return 1.0 / (t * t);
}

/* real-space cross product of prodre3/vector3 (:377-389, mhd :80-92):
 * r7=(r2 r6-r5 r3)tmp, r3=(r3 r4-r6 r1)tmp, r1=(r1 r5-r4 r2)tmp */
static void cross_real(const og_grid *g, double *r1, double *r2, double *r3, const double *r4,
                       const double *r5, const double *r6, double *r7) {
  const double tmp = nsix(g);
  const size_t n = nreal(g);
  OMP_FOR
  for (size_t m = 0; m < n; ++m) {
    double a = (r2[m] * r6[m] - r5[m] * r3[m]) * tmp;
    double b = (r3[m] * r4[m] - r6[m] * r1[m]) * tmp;
    double c = (r1[m] * r5[m] - r4[m] * r2[m]) * tmp;
    r7[m] = a; r3[m] = b; r1[m] = c;
  }
}

/* prodre3 (:316-396): FFT[curl(A) x A] / N^6 */
void og_prodre3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, ocplx *d,
                ocplx *e, ocplx *f) {
  const size_t nr = nreal(g), nc = ncplx(g);
  double *r = (double *)malloc(sizeof(double) * nr * 7);
  ocplx *t = (ocplx *)malloc(sizeof(ocplx) * nc);
  og_rotor3(g, b, c, t, 1); og_c2r(g, t, r + 0 * nr);
  og_rotor3(g, a, c, t, 2); og_c2r(g, t, r + 1 * nr);
  og_rotor3(g, a, b, t, 3); og_c2r(g, t, r + 2 * nr);
  og_c2r(g, a, r + 3 * nr);
  og_c2r(g, b, r + 4 * nr);
  og_c2r(g, c, r + 5 * nr);
  cross_real(g, r, r + nr, r + 2 * nr, r + 3 * nr, r + 4 * nr, r + 5 * nr, r + 6 * nr);
  og_r2c(g, r + 6 * nr, d);
  og_r2c(g, r + 2 * nr, e);
  og_r2c(g, r + 0 * nr, f);
  free(t);
  free(r);
}

/* nonlhd3 (:399-509): g = -a_dir + k_dir (k.a)/k^2, with g = -a_dir on the
 * k_dir = 0 plane (index 1 in GHOST) */
void og_nonlhd3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, ocplx *gout,
                int dir) {
  OMP_FOR
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        size_t m = IDX(g, i, j, k);
        ocplx s = dir == 1 ? a[m] : (dir == 2 ? b[m] : c[m]);
        int plane = dir == 1 ? i : (dir == 2 ? j : k);
        if (plane == 0) {
          gout[m] = -s;
        } else {
          double kd = dir == 1 ? g->kx[i] : (dir == 2 ? g->ky[j] : g->kz[k]);
          ocplx dot = g->kx[i] * a[m] + g->ky[j] * b[m] + g->kz[k] * c[m];
          gout[m] = -s + kd * dot / g->kk2[m];
        }
      }
}

/* nonlin3 (mhd :108-227): g = -a_dir + d_dir + k_dir (k.(A-D))/k^2 */
void og_nonlin3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                const ocplx *e, const ocplx *f, ocplx *gout, int dir) {
  OMP_FOR
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        size_t m = IDX(g, i, j, k);
        ocplx s = dir == 1 ? a[m] : (dir == 2 ? b[m] : c[m]);
        ocplx sd = dir == 1 ? d[m] : (dir == 2 ? e[m] : f[m]);
        int plane = dir == 1 ? i : (dir == 2 ? j : k);
        if (plane == 0) {
          gout[m] = -s + sd;
        } else {
          double kd = dir == 1 ? g->kx[i] : (dir == 2 ? g->ky[j] : g->kz[k]);
          ocplx dot = g->kx[i] * (a[m] - d[m]) + g->ky[j] * (b[m] - e[m]) + g->kz[k] * (c[m] - f[m]);
          gout[m] = -s + sd + kd * dot / g->kk2[m];
        }
      }
}

/* gauge3 (mhd :230-336): g = a_dir - k_dir (k.a)/k^2 */
void og_gauge3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, ocplx *gout,
               int dir) {
  OMP_FOR
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        size_t m = IDX(g, i, j, k);
        ocplx s = dir == 1 ? a[m] : (dir == 2 ? b[m] : c[m]);
        int plane = dir == 1 ? i : (dir == 2 ? j : k);
        if (plane == 0) {
          gout[m] = s;
        } else {
          double kd = dir == 1 ? g->kx[i] : (dir == 2 ? g->ky[j] : g->kz[k]);
          ocplx dot = g->kx[i] * a[m] + g->ky[j] * b[m] + g->kz[k] * c[m];
          gout[m] = s - kd * dot / g->kk2[m];
        }
      }
}

/* vector3 (mhd :22-105): FFT[A x B] / N^6 */
void og_vector3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                const ocplx *e, const ocplx *f, ocplx *x, ocplx *y, ocplx *z) {
  const size_t nr = nreal(g);
  double *r = (double *)malloc(sizeof(double) * nr * 7);
  og_c2r(g, a, r + 0 * nr);
  og_c2r(g, b, r + 1 * nr);
  og_c2r(g, c, r + 2 * nr);
  og_c2r(g, d, r + 3 * nr);
  og_c2r(g, e, r + 4 * nr);
  og_c2r(g, f, r + 5 * nr);
  cross_real(g, r, r + nr, r + 2 * nr, r + 3 * nr, r + 4 * nr, r + 5 * nr, r + 6 * nr);
  og_r2c(g, r + 6 * nr, x);
  og_r2c(g, r + 2 * nr, y);
  og_r2c(g, r + 0 * nr, z);
  free(r);
}

/* ------------------------------------------------------------------------ */
/* Global quantities                                                        */
/* ------------------------------------------------------------------------ */

static inline double abs2(ocplx z) { return creal(z) * creal(z) + cimag(z) * cimag(z); }

/* Hermitian-weighted sum over modes: weight 1 on the kx=0 plane, 2 elsewhere.
 * Each kx plane is summed by one thread, then the planes in kx order: the
 * result does not depend on the thread count or on the order in which
 * threads finish (an OpenMP reduction clause combines in no fixed order). */
#define HSUM(g, expr, out)                                                            \
  do {                                                                                \
    double *part_ = (double *)malloc(sizeof(double) * (size_t)(g)->nxh);              \
    _Pragma("omp parallel for schedule(static)")                                      \
    for (int i = 0; i < (g)->nxh; ++i) {                                              \
      double w_ = i == 0 ? 1.0 : 2.0, a_ = 0.0;                                       \
      for (int j = 0; j < (g)->ny; ++j)                                               \
        for (int k = 0; k < (g)->nz; ++k) {                                           \
          size_t m = IDX(g, i, j, k);                                                 \
          (void)m;                                                                    \
          a_ += w_ * (expr);                                                          \
        }                                                                             \
      part_[i] = a_;                                                                  \
    }                                                                                 \
    double acc_ = 0.0;                                                                \
    for (int i = 0; i < (g)->nxh; ++i) acc_ += part_[i];                              \
    free(part_);                                                                      \
    out = acc_;                                                                       \
  } while (0)

/* energy (:512-662): kin 1 <v^2>, 0 <|curl v|^2>, 2 <|lap v|^2> */
double og_energy(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, int kin) {
  const double tmp = nsix(g);
  double d;
  if (kin == 1) {
    HSUM(g, (abs2(a[m]) + abs2(b[m]) + abs2(c[m])) * tmp, d);
    return d;
  }
  size_t nc = ncplx(g);
  ocplx *c1 = (ocplx *)malloc(sizeof(ocplx) * nc * 3);
  ocplx *c2 = c1 + nc, *c3 = c1 + 2 * nc;
  if (kin == 0) {
    og_rotor3(g, b, c, c1, 1); og_rotor3(g, a, c, c2, 2); og_rotor3(g, a, b, c3, 3);
  } else {
    og_laplak3(g, a, c1); og_laplak3(g, b, c2); og_laplak3(g, c, c3);
  }
  HSUM(g, (abs2(c1[m]) + abs2(c2[m]) + abs2(c3[m])) * tmp, d);
  free(c1);
  return d;
}

/* helicity (:665-784): <v . curl v> */
double og_helicity(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c) {
  const double tmp = nsix(g);
  size_t nc = ncplx(g);
  ocplx *c1 = (ocplx *)malloc(sizeof(ocplx) * nc);
  double d = 0, t;
  og_rotor3(g, b, c, c1, 1);
  HSUM(g, creal(a[m] * conj(c1[m])) * tmp, t); d += t;
  og_rotor3(g, a, c, c1, 2);
  HSUM(g, creal(b[m] * conj(c1[m])) * tmp, t); d += t;
  og_rotor3(g, a, b, c1, 3);
  HSUM(g, creal(c[m] * conj(c1[m])) * tmp, t); d += t;
  free(c1);
  return d;
}

/* cross (:787-907): kin 1 <A.B>, kin 0 sum k^2 A.B */
double og_cross(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                const ocplx *e, const ocplx *f, int kin) {
  const double tmp = nsix(g);
  double r;
  if (kin == 1) {
    HSUM(g, creal(a[m] * conj(d[m]) + b[m] * conj(e[m]) + c[m] * conj(f[m])) * tmp, r);
  } else {
    HSUM(g, creal(a[m] * conj(d[m]) + b[m] * conj(e[m]) + c[m] * conj(f[m])) * g->kk2[m] * tmp, r);
  }
  return r;
}

/* hdcheck divergence (:1036-1071): <(div v)^2> */
double og_divergence(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c) {
  const double tmq = nsix(g);
  double r;
  HSUM(g, abs2(ika(g->kx[i], a[m]) + ika(g->ky[j], b[m]) + ika(g->kz[k], c[m])) * tmq, r);
  return r;
}

/* maxabs (:910-981) */
double og_maxabs(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, int kin) {
  size_t nc = ncplx(g), nr = nreal(g);
  ocplx *cc = (ocplx *)malloc(sizeof(ocplx) * nc * 3);
  double *r = (double *)malloc(sizeof(double) * nr * 3);
  if (kin == 0) {
    og_rotor3(g, b, c, cc, 1); og_rotor3(g, a, c, cc + nc, 2); og_rotor3(g, a, b, cc + 2 * nc, 3);
  } else if (kin == 1) {
    og_laplak3(g, a, cc); og_laplak3(g, b, cc + nc); og_laplak3(g, c, cc + 2 * nc);
  } else {
    memcpy(cc, a, sizeof(ocplx) * nc);
    memcpy(cc + nc, b, sizeof(ocplx) * nc);
    memcpy(cc + 2 * nc, c, sizeof(ocplx) * nc);
  }
  og_c2r(g, cc, r); og_c2r(g, cc + nc, r + nr); og_c2r(g, cc + 2 * nc, r + 2 * nr);
  double dl = 0.0;
  for (size_t m = 0; m < nr; ++m) {
    double v = sqrt(r[m] * r[m] + r[nr + m] * r[nr + m] + r[2 * nr + m] * r[2 * nr + m]);
    if (v > dl) dl = v;
  }
  free(cc); free(r);
  return dl / ((double)g->nx * (double)g->ny * (double)g->nz);
}

/* spectrumc (:1190-1415).  The shell offset .501 is a default-real literal in
 * GHOST, i.e. the single-precision value promoted to double. */
void og_spectrumc(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, int kin,
                  int hel, double *Ek, double *Hk) {
  const double tmp = nsix(g);
  const double off = (double)0.501f;
  const int nb = g->nmax / 2 + 1;
  size_t nc = ncplx(g);
  ocplx *cu = NULL;
  if (kin == 0 || hel == 1) {
    cu = (ocplx *)malloc(sizeof(ocplx) * nc * 3);
    og_rotor3(g, b, c, cu, 1); og_rotor3(g, a, c, cu + nc, 2); og_rotor3(g, a, b, cu + 2 * nc, 3);
  }
  if (kin <= 1 && Ek) {
    for (int q = 0; q < nb; ++q) Ek[q] = 0.0;
    const ocplx *x = kin == 1 ? a : cu, *y = kin == 1 ? b : cu + nc, *z = kin == 1 ? c : cu + 2 * nc;
    for (int i = 0; i < g->nxh; ++i)
      for (int j = 0; j < g->ny; ++j)
        for (int k = 0; k < g->nz; ++k) {
          size_t m = IDX(g, i, j, k);
          int kmn = (int)(sqrt(g->kk2[m]) / g->Dkk + off);
          if (kmn > 0 && kmn <= nb) {
            double w = i == 0 ? 1.0 : 2.0;
            Ek[kmn - 1] += w * (abs2(x[m]) + abs2(y[m]) + abs2(z[m])) * tmp;
          }
        }
  }
  if (hel == 1 && Hk) {
    for (int q = 0; q < nb; ++q) Hk[q] = 0.0;
    for (int i = 0; i < g->nxh; ++i)
      for (int j = 0; j < g->ny; ++j)
        for (int k = 0; k < g->nz; ++k) {
          size_t m = IDX(g, i, j, k);
          int kmn = (int)(sqrt(g->kk2[m]) / g->Dkk + off);
          if (kmn > 0 && kmn <= nb) {
            double w = i == 0 ? 1.0 : 2.0;
            Hk[kmn - 1] += w * (creal(a[m] * conj(cu[m])) + creal(b[m] * conj(cu[nc + m])) +
                                creal(c[m] * conj(cu[2 * nc + m]))) * tmp;
          }
        }
  }
  free(cu);
}

/* phaseshift (:2084-2150).  The Nyquist entries of the kx=0 plane are hit
 * twice (by cdump and by its conjugate), as in GHOST's loops. */
void og_phaseshift(const og_grid *g, ocplx *a, double cre, double cim) {
  const ocplx cd = cre + I * cim, jd = conj(cd);
  const int ny = g->ny, nz = g->nz;
#define A0(k, j) a[IDX(g, 0, (j), (k))]
  for (int j = 2; j <= ny / 2 + 1; ++j) {
    A0(0, j - 1) *= cd;
    A0(0, ny - j + 1) *= jd;
  }
  for (int k = 2; k <= nz / 2 + 1; ++k) {
    A0(k - 1, 0) *= cd;
    A0(nz - k + 1, 0) *= jd;
  }
  for (int j = 2; j <= ny; ++j)
    for (int k = 2; k <= nz / 2 + 1; ++k) {
      A0(k - 1, j - 1) *= cd;
      A0(nz - k + 1, ny - j + 1) *= jd;
    }
#undef A0
  OMP_FOR
  for (int i = 1; i < g->nxh; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) a[IDX(g, i, j, k)] *= cd;
}

/* normalize (module_dns.f90:13-41): f = f*f0/sqrt(energy(f,kin)) */
void og_normalize(const og_grid *g, ocplx *fx, ocplx *fy, ocplx *fz, double f0, int kin) {
  double tmp = og_energy(g, fx, fy, fz, kin);
  double s = sqrt(tmp);
  const size_t n = ncplx(g);
  OMP_FOR
  for (size_t m = 0; m < n; ++m) {
    fx[m] = fx[m] * f0 / s;
    fy[m] = fy[m] * f0 / s;
    fz[m] = fz[m] * f0 / s;
  }
}

/* randu (pseudospec3D_mod.fpp:130-148): Park-Miller minimal standard with
 * Schrage factorisation and an XOR mask; am = 1./im is a default-real
 * (single precision) constant. */
double og_randu(int32_t *idum) {
  const int32_t iq = 127773, ir = 2836, mask = 123459876, ia = 16807, im = 2147483647;
  const double am = (double)(1.0f / (float)im);
  int32_t s = *idum ^ mask;
  int32_t k = s / iq;
  s = ia * (s - k * iq) - ir * k;
  if (s < 0) s += im;
  double r = am * (double)s;
  r = (r - 0.5) * 2;
  *idum = s ^ mask;
  return r;
}

/* sharp (pseudospec3D_filt.f90:200-241) */
void og_sharp(const og_grid *g, const ocplx *a, ocplx *b, double delta) {
  const double pi = 3.14159265358979323846;
  const double tmp = (pi / delta) * (pi / delta);
  const size_t n = ncplx(g);
  OMP_FOR
  for (size_t m = 0; m < n; ++m) b[m] = g->kn2[m] <= tmp ? a[m] : 0.0;
}

/* ------------------------------------------------------------------------ */
/* Initial conditions and forcing                                           */
/* ------------------------------------------------------------------------ */

/* examples/initialv.f90_tg:22-61 */
void og_initialv_tg(const og_grid *g, double kdn, double kup, double u0, ocplx *vx, ocplx *vy,
                    ocplx *vz) {
  const double pi = 3.14159265358979323846;
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  size_t nr = nreal(g);
  double *R1 = (double *)malloc(sizeof(double) * nr);
  double *R2 = (double *)malloc(sizeof(double) * nr);
  OMP_FOR
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double r1 = 0.0, r2 = 0.0;
        for (int ki = (int)kdn; ki <= (int)kup; ++ki) {
          double ax = 2 * pi * ki * ((double)i - 1) / (double)nx;
          double ay = 2 * pi * ki * ((double)j - 1) / (double)ny;
          double az = 2 * pi * ki * ((double)k - 1) / (double)nz;
          r1 = r1 + sin(ax) * cos(ay) * cos(az);
          r2 = r2 - cos(ax) * sin(ay) * cos(az);
        }
        R1[RIDX(g, i - 1, j - 1, k - 1)] = r1;
        R2[RIDX(g, i - 1, j - 1, k - 1)] = r2;
      }
  memset(vz, 0, sizeof(ocplx) * ncplx(g));
  og_r2c(g, R1, vx);
  og_r2c(g, R2, vy);
  free(R1);
  free(R2);
  og_normalize(g, vx, vy, vz, u0, 1);
}

/* Random-phase fill of three complex fields in the loop and Hermitian-mirror
 * order of examples/initialfv.f90_random:16-135.  amp(kk2) gives the modulus
 * for kx=0 modes (kx>0 modes get 2*amp), inband(m) selects the modes drawn. */
typedef double (*amp_fn)(double kk2, const void *ctx);
typedef int (*band_fn)(const og_grid *g, size_t m, const void *ctx);

static void random_fill(const og_grid *g, amp_fn amp, band_fn inband, const void *ctx,
                        int32_t *seed, ocplx *C1, ocplx *C2, ocplx *C3) {
  const double pi = 3.14159265358979323846;
  const int ny = g->ny, nz = g->nz;
#define DRAW(dst, m, fac, dump)                                                   \
  do {                                                                            \
    double phase = 2 * pi * og_randu(seed);                                       \
    (dst)[m] = (fac) * (cos(phase) + I * sin(phase)) * (dump);                    \
  } while (0)
#define M0(k, j) IDX(g, 0, (j), (k))
  C1[M0(0, 0)] = 0; C2[M0(0, 0)] = 0; C3[M0(0, 0)] = 0;
  for (int j = 2; j <= ny / 2 + 1; ++j) {
    size_t m = M0(0, j - 1), mm = M0(0, ny - j + 1);
    if (inband(g, m, ctx)) {
      double dump = amp(g->kk2[m], ctx);
      DRAW(C1, m, 1.0, dump); C1[mm] = conj(C1[m]);
      DRAW(C2, m, 1.0, dump); C2[mm] = conj(C2[m]);
      DRAW(C3, m, 1.0, dump); C3[mm] = conj(C3[m]);
    } else {
      C1[m] = 0; C1[mm] = 0; C2[m] = 0; C2[mm] = 0; C3[m] = 0; C3[mm] = 0;
    }
  }
  for (int k = 2; k <= nz / 2 + 1; ++k) {
    size_t m = M0(k - 1, 0), mm = M0(nz - k + 1, 0);
    if (inband(g, m, ctx)) {
      double dump = amp(g->kk2[m], ctx);
      DRAW(C1, m, 1.0, dump); C1[mm] = conj(C1[m]);
      DRAW(C2, m, 1.0, dump); C2[mm] = conj(C2[m]);
      DRAW(C3, m, 1.0, dump); C3[mm] = conj(C3[m]);
    } else {
      C1[m] = 0; C1[mm] = 0; C2[m] = 0; C2[mm] = 0; C3[m] = 0; C3[mm] = 0;
    }
  }
  for (int j = 2; j <= ny; ++j)
    for (int k = 2; k <= nz / 2 + 1; ++k) {
      size_t m = M0(k - 1, j - 1), mm = M0(nz - k + 1, ny - j + 1);
      if (inband(g, m, ctx)) {
        double dump = amp(g->kk2[m], ctx);
        DRAW(C1, m, 1.0, dump); C1[mm] = conj(C1[m]);
        DRAW(C2, m, 1.0, dump); C2[mm] = conj(C2[m]);
        DRAW(C3, m, 1.0, dump); C3[mm] = conj(C3[m]);
      } else {
        C1[m] = 0; C1[mm] = 0; C2[m] = 0; C2[mm] = 0; C3[m] = 0; C3[mm] = 0;
      }
    }
  for (int i = 1; i < g->nxh; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        size_t m = IDX(g, i, j, k);
        if (inband(g, m, ctx)) {
          double dump = amp(g->kk2[m], ctx);
          DRAW(C1, m, 2, dump);
          DRAW(C2, m, 2, dump);
          DRAW(C3, m, 2, dump);
        } else {
          C1[m] = 0; C2[m] = 0; C3[m] = 0;
        }
      }
#undef M0
#undef DRAW
}

typedef struct { double kdn, kup, k0; } bandctx;
static double amp_invk(double kk2, const void *ctx) { (void)ctx; return 1. / sqrt(kk2); }
static int band_shell(const og_grid *g, size_t m, const void *ctx) {
  const bandctx *b = (const bandctx *)ctx;
  return g->kk2[m] <= b->kup * b->kup && g->kk2[m] >= b->kdn * b->kdn;
}
static double amp_gauss(double kk2, const void *ctx) {
  const bandctx *b = (const bandctx *)ctx;
  return exp(-kk2 / (b->k0 * b->k0));
}
static int band_dealiased(const og_grid *g, size_t m, const void *ctx) {
  (void)ctx;
  return g->kn2[m] <= g->kmax && g->kn2[m] >= g->tiny;
}

static void curl_normalize(const og_grid *g, ocplx *C, double f0, ocplx *fx, ocplx *fy,
                           ocplx *fz) {
  size_t nc = ncplx(g);
  og_rotor3(g, C + nc, C + 2 * nc, fx, 1);
  og_rotor3(g, C, C + 2 * nc, fy, 2);
  og_rotor3(g, C, C + nc, fz, 3);
  og_normalize(g, fx, fy, fz, f0, 1);
}

/* examples/initialfv.f90_random:16-139 */
void og_initialfv_random(const og_grid *g, double kdn, double kup, double f0, int32_t *seed,
                         ocplx *fx, ocplx *fy, ocplx *fz) {
  size_t nc = ncplx(g);
  ocplx *C = (ocplx *)calloc(nc * 3, sizeof(ocplx));
  bandctx b = {kdn, kup, 0};
  random_fill(g, amp_invk, band_shell, &b, seed, C, C + nc, C + 2 * nc);
  curl_normalize(g, C, f0, fx, fy, fz);
  free(C);
}

/* Random-phase solenoidal velocity with shell spectrum E(k) ~ k^4 exp(-2 (k/k0)^2):
 * a vector potential of modulus exp(-(k/k0)^2) on every dealiased mode, filled
 * in the order of examples/initialfv.f90_random, then curl and normalize. */
void og_initialv_k4exp(const og_grid *g, double k0, double u0, int32_t *seed, ocplx *vx,
                       ocplx *vy, ocplx *vz) {
  size_t nc = ncplx(g);
  ocplx *C = (ocplx *)calloc(nc * 3, sizeof(ocplx));
  bandctx b = {0, 0, k0};
  random_fill(g, amp_gauss, band_dealiased, &b, seed, C, C + nc, C + 2 * nc);
  curl_normalize(g, C, u0, vx, vy, vz);
  free(C);
}

/* ------------------------------------------------------------------------ */
/* Runge-Kutta steps                                                        */
/* ------------------------------------------------------------------------ */

/* Coriolis force of include/roth_rkstep2.f90:6-14 added to the products
 * before the projection: C4 - 2 omegaz vy, C5 + 2 omegaz vx */
static void coriolis(const og_grid *g, double omegaz, const ocplx *vx, const ocplx *vy, ocplx *C4, ocplx *C5) {
  const size_t nc = ncplx(g);
  OMP_FOR
  for (size_t m = 0; m < nc; ++m) {
    C4[m] = C4[m] - 2 * omegaz * vy[m];
    C5[m] = C5[m] + 2 * omegaz * vx[m];
  }
}

/* include/hd_rkstep2.f90:3-32 for one value of o; with omegaz != 0 the
 * rotating-frame step include/roth_rkstep2.f90:1-49 */
static void hd_substep_rot(const og_grid *g, int o, double dt, double nu, double omegaz, ocplx *vx, ocplx *vy,
                           ocplx *vz, const ocplx *c1, const ocplx *c2, const ocplx *c3, const ocplx *fx,
                           const ocplx *fy, const ocplx *fz) {
  size_t nc = ncplx(g);
  ocplx *W = (ocplx *)malloc(sizeof(ocplx) * nc * 5);
  ocplx *C4 = W, *C5 = W + nc, *C6 = W + 2 * nc, *C7 = W + 3 * nc, *C8 = W + 4 * nc;
  og_prodre3(g, vx, vy, vz, C4, C5, C6);
  if (omegaz != 0.0) coriolis(g, omegaz, vx, vy, C4, C5);
  og_nonlhd3(g, C4, C5, C6, C7, 1);
  og_nonlhd3(g, C4, C5, C6, C8, 2);
  og_nonlhd3(g, C4, C5, C6, C4, 3);
  og_laplak3(g, vx, vx);
  og_laplak3(g, vy, vy);
  og_laplak3(g, vz, vz);
  const double rmp = 1. / (double)o;
  OMP_FOR
  for (size_t m = 0; m < nc; ++m) {
    if (g->kn2[m] <= g->kmax && g->kn2[m] >= g->tiny) {
      ocplx f1 = fx ? fx[m] : 0, f2 = fy ? fy[m] : 0, f3 = fz ? fz[m] : 0;
      vx[m] = c1[m] + dt * (nu * vx[m] + C7[m] + f1) * rmp;
      vy[m] = c2[m] + dt * (nu * vy[m] + C8[m] + f2) * rmp;
      vz[m] = c3[m] + dt * (nu * vz[m] + C4[m] + f3) * rmp;
    } else {
      vx[m] = 0; vy[m] = 0; vz[m] = 0;
    }
  }
  free(W);
}

void og_hd_substep(const og_grid *g, int o, double dt, double nu, ocplx *vx, ocplx *vy, ocplx *vz,
                   const ocplx *c1, const ocplx *c2, const ocplx *c3, const ocplx *fx,
                   const ocplx *fy, const ocplx *fz) {
  hd_substep_rot(g, o, dt, nu, 0.0, vx, vy, vz, c1, c2, c3, fx, fy, fz);
}

void og_roth_substep(const og_grid *g, int o, double dt, double nu, double omegaz, ocplx *vx, ocplx *vy,
                     ocplx *vz, const ocplx *c1, const ocplx *c2, const ocplx *c3, const ocplx *fx,
                     const ocplx *fy, const ocplx *fz) {
  hd_substep_rot(g, o, dt, nu, omegaz, vx, vy, vz, c1, c2, c3, fx, fy, fz);
}

/* main3D.fpp:3202-3231 with include/roth_rkstep1.f90 and roth_rkstep2.f90 */
void og_roth_step(const og_grid *g, int ord, double dt, double nu, double omegaz, ocplx *vx, ocplx *vy,
                  ocplx *vz, const ocplx *fx, const ocplx *fy, const ocplx *fz) {
  size_t nc = ncplx(g);
  ocplx *C = (ocplx *)malloc(sizeof(ocplx) * nc * 3);
  memcpy(C, vx, sizeof(ocplx) * nc);
  memcpy(C + nc, vy, sizeof(ocplx) * nc);
  memcpy(C + 2 * nc, vz, sizeof(ocplx) * nc);
  for (int o = ord; o >= 1; --o)
    hd_substep_rot(g, o, dt, nu, omegaz, vx, vy, vz, C, C + nc, C + 2 * nc, fx, fy, fz);
  free(C);
}

/* main3D.fpp:3202-3231 with include/hd_rkstep1.f90 */
void og_hd_step(const og_grid *g, int ord, double dt, double nu, ocplx *vx, ocplx *vy, ocplx *vz,
                const ocplx *fx, const ocplx *fy, const ocplx *fz) {
  size_t nc = ncplx(g);
  ocplx *C = (ocplx *)malloc(sizeof(ocplx) * nc * 3);
  memcpy(C, vx, sizeof(ocplx) * nc);
  memcpy(C + nc, vy, sizeof(ocplx) * nc);
  memcpy(C + 2 * nc, vz, sizeof(ocplx) * nc);
  for (int o = ord; o >= 1; --o)
    og_hd_substep(g, o, dt, nu, vx, vy, vz, C, C + nc, C + 2 * nc, fx, fy, fz);
  free(C);
}

/* main3D.fpp:3202-3231 with include/mhd_rkstep1.f90 and mhd_rkstep2.f90:4-65 */
void og_mhd_step(const og_grid *g, int ord, double dt, double nu, double mu, ocplx *vx, ocplx *vy,
                 ocplx *vz, ocplx *ax, ocplx *ay, ocplx *az, const ocplx *fx, const ocplx *fy,
                 const ocplx *fz, const ocplx *mx, const ocplx *my, const ocplx *mz) {
  size_t nc = ncplx(g);
  ocplx *W = (ocplx *)malloc(sizeof(ocplx) * nc * 17);
  ocplx *C[18];
  for (int q = 1; q <= 17; ++q) C[q] = W + (size_t)(q - 1) * nc;
  memcpy(C[1], vx, sizeof(ocplx) * nc); memcpy(C[2], vy, sizeof(ocplx) * nc);
  memcpy(C[3], vz, sizeof(ocplx) * nc); memcpy(C[4], ax, sizeof(ocplx) * nc);
  memcpy(C[5], ay, sizeof(ocplx) * nc); memcpy(C[6], az, sizeof(ocplx) * nc);
  for (int o = ord; o >= 1; --o) {
    og_rotor3(g, ay, az, C[7], 1);
    og_rotor3(g, ax, az, C[8], 2);
    og_rotor3(g, ax, ay, C[9], 3);
    og_prodre3(g, vx, vy, vz, C[10], C[11], C[12]);
    og_prodre3(g, C[7], C[8], C[9], C[13], C[14], C[15]);
    og_nonlin3(g, C[10], C[11], C[12], C[13], C[14], C[15], C[16], 1);
    og_nonlin3(g, C[10], C[11], C[12], C[13], C[14], C[15], C[17], 2);
    og_nonlin3(g, C[10], C[11], C[12], C[13], C[14], C[15], C[10], 3);
    og_vector3(g, vx, vy, vz, C[7], C[8], C[9], C[11], C[12], C[13]);
    og_gauge3(g, C[11], C[12], C[13], C[7], 1);
    og_gauge3(g, C[11], C[12], C[13], C[8], 2);
    og_gauge3(g, C[11], C[12], C[13], C[9], 3);
    og_laplak3(g, vx, vx); og_laplak3(g, vy, vy); og_laplak3(g, vz, vz);
    og_laplak3(g, ax, ax); og_laplak3(g, ay, ay); og_laplak3(g, az, az);
    const double rmp = 1. / (double)o;
    OMP_FOR
    for (size_t m = 0; m < nc; ++m) {
      if (g->kn2[m] <= g->kmax && g->kn2[m] >= g->tiny) {
        ocplx f1 = fx ? fx[m] : 0, f2 = fy ? fy[m] : 0, f3 = fz ? fz[m] : 0;
        ocplx m1 = mx ? mx[m] : 0, m2 = my ? my[m] : 0, m3 = mz ? mz[m] : 0;
        vx[m] = C[1][m] + dt * (nu * vx[m] + C[16][m] + f1) * rmp;
        vy[m] = C[2][m] + dt * (nu * vy[m] + C[17][m] + f2) * rmp;
        vz[m] = C[3][m] + dt * (nu * vz[m] + C[10][m] + f3) * rmp;
        ax[m] = C[4][m] + dt * (mu * ax[m] + C[7][m] + m1) * rmp;
        ay[m] = C[5][m] + dt * (mu * ay[m] + C[8][m] + m2) * rmp;
        az[m] = C[6][m] + dt * (mu * az[m] + C[9][m] + m3) * rmp;
      } else {
        vx[m] = 0; vy[m] = 0; vz[m] = 0; ax[m] = 0; ay[m] = 0; az[m] = 0;
      }
    }
  }
  free(W);
}

/* ======================================================================
 * Lagrangian tracers (3D/src/lagpart/), single rank, cubic splines,
 * voxel-database exchange (GHOST's default ilgexchtype), no internal
 * acceleration (dopacc = 0).  Positions are in grid units, [0, n).
 * ====================================================================== */

/* Factorisation of the cyclic tridiagonal system (1/6, 2/3, 1/6) of the
 * periodic cubic B-spline, GPSplineInt_MatInvQ (lagpart/gpsplineint_mod.fpp:
 * 1077-1125).  Index 1..n as in GHOST; element 0 unused. */
typedef struct {
  int n;
  double *bet, *p, *gam, *xx;
  double zeta, a, c;
} cyc_fac;

static cyc_fac *cyc_new(int n) {
  cyc_fac *f = (cyc_fac *)calloc(1, sizeof(cyc_fac));
  f->n = n;
  f->bet = (double *)calloc((size_t)n + 2, sizeof(double));
  f->p = (double *)calloc((size_t)n + 2, sizeof(double));
  f->gam = (double *)calloc((size_t)n + 2, sizeof(double));
  f->xx = (double *)calloc((size_t)n + 2, sizeof(double));
  const double a = 1.0 / 6.0, c = 1.0 / 6.0, b = 2.0 / 3.0; /* sub, super, diagonal */
  f->a = a; f->c = c;
  double beta;
  f->bet[1] = 1. / b;
  f->p[1] = a * f->bet[1];
  f->xx[1] = c;
  beta = b;
  for (int i = 2; i <= n - 2; ++i) {
    f->gam[i] = c * f->bet[i - 1];
    f->bet[i] = 1. / (b - a * f->gam[i]);
    f->p[i] = -f->p[i - 1] * a * f->bet[i];
    beta = beta - f->xx[i - 1] * f->p[i - 1];
    f->xx[i] = -f->xx[i - 1] * f->gam[i];
  }
  f->gam[n - 1] = c * f->bet[n - 2];
  f->bet[n - 1] = 1. / (b - a * f->gam[n - 1]);
  f->gam[n] = (c - f->p[n - 2] * a) * f->bet[n - 1];
  f->zeta = a - f->xx[n - 2] * f->gam[n - 1];
  beta = beta - f->xx[n - 2] * f->p[n - 2];
  f->bet[n] = 1. / (beta - f->zeta * f->gam[n]);
  return f;
}
static void cyc_free(cyc_fac *f) {
  free(f->bet); free(f->p); free(f->gam); free(f->xx); free(f);
}

/* Solve one line in place (elements s[(i-1)*st], i = 1..n), in the order of
 * the x-sweep of GPSplineInt_CompSpline3D (gpsplineint_mod.fpp:1200-1262).
 * ghost != 0 keeps GHOST's row-n update with xx(n-2)*s(n-1) (:1240); ghost
 * == 0 uses xx(n-2)*s(n-2), which makes the solve exact.  The two differ by
 * a multiple of xx(n-2) ~ (2-sqrt 3)^(n-2): below 1e-17 for n >= 32. */
static void cyc_solve(const cyc_fac *f, double *s0, size_t st, double *w, int ghost) {
  const int n = f->n;
  double *s = w; /* 1-based work copy */
  for (int i = 1; i <= n; ++i) s[i] = s0[(size_t)(i - 1) * st];
  double fn1 = s[n - 1];
  s[1] = s[1] * f->bet[1];
  /* s[n] keeps f(n) */
  for (int i = 2; i <= n - 2; ++i) s[i] = (s[i] - f->a * s[i - 1]) * f->bet[i];
  for (int i = 2; i <= n - 2; ++i) s[n] = s[n] - f->xx[i - 1] * s[i - 1];
  s[n - 1] = (fn1 - f->a * s[n - 2]) * f->bet[n - 1];
  s[n] = s[n] - f->xx[n - 2] * (ghost ? s[n - 1] : s[n - 2]);
  s[n] = (s[n] - s[n - 1] * f->zeta) * f->bet[n];
  s[n - 1] = s[n - 1] - f->gam[n] * s[n];
  for (int i = n - 2; i >= 1; --i) s[i] = s[i] - f->gam[i + 1] * s[i + 1] - f->p[i] * s[n];
  for (int i = 1; i <= n; ++i) s0[(size_t)(i - 1) * st] = s[i];
}

/* GPSplineInt_CompSpline3D (gpsplineint_mod.fpp:1158-1579): spline control
 * points of a real field r[z][y][x], solved along x, then y, then z.  The
 * z-sweep runs on GHOST's transposed copy there; the arithmetic per line is
 * the same.  coef may alias field.  Needs n >= 4 along each axis. */
void og_spline_coef(const og_grid *g, const double *field, double *coef, int ghost) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const size_t nxy = (size_t)nx * ny;
  if (coef != field) memcpy(coef, field, sizeof(double) * nxy * nz);
  cyc_fac *fx = cyc_new(nx), *fy = cyc_new(ny), *fz = cyc_new(nz);
  const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
#pragma omp parallel
  {
    double *w = (double *)malloc(sizeof(double) * ((size_t)nmax + 2));
#pragma omp for schedule(static)
    for (long l = 0; l < (long)ny * nz; ++l) cyc_solve(fx, coef + (size_t)l * nx, 1, w, ghost);
#pragma omp for schedule(static)
    for (long l = 0; l < (long)nx * nz; ++l) {
      long k = l / nx, i = l % nx;
      cyc_solve(fy, coef + (size_t)k * nxy + i, (size_t)nx, w, ghost);
    }
#pragma omp for schedule(static)
    for (long l = 0; l < (long)nxy; ++l) cyc_solve(fz, coef + l, nxy, w, ghost);
    free(w);
  }
  cyc_free(fx); cyc_free(fy); cyc_free(fz);
}

/* Interpolation weights of GPSplineInt_Interp3D (gpsplineint_mod.fpp:268-283):
 * w[0..2] in GHOST's expressions, w[3] = 1 - w0 - w1 - w2 (:393-404). */
static inline void spline_w(double t, double w[4]) {
  const double sixth = 1.0 / 6.0, four = 4.0, three = 3.0, six = 6.0;
  const double tm = 1.0 - t;
  w[0] = sixth * tm * tm * tm;
  w[1] = sixth * (four + t * t * (three * t - six));
  w[2] = sixth * (four + tm * tm * (three * tm - six));
  w[3] = 1.0 - w[0] - w[1] - w[2];
}

/* Stencil of GPSplineInt_PartUpdate3D (gpsplineint_mod.fpp:560-667) for one
 * coordinate.  x and y: points floor-1 .. floor+2 (periodic), t = x - floor.
 * z with zcompat: GHOST builds the spline with nzghost = intorder = 3
 * (gparts_mod.f90:291 passes intorder where GPSplineInt_ctor expects the
 * ghost count, gpsplineint_mod.fpp:159) while the extended field holds 2
 * ghost planes per side (gpartcomm_mod.f90:1005-1057, :2632-2670): the
 * control-point index is klg = clamp(int(z + 3), 1, nz + 1) into planes that
 * map to field plane klg - 3 (0-based), i.e. points floor .. floor+3.  Below
 * z = nz - 1 that interpolates the field at z + 1; in the top cell the clamp
 * gives t in [1, 2).  zcompat == 0 treats z like x and y. */
static inline void stencil_xy(double x, int n, int idx[4], double *t) {
  int i0 = (int)x; /* Fortran INT: truncation, x >= 0 */
  *t = x - (double)i0;
  idx[1] = i0 % n;
  idx[2] = (idx[1] + 1) % n;
  idx[3] = (idx[2] + 1) % n;
  idx[0] = (n + idx[1] - 1) % n;
}
static inline void stencil_z(double z, int n, int zcompat, int idx[4], double *t) {
  if (!zcompat) { stencil_xy(z, n, idx, t); return; }
  const double zs = z + 3.0; /* (zp - xbnds(3,1)) * dxi with xbnds(3,1) = -3, dxi = 1 */
  int kl = (int)zs;
  if (kl > n + 1) kl = n + 1;
  if (kl < 1) kl = 1;
  *t = zs - (double)kl;
  for (int m = 0; m < 4; ++m) idx[m] = (((kl - 3 + m) % n) + n) % n;
}

/* GPSplineInt_Interp3D (gpsplineint_mod.fpp:227-480) of control points coef
 * at np points, summed in GHOST's order (x fastest, then y, then z). */
void og_spline_interp(const og_grid *g, const double *coef, long np, const double *px, const double *py,
                      const double *pz, int zcompat, double *out) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const size_t nxy = (size_t)nx * ny;
  OMP_FOR
  for (long q = 0; q < np; ++q) {
    int ix[4], iy[4], iz[4];
    double tx, ty, tz, wx[4], wy[4], wz[4];
    stencil_xy(px[q], nx, ix, &tx);
    stencil_xy(py[q], ny, iy, &ty);
    stencil_z(pz[q], nz, zcompat, iz, &tz);
    spline_w(tx, wx);
    spline_w(ty, wy);
    spline_w(tz, wz);
    double acc = 0.0;
    for (int c = 0; c < 4; ++c)
      for (int b = 0; b < 4; ++b)
        for (int a = 0; a < 4; ++a)
          acc = acc + coef[iz[c] * nxy + (size_t)iy[b] * nx + ix[a]] * wx[a] * wy[b] * wz[c];
    out[q] = acc;
  }
}

/* GPart_MakePeriodicP (gparts_mod.f90:2270-2318): p = modulo(p + 2 n, n). */
static inline double wrap_p(double p, double n) { return fmod(p + 2.0 * n, n); }

/* One tracer stage, main3D.fpp:3233-3267 + GPart_StepRKK (gparts_mod.f90:
 * 2046-2145) + GPart_EndStageRKK (:1911-2043, voxel-database branch):
 *   R_c = c2r(v_c / (nx ny nz)); u_c = Interp3D(CompSpline3D(R_c)) at the
 *   current positions (all three components at the positions the stage
 *   started from); x_c = x0_c + (dt xk n_c / (2 pi L_c)) u_c; x and y are
 *   wrapped twice (StepRKK, then EndStageRKK), z once.
 * lv (3*np, optional) receives the interpolated velocities. */
void og_tracer_stage(const og_grid *g, const ocplx *vx, const ocplx *vy, const ocplx *vz, long np,
                     double *px, double *py, double *pz, const double *x0, const double *y0,
                     const double *z0, double dt, double xk, int zcompat, int ghost_solve, double *lv) {
  const size_t nc = (size_t)g->nxh * g->ny * g->nz, nr = (size_t)g->nx * g->ny * g->nz;
  const double rmp = 1.0 / ((double)g->nx * (double)g->ny * (double)g->nz);
  ocplx *C7 = (ocplx *)malloc(sizeof(ocplx) * nc);
  double *R = (double *)malloc(sizeof(double) * nr);
  double *u = (double *)malloc(sizeof(double) * 3 * (size_t)np);
  const ocplx *v[3] = {vx, vy, vz};
  for (int c = 0; c < 3; ++c) {
    OMP_FOR
    for (size_t m = 0; m < nc; ++m) C7[m] = v[c][m] * rmp;
    og_c2r(g, C7, R);
    og_spline_coef(g, R, R, ghost_solve);
    og_spline_interp(g, R, np, px, py, pz, zcompat, u + (size_t)c * np);
  }
  const double L[3] = {g->Lx, g->Ly, g->Lz};
  const int n[3] = {g->nx, g->ny, g->nz};
  double dtf[3];
  for (int c = 0; c < 3; ++c) dtf[c] = dt * xk * ((double)n[c] / (2 * M_PI * L[c]));
  OMP_FOR
  for (long q = 0; q < np; ++q) {
    double x = x0[q] + dtf[0] * u[q];
    double y = y0[q] + dtf[1] * u[np + q];
    double z = z0[q] + dtf[2] * u[2 * np + q];
    x = wrap_p(x, (double)n[0]);
    y = wrap_p(y, (double)n[1]);
    x = wrap_p(x, (double)n[0]);
    y = wrap_p(y, (double)n[1]);
    z = wrap_p(z, (double)n[2]);
    px[q] = x; py[q] = y; pz[q] = z;
  }
  if (lv) memcpy(lv, u, sizeof(double) * 3 * (size_t)np);
  free(C7); free(R); free(u);
}

/* HD RK step with tracers: GPart_SetStepRKK (gparts_mod.f90:1839-1908)
 * stores x0 = x, then every fluid substep o = ord..1 is followed by a tracer
 * stage with xk = 1/o on the velocity that substep produced
 * (main3D.fpp:3202-3267). */
void og_hd_step_tracers(const og_grid *g, int ord, double dt, double nu, ocplx *vx, ocplx *vy, ocplx *vz,
                        const ocplx *fx, const ocplx *fy, const ocplx *fz, long np, double *px, double *py,
                        double *pz, int zcompat, int ghost_solve) {
  size_t nc = ncplx(g);
  ocplx *C = (ocplx *)malloc(sizeof(ocplx) * nc * 3);
  double *x0 = (double *)malloc(sizeof(double) * 3 * (size_t)np);
  memcpy(C, vx, sizeof(ocplx) * nc);
  memcpy(C + nc, vy, sizeof(ocplx) * nc);
  memcpy(C + 2 * nc, vz, sizeof(ocplx) * nc);
  memcpy(x0, px, sizeof(double) * np);
  memcpy(x0 + np, py, sizeof(double) * np);
  memcpy(x0 + 2 * np, pz, sizeof(double) * np);
  for (int o = ord; o >= 1; --o) {
    og_hd_substep(g, o, dt, nu, vx, vy, vz, C, C + nc, C + 2 * nc, fx, fy, fz);
    og_tracer_stage(g, vx, vy, vz, np, px, py, pz, x0, x0 + np, x0 + 2 * np, dt, 1.0 / (double)o, zcompat,
                    ghost_solve, NULL);
  }
  free(C);
  free(x0);
}

/* ======================================================================
 * Boussinesq (pseudo/pseudospec3D_phd.f90, include/bouss_rkstep{1,2}.f90)
 * ====================================================================== */

/* advect3 (pseudospec3D_phd.f90:23-115): e = FFT[-(a . grad d)] / N^6, the
 * three products accumulated in GHOST's order (x, then y, then z, then the
 * scale). */
void og_advect3(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                ocplx *e) {
  const size_t nr = nreal(g), nc = ncplx(g);
  double *r1 = (double *)malloc(sizeof(double) * nr);
  double *r2 = (double *)malloc(sizeof(double) * nr);
  double *r3 = (double *)malloc(sizeof(double) * nr);
  ocplx *c2 = (ocplx *)malloc(sizeof(ocplx) * nc);
  const ocplx *comp[3] = {a, b, c};
  const double t = (double)g->nx * (double)g->ny * (double)g->nz;
  const double tmp = -1.0 / (t * t);
  for (int dir = 1; dir <= 3; ++dir) {
    og_derivk3(g, d, c2, dir);
    og_c2r(g, comp[dir - 1], r1);
    og_c2r(g, c2, r2);
    OMP_FOR
    for (size_t m = 0; m < nr; ++m) {
      if (dir == 1) r3[m] = r1[m] * r2[m];
      else if (dir == 2) r3[m] = r3[m] + r1[m] * r2[m];
      else r3[m] = (r3[m] + r1[m] * r2[m]) * tmp;
    }
  }
  og_r2c(g, r3, e);
  free(r1); free(r2); free(r3); free(c2);
}

/* One Boussinesq substep o, include/bouss_rkstep2.f90:1-71 (the ord = 3
 * correction of :73-145 is og_bouss_ord3_correction).  State (vx,vy,vz,th)
 * updated in place from the step-start copies (c1,c2,c3,c20); forcing
 * (fx,fy,fz,fs) may be NULL. */
static void bouss_substep_rot(const og_grid *g, int o, double dt, double nu, double kappa, double xmom,
                              double xtemp, double omegaz, ocplx *vx, ocplx *vy, ocplx *vz, ocplx *th,
                              const ocplx *c1, const ocplx *c2, const ocplx *c3, const ocplx *c20,
                              const ocplx *fx, const ocplx *fy, const ocplx *fz, const ocplx *fs) {
  size_t nc = ncplx(g);
  ocplx *W = (ocplx *)malloc(sizeof(ocplx) * nc * 5);
  ocplx *C4 = W, *C5 = W + nc, *C6 = W + 2 * nc, *C7 = W + 3 * nc, *C8 = W + 4 * nc;
  og_prodre3(g, vx, vy, vz, C4, C5, C6);
  /* include/rotbouss_rkstep2.f90:6-15: Coriolis, then gravity */
  if (omegaz != 0.0) coriolis(g, omegaz, vx, vy, C4, C5);
  OMP_FOR
  for (size_t m = 0; m < nc; ++m) C6[m] = C6[m] + xmom * th[m]; /* gravity */
  og_nonlhd3(g, C4, C5, C6, C7, 1);
  og_nonlhd3(g, C4, C5, C6, C8, 2);
  og_nonlhd3(g, C4, C5, C6, C4, 3);
  og_advect3(g, vx, vy, vz, th, C5);
  OMP_FOR
  for (size_t m = 0; m < nc; ++m) C5[m] = C5[m] + xtemp * vz[m]; /* heat current */
  og_laplak3(g, vx, vx);
  og_laplak3(g, vy, vy);
  og_laplak3(g, vz, vz);
  og_laplak3(g, th, th);
  const double rmp = 1. / (double)o;
  OMP_FOR
  for (size_t m = 0; m < nc; ++m) {
    const double k2 = g->kn2[m];
    if (k2 <= g->kmax && k2 >= g->tiny) {
      ocplx f1 = fx ? fx[m] : 0, f2 = fy ? fy[m] : 0, f3 = fz ? fz[m] : 0, f4 = fs ? fs[m] : 0;
      vx[m] = c1[m] + dt * (nu * vx[m] + C7[m] + f1) * rmp;
      vy[m] = c2[m] + dt * (nu * vy[m] + C8[m] + f2) * rmp;
      vz[m] = c3[m] + dt * (nu * vz[m] + C4[m] + f3) * rmp;
      th[m] = c20[m] + dt * (kappa * th[m] + C5[m] + f4) * rmp;
    } else if (k2 > g->kmax) {
      vx[m] = 0; vy[m] = 0; vz[m] = 0; th[m] = 0;
    } else { /* kn2 < tiny: the mean of th is kept */
      vx[m] = 0; vy[m] = 0; vz[m] = 0; th[m] = c20[m];
    }
  }
  free(W);
}

void og_bouss_substep(const og_grid *g, int o, double dt, double nu, double kappa, double xmom, double xtemp,
                      ocplx *vx, ocplx *vy, ocplx *vz, ocplx *th, const ocplx *c1, const ocplx *c2,
                      const ocplx *c3, const ocplx *c20, const ocplx *fx, const ocplx *fy, const ocplx *fz,
                      const ocplx *fs) {
  bouss_substep_rot(g, o, dt, nu, kappa, xmom, xtemp, 0.0, vx, vy, vz, th, c1, c2, c3, c20, fx, fy, fz, fs);
}

/* include/bouss_rkstep2.f90:73-145, applied after the last substep when
 * ord == 3: u += 2 N(L(C) + N(C)) dt^3/24 with N the projected nonlinear term
 * including the buoyancy of the step-start temperature c20, L the Laplacian
 * (without nu, as written there).  c1..c3 are overwritten. */
static void bouss_ord3_correction(const og_grid *g, double dt, double xmom, ocplx *vx, ocplx *vy, ocplx *vz,
                                  ocplx *c1, ocplx *c2, ocplx *c3, const ocplx *c20) {
  size_t nc = ncplx(g);
  ocplx *W = (ocplx *)malloc(sizeof(ocplx) * nc * 5);
  ocplx *C4 = W, *C5 = W + nc, *C6 = W + 2 * nc, *C7 = W + 3 * nc, *C8 = W + 4 * nc;
  og_prodre3(g, c1, c2, c3, C4, C5, C6);
  OMP_FOR
  for (size_t m = 0; m < nc; ++m) C6[m] = C6[m] + xmom * c20[m];
  og_nonlhd3(g, C4, C5, C6, C7, 1);
  og_nonlhd3(g, C4, C5, C6, C8, 2);
  og_nonlhd3(g, C4, C5, C6, C4, 3);
  og_laplak3(g, c1, c1);
  og_laplak3(g, c2, c2);
  og_laplak3(g, c3, c3);
  OMP_FOR
  for (size_t m = 0; m < nc; ++m) {
    if (g->kn2[m] <= g->kmax && g->kn2[m] >= g->tiny) {
      c1[m] = c1[m] + C7[m];
      c2[m] = c2[m] + C8[m];
      c3[m] = c3[m] + C4[m];
    } else {
      c1[m] = 0; c2[m] = 0; c3[m] = 0;
    }
  }
  og_prodre3(g, c1, c2, c3, C4, C5, C6);
  OMP_FOR
  for (size_t m = 0; m < nc; ++m) C6[m] = C6[m] + xmom * c20[m];
  og_nonlhd3(g, C4, C5, C6, C7, 1);
  og_nonlhd3(g, C4, C5, C6, C8, 2);
  og_nonlhd3(g, C4, C5, C6, C4, 3);
  const double rmp = dt * dt * dt / 24.0;
  OMP_FOR
  for (size_t m = 0; m < nc; ++m) {
    vx[m] = vx[m] + 2.0 * C7[m] * rmp;
    vy[m] = vy[m] + 2.0 * C8[m] * rmp;
    vz[m] = vz[m] + 2.0 * C4[m] * rmp;
  }
  free(W);
}

/* main3D.fpp:3202-3231 with include/bouss_rkstep1.f90 and bouss_rkstep2.f90 */
void og_bouss_step(const og_grid *g, int ord, double dt, double nu, double kappa, double xmom, double xtemp,
                   ocplx *vx, ocplx *vy, ocplx *vz, ocplx *th, const ocplx *fx, const ocplx *fy,
                   const ocplx *fz, const ocplx *fs) {
  size_t nc = ncplx(g);
  ocplx *C = (ocplx *)malloc(sizeof(ocplx) * nc * 4);
  memcpy(C, vx, sizeof(ocplx) * nc);
  memcpy(C + nc, vy, sizeof(ocplx) * nc);
  memcpy(C + 2 * nc, vz, sizeof(ocplx) * nc);
  memcpy(C + 3 * nc, th, sizeof(ocplx) * nc);
  for (int o = ord; o >= 1; --o) {
    og_bouss_substep(g, o, dt, nu, kappa, xmom, xtemp, vx, vy, vz, th, C, C + nc, C + 2 * nc, C + 3 * nc, fx,
                     fy, fz, fs);
    if (ord == 3 && o <= 1)
      bouss_ord3_correction(g, dt, xmom, vx, vy, vz, C, C + nc, C + 2 * nc, C + 3 * nc);
  }
  free(C);
}

/* main3D.fpp:3202-3231 with include/rotbouss_rkstep1.f90 and
 * rotbouss_rkstep2.f90 (no ord = 3 correction there) */
void og_rotbouss_step(const og_grid *g, int ord, double dt, double nu, double kappa, double xmom, double xtemp,
                      double omegaz, ocplx *vx, ocplx *vy, ocplx *vz, ocplx *th, const ocplx *fx,
                      const ocplx *fy, const ocplx *fz, const ocplx *fs) {
  size_t nc = ncplx(g);
  ocplx *C = (ocplx *)malloc(sizeof(ocplx) * nc * 4);
  memcpy(C, vx, sizeof(ocplx) * nc);
  memcpy(C + nc, vy, sizeof(ocplx) * nc);
  memcpy(C + 2 * nc, vz, sizeof(ocplx) * nc);
  memcpy(C + 3 * nc, th, sizeof(ocplx) * nc);
  for (int o = ord; o >= 1; --o)
    bouss_substep_rot(g, o, dt, nu, kappa, xmom, xtemp, omegaz, vx, vy, vz, th, C, C + nc, C + 2 * nc,
                      C + 3 * nc, fx, fy, fz, fs);
  free(C);
}

/* analysis/trunc3D.fpp:190-320 spectral core: the spectrum c of a field on
 * grid g (r2c, unnormalised) is truncated to the modes with
 * kx^2/nxt^2 + ky^2/nyt^2 + kz^2/nzt^2 <= 1/9 (wavenumbers of g, normalised by
 * the new sizes), scaled by 1/(nx ny nz) and stored on the smaller grid t
 * (kz and ky re-indexed as trunc3D's T1 fill), so that og_c2r on t gives the
 * truncated field sampled on t's points. */
void og_truncate(const og_grid *g, const ocplx *c, const og_grid *t, ocplx *out) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int nxt = t->nx, nyt = t->ny, nzt = t->nz, nxht = t->nxh;
  const double kmax = 1.0 / 9.0;
  const double rmp = 1.0 / ((double)nxt * (double)nxt), rmq = 1.0 / ((double)nyt * (double)nyt),
               rms = 1.0 / ((double)nzt * (double)nzt);
  const double fact = 1.0 / ((double)nx * (double)ny * (double)nz);
  OMP_FOR
  for (long i = 0; i < nxht; ++i)
    for (int jt = 0; jt < nyt; ++jt)
      for (int kt = 0; kt < nzt; ++kt) {
        const int j = jt < nyt / 2 ? jt : jt - nyt + ny;
        const int k = kt < nzt / 2 ? kt : kt - nzt + nz;
        /* trunc3D's kx, ky, kz: the integer wavenumbers of the old grid */
        const double kx = (double)(i < nx / 2 ? i : i - nx), ky = (double)(j < ny / 2 ? j : j - ny),
                     kz = (double)(k < nz / 2 ? k : k - nz);
        const double kn2 = rmp * (kx * kx) + rmq * (ky * ky) + rms * (kz * kz);
        ocplx v = c[((size_t)i * ny + j) * nz + k];
        if (kn2 > kmax) v = 0;
        out[((size_t)i * nyt + jt) * nzt + kt] = v * fact;
      }
}

/* ------------------------------------------------------------------------
 * The closure network of the learned LES (no counterpart in GHOST): a
 * 10 -> 64 -> 64 -> 6 perceptron, h = GELU(W x + b) with the exact GELU
 * x Phi(x), parameters W1[64][10], b1[64], W2[64][64], b2[64], W3[6][64],
 * b3[6] in one float array.  With bf16 != 0 every operand of a matrix product
 * is first rounded to bfloat16 (round to nearest, ties to even), as the matrix
 * cores see it: the inputs, the weights, the hidden activations and, in the
 * backward pass, the output error and the pre-activation errors of the
 * hidden layers; sums are taken in double.  Biases and their gradients are
 * not rounded.  bf16 = 0 is plain double arithmetic.
 * ------------------------------------------------------------------------ */
enum { MNI = 10, MNH = 64, MNO = 6 };
enum { MW1 = 0, MB1 = MW1 + MNH * MNI, MW2 = MB1 + MNH, MB2 = MW2 + MNH * MNH, MW3 = MB2 + MNH,
       MB3 = MW3 + MNO * MNH, MNPAR = MB3 + MNO };

static double bf16_round(double x, int on) {
  if (!on) return x;
  float f = (float)x; /* the GPU holds these values in FP32 before packing */
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return f; /* inf / nan unchanged */
  const uint32_t lsb = (u >> 16) & 1u;
  u = (u + 0x7fffu + lsb) & 0xffff0000u;
  memcpy(&f, &u, 4);
  return f;
}
static double gelu_o(double x) { return 0.5 * x * (1.0 + erf(x * 0.70710678118654752440)); }
static double gelu_do(double x) {
  return 0.5 * (1.0 + erf(x * 0.70710678118654752440)) + x * 0.39894228040143267794 * exp(-0.5 * x * x);
}

/* one point: z1, h1, z2, h2 and y (y, z1, z2 in double) */
static void mlp_point(const float *P, const float *x, int bf, double *z1, double *h1, double *z2, double *h2,
                      double *y) {
  double xb[MNI];
  for (int k = 0; k < MNI; ++k) xb[k] = bf16_round(x[k], bf);
  for (int m = 0; m < MNH; ++m) {
    double s = 0.0;
    for (int k = 0; k < MNI; ++k) s += bf16_round(P[MW1 + m * MNI + k], bf) * xb[k];
    z1[m] = s + P[MB1 + m];
    h1[m] = bf16_round(gelu_o(z1[m]), bf);
  }
  for (int m = 0; m < MNH; ++m) {
    double s = 0.0;
    for (int k = 0; k < MNH; ++k) s += bf16_round(P[MW2 + m * MNH + k], bf) * h1[k];
    z2[m] = s + P[MB2 + m];
    h2[m] = bf16_round(gelu_o(z2[m]), bf);
  }
  for (int j = 0; j < MNO; ++j) {
    double s = 0.0;
    for (int k = 0; k < MNH; ++k) s += bf16_round(P[MW3 + j * MNH + k], bf) * h2[k];
    y[j] = s + P[MB3 + j];
  }
}

void og_mlp_forward(const float *P, long n, const float *X, int bf16, double *Y) {
  OMP_FOR
  for (long p = 0; p < n; ++p) {
    double z1[MNH], h1[MNH], z2[MNH], h2[MNH];
    mlp_point(P, X + p * MNI, bf16, z1, h1, z2, h2, Y + p * MNO);
  }
}

/* loss = mean over the n points and 6 outputs of (y - t)^2; grad (MNPAR
 * doubles) receives d loss / d parameter */
double og_mlp_grad(const float *P, long n, const float *X, const float *T, int bf16, double *grad) {
  memset(grad, 0, sizeof(double) * MNPAR);
  double loss = 0.0;
  const double sc = 2.0 / ((double)n * MNO);
  for (long p = 0; p < n; ++p) {
    double z1[MNH], h1[MNH], z2[MNH], h2[MNH], y[MNO], dy[MNO], dyb[MNO], dz2[MNH], dz2b[MNH], dz1[MNH];
    const float *x = X + p * MNI;
    mlp_point(P, x, bf16, z1, h1, z2, h2, y);
    for (int j = 0; j < MNO; ++j) {
      const double e = y[j] - T[p * MNO + j];
      loss += e * e;
      dy[j] = (double)(float)(sc * e); /* the GPU forms this in FP32 */
      dyb[j] = bf16_round(dy[j], bf16);
    }
    for (int m = 0; m < MNH; ++m) {
      double s = 0.0;
      for (int j = 0; j < MNO; ++j) s += bf16_round(P[MW3 + j * MNH + m], bf16) * dyb[j];
      dz2[m] = s * gelu_do(z2[m]);
      dz2b[m] = bf16_round(dz2[m], bf16);
    }
    for (int m = 0; m < MNH; ++m) {
      double s = 0.0;
      for (int u = 0; u < MNH; ++u) s += bf16_round(P[MW2 + u * MNH + m], bf16) * dz2b[u];
      dz1[m] = s * gelu_do(z1[m]);
    }
    for (int j = 0; j < MNO; ++j) {
      grad[MB3 + j] += dy[j];
      for (int k = 0; k < MNH; ++k) grad[MW3 + j * MNH + k] += dyb[j] * h2[k];
    }
    for (int m = 0; m < MNH; ++m) {
      grad[MB2 + m] += dz2[m];
      for (int k = 0; k < MNH; ++k) grad[MW2 + m * MNH + k] += dz2b[m] * h1[k];
      grad[MB1 + m] += dz1[m];
      const double d1 = bf16_round(dz1[m], bf16);
      for (int k = 0; k < MNI; ++k) grad[MW1 + m * MNI + k] += d1 * bf16_round(x[k], bf16);
    }
  }
  return loss / ((double)n * MNO);
}

/* ------------------------------------------------------------------------
 * Transfer and 1D spectra, MHD checks (pseudospec3D_hd.f90:1418-1949,
 * pseudospec3D_mhd.f90:339-712).  The shell sums return the reduced totals
 * (GHOST writes them divided by Dkk, or 0.5/Dn for spectr1d, to its files).
 * ------------------------------------------------------------------------ */
/* sum over shells int(sqrt(kk2)/Dkk + off) in 1..nmax/2+1 of
 * w (Re(x d*) + Re(y e*) + Re(z f*)) tmp, times kk2 when kk2w */
static void shell_dot(const og_grid *g, const ocplx *x, const ocplx *y, const ocplx *z, const ocplx *d,
                      const ocplx *e, const ocplx *f, int kk2w, double off, double *out) {
  const double tmp = nsix(g);
  const int nb = g->nmax / 2 + 1;
  for (int q = 0; q < nb; ++q) out[q] = 0.0;
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        size_t m = IDX(g, i, j, k);
        int kmn = (int)(sqrt(g->kk2[m]) / g->Dkk + off);
        if (kmn > 0 && kmn <= nb) {
          double s = creal(x[m] * conj(d[m])) + creal(y[m] * conj(e[m])) + creal(z[m] * conj(f[m]));
          if (kk2w) s = g->kk2[m] * s;
          out[kmn - 1] += (i == 0 ? 1.0 : 2.0) * s * tmp;
        }
      }
}

/* entrans (:1629-1820): kin = 0 weights each mode by kk2 */
void og_entrans(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                const ocplx *e, const ocplx *f, int kin, double *Ek) {
  shell_dot(g, a, b, c, d, e, f, kin == 0, (double)0.501f, Ek);
}

/* heltrans (:1823-1949): curl(a,b,c) against (d,e,f) */
void og_heltrans(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                 const ocplx *e, const ocplx *f, double *Hk) {
  size_t nc = ncplx(g);
  ocplx *cu = (ocplx *)malloc(sizeof(ocplx) * nc * 3);
  og_rotor3(g, b, c, cu, 1); og_rotor3(g, a, c, cu + nc, 2); og_rotor3(g, a, b, cu + 2 * nc, 3);
  shell_dot(g, cu, cu + nc, cu + 2 * nc, d, e, f, 0, (double)0.501f, Hk);
  free(cu);
}

/* crosspecc (pseudospec3D_mhd.f90:603-712): Re(v . conj(curl A)), shells
 * with the double-precision offset shift + .501 */
void og_crosspecc(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
                  const ocplx *e, const ocplx *f, double shift, double *Ck) {
  size_t nc = ncplx(g);
  ocplx *cu = (ocplx *)malloc(sizeof(ocplx) * nc * 3);
  og_rotor3(g, e, f, cu, 1); og_rotor3(g, d, f, cu + nc, 2); og_rotor3(g, d, e, cu + 2 * nc, 3);
  shell_dot(g, a, b, c, cu, cu + nc, cu + 2 * nc, 0, shift + 0.501, Ck);
  free(cu);
}

/* spectr1d (:1418-1626): planes of k_dir, bin int(|k_dir| L_dir + 1) in
 * 1..n_dir/2+1, |a|^2 / N^6 with the Hermitian weight; Ek has nmax/2+1 */
void og_spectr1d(const og_grid *g, const ocplx *a, int dir, double *Ek) {
  const double tmp = nsix(g);
  for (int q = 0; q < g->nmax / 2 + 1; ++q) Ek[q] = 0.0;
  const int lim = dir == 1 ? g->nx / 2 + 1 : dir == 2 ? g->ny / 2 + 1 : g->nz / 2 + 1;
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        const double kd = dir == 1 ? fabs(g->kx[i]) * g->Lx : dir == 2 ? fabs(g->ky[j]) * g->Ly : fabs(g->kz[k]) * g->Lz;
        const int kmn = (int)(kd + 1);
        if (kmn <= lim) Ek[kmn - 1] += (i == 0 ? 1.0 : 2.0) * abs2(a[IDX(g, i, j, k)]) * tmp;
      }
}

/* ---- scalar diagnostics (pseudospec3D_phd.f90) ------------------------- */

/* variance (:116-217): <th^2> (kin = 1) or <kk2 |th|^2> (kin = 0) */
double og_variance(const og_grid *g, const ocplx *a, int kin) {
  const double tmp = nsix(g);
  double acc = 0.0;
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        const size_t m = IDX(g, i, j, k);
        const double w = i == 0 ? 1.0 : 2.0;
        acc += kin == 1 ? w * tmp * abs2(a[m]) : w * tmp * g->kk2[m] * abs2(a[m]);
      }
  return acc;
}

/* product (:219-284): <a b> = sum Re(a b*) / N^6 */
double og_product(const og_grid *g, const ocplx *a, const ocplx *b) {
  const double tmp = nsix(g);
  double acc = 0.0;
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        const size_t m = IDX(g, i, j, k);
        acc += (i == 0 ? 1.0 : 2.0) * tmp * creal(a[m] * conj(b[m]));
      }
  return acc;
}

/* spectrscc (:516-610): |th|^2 per shell int(sqrt(kk2)/Dkk + shift + .501),
 * the offset in working precision; sctrans (:612-727): Re(a b*) per shell
 * with the single-precision .501 (dot != 0) */
static void scalar_shells(const og_grid *g, const ocplx *a, const ocplx *b, double off, double *Ek) {
  const double tmp = nsix(g);
  const int nb = g->nmax / 2 + 1;
  for (int q = 0; q < nb; ++q) Ek[q] = 0.0;
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        const size_t m = IDX(g, i, j, k);
        const int kmn = (int)(sqrt(g->kk2[m]) / g->Dkk + off);
        if (kmn > 0 && kmn <= nb) {
          const double s = b ? creal(a[m] * conj(b[m])) : abs2(a[m]);
          Ek[kmn - 1] += (i == 0 ? 1.0 : 2.0) * tmp * s;
        }
      }
}
void og_spectrscc(const og_grid *g, const ocplx *a, double shift, double *Ek) {
  scalar_shells(g, a, NULL, shift + 0.501, Ek);
}
void og_sctrans(const og_grid *g, const ocplx *a, const ocplx *b, double *Ek) {
  scalar_shells(g, a, b, (double)0.501f, Ek);
}

/* difucx / difucz (:729-940): plane means of v_i, th and v_i th as functions
 * of x (dir = 1) or z (dir = 3); out = {<v>(s), <th>(s), <v th>(s)}, n each */
void og_difuc(const og_grid *g, const ocplx *a, const ocplx *b, int dir, double *out) {
  const size_t nr = (size_t)g->nx * g->ny * g->nz;
  double *r1 = (double *)malloc(sizeof(double) * nr), *r2 = (double *)malloc(sizeof(double) * nr);
  og_c2r(g, a, r1);
  og_c2r(g, b, r2);
  const int n = dir == 1 ? g->nx : g->nz;
  for (int q = 0; q < 3 * n; ++q) out[q] = 0.0;
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        const size_t m = ((size_t)k * g->ny + j) * g->nx + i;
        const int s = dir == 1 ? i : k;
        out[s] += r1[m];
        out[n + s] += r2[m];
        out[2 * n + s] += r1[m] * r2[m];
      }
  const double nx = g->nx, ny = g->ny, nz = g->nz;
  const double tmp = dir == 1 ? 1.0 / (nx * (ny * ny) * (nz * nz)) : 1.0 / ((nx * nx) * (ny * ny) * nz);
  const double tmq = dir == 1 ? 1.0 / ((nx * nx) * (ny * ny * ny) * (nz * nz * nz))
                              : 1.0 / ((nx * nx * nx) * (ny * ny * ny) * (nz * nz));
  for (int q = 0; q < n; ++q) {
    out[q] *= tmp;
    out[n + q] *= tmp;
    out[2 * n + q] *= tmq;
  }
  free(r1);
  free(r2);
}

/* ---- Boussinesq diagnostics (pseudospec3D_bouss.f90) ------------------- */

/* horizontal (x, y) sums per z plane of f(r1, r2, r3, r4) for kinds
 * 0: r1^2 + r2^2, 1: r1^2, 2: r1 r2, 3: -r1 r2 + r3 r4, 4: r3 (r1 - r2) */
static void zplane_sums(const og_grid *g, int kind, const double *r1, const double *r2, const double *r3,
                        const double *r4, double *sh) {
  const size_t pl = (size_t)g->nx * g->ny;
  for (int k = 0; k < g->nz; ++k) {
    double acc = 0.0;
    for (size_t t = 0; t < pl; ++t) {
      const size_t m = (size_t)k * pl + t;
      switch (kind) {
        case 0: acc += r1[m] * r1[m] + r2[m] * r2[m]; break;
        case 1: acc += r1[m] * r1[m]; break;
        case 2: acc += r1[m] * r2[m]; break;
        case 3: acc += -(r1[m] * r2[m]) + r3[m] * r4[m]; break;
        default: acc += r3[m] * (r1[m] - r2[m]); break;
      }
    }
    sh[k] = acc;
  }
}

/* havgcomp (:19-555): horizontal averages as functions of z, gsh[nz].
 * itype 0 shear <(du_perp/dz)^2>, 1 <(dth/dz)^2>, 2 <u_z dth/dz>, 3
 * <u^2 + v^2>, 4 <w^2>, 5 <u_perp . (curl u)_perp>, 6 <omega_z th>, 7 <PV^2>
 * with PV = fo dth/dz - bv omega_z + omega . grad th - fo bv, 8 <omega . curl
 * omega>, 9 the shear of itype 0 (GHOST's first branch serves 0 and 9 and
 * returns before the Richardson number is formed), 10 <u_z th>.  For itype 2
 * GHOST multiplies a work array that this call never fills; here the
 * documented correlation with dth/dz is formed. */
void og_havgcomp(const og_grid *g, const ocplx *u, const ocplx *v, const ocplx *w, const ocplx *s, double fo,
                 double bv, int itype, double *gsh) {
  const size_t nc = ncplx(g), nr = (size_t)g->nx * g->ny * g->nz;
  const double nx = g->nx, ny = g->ny, nz = g->nz;
  const double tmp = 1.0 / (nx * nx * nx * ny * ny * ny * nz * nz);
  ocplx *c = (ocplx *)malloc(sizeof(ocplx) * nc * 4);
  ocplx *c1 = c, *c2 = c + nc, *c3 = c + 2 * nc, *c4 = c + 3 * nc;
  double *r = (double *)malloc(sizeof(double) * nr * 4);
  double *r1 = r, *r2 = r + nr, *r3 = r + 2 * nr, *r4 = r + 3 * nr;
  if (itype == 0 || itype == 9) {
    og_derivk3(g, u, c1, 3); og_c2r(g, c1, r1);
    og_derivk3(g, v, c1, 3); og_c2r(g, c1, r2);
    zplane_sums(g, 0, r1, r2, NULL, NULL, gsh);
  } else if (itype == 1 || itype == 2) {
    og_derivk3(g, s, c1, 3); og_c2r(g, c1, r3);
    if (itype == 1) {
      zplane_sums(g, 1, r3, NULL, NULL, NULL, gsh);
    } else {
      og_c2r(g, w, r2);
      zplane_sums(g, 2, r3, r2, NULL, NULL, gsh);
    }
  } else if (itype == 3) {
    og_c2r(g, u, r1); og_c2r(g, v, r2);
    zplane_sums(g, 0, r1, r2, NULL, NULL, gsh);
  } else if (itype == 4) {
    og_c2r(g, w, r1);
    zplane_sums(g, 1, r1, NULL, NULL, NULL, gsh);
  } else if (itype == 5) {
    og_c2r(g, u, r1); og_derivk3(g, v, c1, 3); og_c2r(g, c1, r2);
    og_c2r(g, v, r3); og_derivk3(g, u, c1, 3); og_c2r(g, c1, r4);
    zplane_sums(g, 3, r1, r2, r3, r4, gsh);
  } else if (itype == 6 || itype == 10) {
    if (itype == 6) { og_rotor3(g, u, v, c1, 3); og_c2r(g, c1, r1); }
    else og_c2r(g, w, r1);
    og_c2r(g, s, r2);
    zplane_sums(g, 2, r1, r2, NULL, NULL, gsh);
  } else if (itype == 7) {
    const double t = 1.0 / (nx * ny * nz);
    og_derivk3(g, s, c1, 3); og_c2r(g, c1, r1);          /* dth/dz */
    og_rotor3(g, u, v, c1, 3); og_c2r(g, c1, r3);        /* omega_z */
    og_rotor3(g, v, w, c1, 1); og_rotor3(g, u, w, c2, 2); og_rotor3(g, u, v, c3, 3);
    og_advect3(g, c1, c2, c3, s, c4); og_c2r(g, c4, r4); /* -omega . grad th */
    for (size_t m = 0; m < nr; ++m) {
      double pv = 0.0 + fo * r1[m] * t;
      pv = pv - bv * r3[m] * t;
      pv = pv + r4[m] * (-t);
      pv = pv - fo * bv;
      r2[m] = pv;
    }
    zplane_sums(g, 1, r2, NULL, NULL, NULL, gsh);
    for (int k = 0; k < g->nz; ++k) gsh[k] *= 1.0 / (nx * ny);
    free(c); free(r);
    return;
  } else if (itype == 8) {
    double *sa = (double *)malloc(sizeof(double) * g->nz * 3);
    /* omega_x (d omega_z/dy - d omega_y/dz) */
    og_rotor3(g, u, v, c1, 3); og_derivk3(g, c1, c2, 2); og_c2r(g, c2, r1);
    og_rotor3(g, u, w, c1, 2); og_derivk3(g, c1, c2, 3); og_c2r(g, c2, r2);
    og_rotor3(g, v, w, c1, 1); og_c2r(g, c1, r3);
    zplane_sums(g, 4, r1, r2, r3, NULL, sa);
    /* omega_y (d omega_x/dz - d omega_z/dx) */
    og_rotor3(g, v, w, c1, 1); og_derivk3(g, c1, c2, 3); og_c2r(g, c2, r1);
    og_rotor3(g, u, v, c1, 3); og_derivk3(g, c1, c2, 1); og_c2r(g, c2, r2);
    og_rotor3(g, u, w, c1, 2); og_c2r(g, c1, r3);
    zplane_sums(g, 4, r1, r2, r3, NULL, sa + g->nz);
    /* omega_z (d omega_y/dx - d omega_x/dy) */
    og_rotor3(g, u, w, c1, 2); og_derivk3(g, c1, c2, 1); og_c2r(g, c2, r1);
    og_rotor3(g, v, w, c1, 1); og_derivk3(g, c1, c2, 2); og_c2r(g, c2, r2);
    og_rotor3(g, u, v, c1, 3); og_c2r(g, c1, r3);
    zplane_sums(g, 4, r1, r2, r3, NULL, sa + 2 * g->nz);
    for (int k = 0; k < g->nz; ++k) gsh[k] = sa[k] + sa[g->nz + k] + sa[2 * g->nz + k];
    free(sa);
  } else {
    for (int k = 0; k < g->nz; ++k) gsh[k] = 0.0;
  }
  for (int k = 0; k < g->nz; ++k) gsh[k] *= tmp;
  free(c);
  free(r);
}

/* spectpv (:712-812): |omega . grad th|^2 over planes of |kz|, nz/2+1 bins */
void og_spectpv(const og_grid *g, const ocplx *u, const ocplx *v, const ocplx *w, const ocplx *s, double *Ek) {
  const size_t nc = ncplx(g);
  ocplx *c = (ocplx *)malloc(sizeof(ocplx) * nc * 4);
  og_rotor3(g, v, w, c, 1); og_rotor3(g, u, w, c + nc, 2); og_rotor3(g, u, v, c + 2 * nc, 3);
  og_advect3(g, c, c + nc, c + 2 * nc, s, c + 3 * nc);
  const double tmp = nsix(g);
  const int nb = g->nz / 2 + 1;
  for (int q = 0; q < nb; ++q) Ek[q] = 0.0;
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        const int kmn = (int)(fabs(g->kz[k]) * g->Lz + 1);
        if (kmn > 0 && kmn <= nb) Ek[kmn - 1] += (i == 0 ? 1.0 : 2.0) * tmp * abs2(c[3 * nc + IDX(g, i, j, k)]);
      }
  free(c);
}

/* ---- axisymmetric spectra (pseudospec3D_rot.f90), preferred axis z ----- */

/* nmaxperp of main3D.fpp:1729 */
int og_nmaxperp(const og_grid *g) {
  const double a = g->nx * g->Dkx, b = g->ny * g->Dky;
  return (int)((a > b ? a : b) / g->Dkk);
}

/* bins = 0: planes kmn = int(|kz| Lz + 1), 1..nz/2+1 (specpara :25-335);
 * bins = 1: cylinders kmn = int(sqrt(kx^2 + ky^2)/Dkk + 1), 1..nmaxperp/2+1
 * (specperp :338-686).  Per mode, with the Hermitian weight and 1/N^6:
 * kind 0: |a|^2 + |b|^2 (xy) and |c|^2 (z); kind 1: Re(a d*) + Re(b e*) and
 * Re(c f*); kind 2: kind 1 times kk2 (entpara's kin = 0).  tot gets xy + z;
 * xy and z (either may be NULL) get every mode for planes and only the
 * kz = 0 plane for cylinders, as the two files write them. */
void og_axispec(const og_grid *g, int bins, int kind, const ocplx *a, const ocplx *b, const ocplx *c,
                const ocplx *d, const ocplx *e, const ocplx *f, double *tot, double *xy, double *zz) {
  const double tmp = nsix(g);
  const int nb = bins == 0 ? g->nz / 2 + 1 : og_nmaxperp(g) / 2 + 1;
  for (int q = 0; q < nb; ++q) {
    tot[q] = 0.0;
    if (xy) xy[q] = 0.0;
    if (zz) zz[q] = 0.0;
  }
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j)
      for (int k = 0; k < g->nz; ++k) {
        const size_t m = IDX(g, i, j, k);
        const int kmn = bins == 0 ? (int)(fabs(g->kz[k]) * g->Lz + 1)
                                  : (int)(sqrt(g->kx[i] * g->kx[i] + g->ky[j] * g->ky[j]) / g->Dkk + 1);
        if (kmn <= 0 || kmn > nb) continue;
        const double w = (i == 0 ? 1.0 : 2.0) * tmp;
        double h, v;
        if (kind == 0) {
          h = abs2(a[m]) + abs2(b[m]);
          v = abs2(c[m]);
        } else {
          h = creal(a[m] * conj(d[m])) + creal(b[m] * conj(e[m]));
          v = creal(c[m] * conj(f[m]));
          if (kind == 2) { h *= g->kk2[m]; v *= g->kk2[m]; }
        }
        tot[kmn - 1] += w * (h + v);
        if (bins == 0 || k == 0) {
          if (xy) xy[kmn - 1] += w * h;
          if (zz) zz[kmn - 1] += w * v;
        }
      }
}

/* spec2D (:1331-1629): E(kperp, kpara) in Fortran order (kperp fastest),
 * nmaxperp/2+1 by nz/2+1, kind 0 or 1 as og_axispec */
void og_spec2d(const og_grid *g, int kind, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *d,
               const ocplx *e, const ocplx *f, double *E) {
  const double tmp = nsix(g);
  const int np = og_nmaxperp(g) / 2 + 1, nzb = g->nz / 2 + 1;
  for (int q = 0; q < np * nzb; ++q) E[q] = 0.0;
  for (int i = 0; i < g->nxh; ++i)
    for (int j = 0; j < g->ny; ++j) {
      const int kmn = (int)(sqrt(g->kx[i] * g->kx[i] + g->ky[j] * g->ky[j]) / g->Dkk + 1);
      if (kmn <= 0 || kmn > np) continue;
      for (int k = 0; k < g->nz; ++k) {
        const int kmz = (int)(fabs(g->kz[k]) * g->Lz + 1);
        if (kmz <= 0 || kmz > nzb) continue;
        const size_t m = IDX(g, i, j, k);
        const double s = kind == 0 ? abs2(a[m]) + abs2(b[m]) + abs2(c[m])
                                   : creal(a[m] * conj(d[m])) + creal(b[m] * conj(e[m])) + creal(c[m] * conj(f[m]));
        E[(size_t)(kmz - 1) * np + (kmn - 1)] += (i == 0 ? 1.0 : 2.0) * s * tmp;
      }
    }
}

/* mhdcheck (pseudospec3D_mhd.f90:339-547) without its file output and with
 * crs <= 1: out = {eng, ens, cur, engk, engm, helk, helm, crh, asq, divk, divm}
 * (entries not requested by hel/crs/chk are 0) */
void og_mhdcheck(const og_grid *g, const ocplx *a, const ocplx *b, const ocplx *c, const ocplx *ma,
                 const ocplx *mb, const ocplx *mc, int hel, int crs, int chk, double *out) {
  size_t nc = ncplx(g);
  for (int q = 0; q < 11; ++q) out[q] = 0.0;
  if (chk == 1) {
    out[9] = og_divergence(g, a, b, c);
    out[10] = og_divergence(g, ma, mb, mc);
  }
  const double engk = og_energy(g, a, b, c, 1), ens = og_energy(g, a, b, c, 0);
  ocplx *cu = (ocplx *)malloc(sizeof(ocplx) * nc * 3);
  og_rotor3(g, mb, mc, cu, 1); og_rotor3(g, ma, mc, cu + nc, 2); og_rotor3(g, ma, mb, cu + 2 * nc, 3);
  const double engm = og_energy(g, cu, cu + nc, cu + 2 * nc, 1), cur = og_energy(g, cu, cu + nc, cu + 2 * nc, 0);
  out[0] = engk + engm;
  out[1] = ens;
  out[2] = cur;
  out[3] = engk;
  out[4] = engm;
  if (hel == 1) {
    out[5] = og_helicity(g, a, b, c);
    out[6] = og_helicity(g, ma, mb, mc);
  }
  if (crs >= 1) {
    out[8] = og_energy(g, ma, mb, mc, 1);
    out[7] = og_cross(g, a, b, c, cu, cu + nc, cu + 2 * nc, 1);
  }
  free(cu);
}
