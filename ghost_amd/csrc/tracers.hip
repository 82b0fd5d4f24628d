// tracers.hip -- Lagrangian tracer kernels on gfx950.
//
// One lane per particle.  The stencil and weights follow
// GPSplineInt_PartUpdate3D / GPSplineInt_Interp3D (lagpart/gpsplineint_mod.fpp:
// 227-480, 560-667) and are evaluated in GHOST's operand order without fused
// multiply-adds, so that the interpolant differs from GHOST's only through the
// control points.  The control points come from the spectral prefilter of the
// c2r (one division per mode in place of GHOST's three cyclic tridiagonal
// sweeps), so the kernel only gathers: 4x4x4 points per component.
#include "tracers.h"

namespace ghx {

namespace {

GHX_DEV double mul(double a, double b) { return __dmul_rn(a, b); }
GHX_DEV double add(double a, double b) { return __dadd_rn(a, b); }
GHX_DEV double sub(double a, double b) { return __dsub_rn(a, b); }

// Weights of GPSplineInt_Interp3D (gpsplineint_mod.fpp:268-283, 393-404).
GHX_DEV void weights(double t, double w[4]) {
  const double sixth = 1.0 / 6.0;
  const double tm = sub(1.0, t);
  w[0] = mul(mul(mul(sixth, tm), tm), tm);
  w[1] = mul(sixth, add(4.0, mul(mul(t, t), sub(mul(3.0, t), 6.0))));
  w[2] = mul(sixth, add(4.0, mul(mul(tm, tm), sub(mul(3.0, tm), 6.0))));
  w[3] = sub(sub(sub(1.0, w[0]), w[1]), w[2]);
}

// Integer part of a coordinate, as Fortran INT (truncation).  Non-finite or
// huge values map to 0 so that every index below stays inside the field.
GHX_DEV int trunc_index(double x) {
  if (!(x > -1073741824.0 && x < 1073741824.0)) return 0;
  return (int)x;
}
GHX_DEV int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

// x and y (and z without the compatibility stencil): points i0-1 .. i0+2.
GHX_DEV void stencil_xy(double x, int n, int idx[4], double &t) {
  const int i0 = trunc_index(x);
  t = sub(x, (double)i0);
  idx[1] = pmod(i0, n);
  idx[2] = pmod(idx[1] + 1, n);
  idx[3] = pmod(idx[2] + 1, n);
  idx[0] = pmod(n + idx[1] - 1, n);
}

// GHOST's z stencil: klg = clamp(int(z + 3), 1, nz + 1) into a field that
// carries 2 ghost planes per side while the spline was set up for 3
// (gparts_mod.f90:291, gpsplineint_mod.fpp:159, gpartcomm_mod.f90:1005-1057):
// control point klg maps to field plane klg - 3.
GHX_DEV void stencil_z(double z, int n, int compat, int idx[4], double &t) {
  if (!compat) {
    stencil_xy(z, n, idx, t);
    return;
  }
  const double zs = add(z, 3.0);
  int kl = trunc_index(zs);
  kl = kl > n + 1 ? n + 1 : (kl < 1 ? 1 : kl);
  t = sub(zs, (double)kl);
#pragma unroll
  for (int m = 0; m < 4; ++m) idx[m] = pmod(kl - 3 + m, n);
}

struct Stencil {
  int ix[4], iy[4], iz[4];
  double wx[4], wy[4], wz[4];
};

GHX_DEV void make_stencil(Stencil &s, double x, double y, double z, int nx, int ny, int nz, int zcompat) {
  double tx, ty, tz;
  stencil_xy(x, nx, s.ix, tx);
  stencil_xy(y, ny, s.iy, ty);
  stencil_z(z, nz, zcompat, s.iz, tz);
  weights(tx, s.wx);
  weights(ty, s.wy);
  weights(tz, s.wz);
}

// 64-term sum in GHOST's order: x fastest, then y, then z; each term is
// ((c * wx) * wy) * wz.
GHX_DEV double gather(const double *__restrict__ c, const Stencil &s, int nx, size_t nxy) {
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double *plane = c + (size_t)s.iz[k] * nxy;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double *row = plane + (size_t)s.iy[j] * nx;
#pragma unroll
      for (int i = 0; i < 4; ++i) acc = add(acc, mul(mul(mul(row[s.ix[i]], s.wx[i]), s.wy[j]), s.wz[k]));
    }
  }
  return acc;
}

// The same sums for three interleaved components (c[(row nx + x) 3 + comp]):
// each stencil row of the three components is 96 contiguous bytes.
GHX_DEV void gather3(const double *__restrict__ c, const Stencil &s, int nx, size_t nxy, double acc[3]) {
  acc[0] = acc[1] = acc[2] = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double *plane = c + (size_t)s.iz[k] * nxy * 3;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double *row = plane + (size_t)s.iy[j] * nx * 3;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double *p = row + 3 * s.ix[i];
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q] = add(acc[q], mul(mul(mul(p[q], s.wx[i]), s.wy[j]), s.wz[k]));
      }
    }
  }
}

// GPart_MakePeriodicP (gparts_mod.f90:2270-2318): modulo(p + 2 n, n).
GHX_DEV double wrap(double p, double n) { return fmod(add(p, mul(2.0, n)), n); }

__global__ __launch_bounds__(256) void k_tracer_stage(TracerStage p) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.np) return;
  const size_t nxy = (size_t)p.nx * p.ny;
  Stencil s;
  make_stencil(s, p.pos[0][q], p.pos[1][q], p.pos[2][q], p.nx, p.ny, p.nz, p.zcompat);
  double u[3];
  gather3(p.coef3, s, p.nx, nxy, u);
  // GPart_StepRKK (gparts_mod.f90:2046-2145): x = x0 + dtfact u, then x and
  // y wrapped there and again in GPart_EndStageRKK (:1911-2043), z once.
  double x = add(p.x0[0][q], mul(p.dtf[0], u[0]));
  double y = add(p.x0[1][q], mul(p.dtf[1], u[1]));
  double z = add(p.x0[2][q], mul(p.dtf[2], u[2]));
  const double nx = (double)p.nx, ny = (double)p.ny, nz = (double)p.nz;
  x = wrap(wrap(x, nx), nx);
  y = wrap(wrap(y, ny), ny);
  z = wrap(z, nz);
  p.pos[0][q] = x;
  p.pos[1][q] = y;
  p.pos[2][q] = z;
  if (p.lv[0]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) p.lv[c][q] = u[c];
  }
}

__global__ __launch_bounds__(256) void k_spline_interp(const double *__restrict__ coef, const double *px,
                                                       const double *py, const double *pz, long long np, int nx,
                                                       int ny, int nz, int zcompat, double *out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  Stencil s;
  make_stencil(s, px[q], py[q], pz[q], nx, ny, nz, zcompat);
  out[q] = gather(coef, s, nx, (size_t)nx * ny);
}

__global__ __launch_bounds__(256) void k_tracer_copy(double *a, double *b, double *c, const double *x,
                                                     const double *y, const double *z, long long np) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= np) return;
  a[q] = x[q];
  b[q] = y[q];
  c[q] = z[q];
}

inline unsigned blocks_for(long long np) { return (unsigned)((np + 255) / 256); }

}  // namespace

hipError_t tracer_stage(hipStream_t s, const TracerStage &p) {
  if (p.np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tracer_stage, dim3(blocks_for(p.np)), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t spline_interp(hipStream_t s, const double *coef, const double *const *pos, long long np, int nx, int ny,
                         int nz, int zcompat, double *out) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_spline_interp, dim3(blocks_for(np)), dim3(256), 0, s, coef, pos[0], pos[1], pos[2], np, nx,
                     ny, nz, zcompat, out);
  return hipGetLastError();
}

hipError_t tracer_copy(hipStream_t s, double *const *dst, const double *const *src, long long np) {
  if (np <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tracer_copy, dim3(blocks_for(np)), dim3(256), 0, s, dst[0], dst[1], dst[2], src[0], src[1],
                     src[2], np);
  return hipGetLastError();
}

}  // namespace ghx
