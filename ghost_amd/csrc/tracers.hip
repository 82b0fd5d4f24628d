// tracers.hip -- Lagrangian tracer kernels on gfx950.
//
// The stencil and weights follow
// GPSplineInt_PartUpdate3D / GPSplineInt_Interp3D (lagpart/gpsplineint_mod.fpp:
// 227-480, 560-667); weights and terms are evaluated in GHOST's operand order
// without fused multiply-adds.  k_spline_interp also adds the 64 terms in
// GHOST's order; the tracer stage adds them in a tree (see there), which
// changes the interpolant by rounding only.  The control points come from the spectral prefilter of the
// c2r (one division per mode in place of GHOST's three cyclic tridiagonal
// sweeps), so the kernels only gather: 4x4x4 points per component.
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

// One stage of the tracer RK step, 16 lanes per particle.  The 4 x 4 stencil
// rows of the three interleaved components are 16 rows of 12 doubles (96
// contiguous bytes unless the row wraps in x), 192 doubles in all: lane r of
// the particle's group loads doubles r, r + 16, ..., r + 176, so neighbouring
// lanes read neighbouring doubles and each cache line of a row is requested
// once per load instruction (one lane per particle asked for each of the 192
// doubles separately).  Double d of row w is component (d mod 3) of stencil
// point d / 3 on row (y index w mod 4, z index w / 4); since 16 = 1 mod 3,
// the component of load m on lane r is (r + m) mod 3.  Each lane sums its
// weighted terms per component, and the group adds its 16 partial sums with
// a butterfly; the interpolant is GHOST's 64-term sum, added in that order
// within a lane and pairwise across lanes.
constexpr int kStageLanes = 16;
constexpr int kStageBlock = 256;

GHX_DEV int sel4(int i, const int v[4]) { return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3])); }
GHX_DEV double sel4(int i, const double v[4]) { return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3])); }

__global__ __launch_bounds__(kStageBlock) void k_tracer_stage(TracerStage p) {
  const int r = threadIdx.x % kStageLanes;
  const long long q = (long long)blockIdx.x * (kStageBlock / kStageLanes) + threadIdx.x / kStageLanes;
  if (q >= p.np) return;  // whole groups only: the butterfly stays inside a group
  const size_t nx = (size_t)p.nx, ny = (size_t)p.ny;
  Stencil s;
  make_stencil(s, p.pos[0][q], p.pos[1][q], p.pos[2][q], p.nx, p.ny, p.nz, p.zcompat);
  double v[12];
#pragma unroll
  for (int m = 0; m < 12; ++m) {
    const int sl = r + kStageLanes * m;  // 0..191
    const int w = sl / 12, d = sl - 12 * (sl / 12);
    const int pt = d / 3, comp = d - 3 * (d / 3);
    const size_t row = (size_t)sel4(w >> 2, s.iz) * ny + (size_t)sel4(w & 3, s.iy);
    v[m] = p.coef3[(row * nx + (size_t)sel4(pt, s.ix)) * 3 + comp];
  }
  double b[3] = {0.0, 0.0, 0.0};  // b[t]: component (r + t) mod 3
#pragma unroll
  for (int m = 0; m < 12; ++m) {
    const int sl = r + kStageLanes * m;
    const int w = sl / 12, d = sl - 12 * (sl / 12);
    const int pt = d / 3;
    b[m % 3] = add(b[m % 3], mul(mul(mul(v[m], sel4(pt, s.wx)), sel4(w & 3, s.wy)), sel4(w >> 2, s.wz)));
  }
  const int r3 = r % 3;
  double u[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int t = (c - r3 + 3) % 3;
    double a = t == 0 ? b[0] : (t == 1 ? b[1] : b[2]);
#pragma unroll
    for (int o = kStageLanes / 2; o >= 1; o >>= 1) a = add(a, __shfl_xor(a, o, kStageLanes));
    u[c] = a;
  }
  if (r != 0) return;
  // GPart_StepRKK (gparts_mod.f90:2046-2145): x = x0 + dtfact u, then x and
  // y wrapped there and again in GPart_EndStageRKK (:1911-2043), z once.
  double x = add(p.x0[0][q], mul(p.dtf[0], u[0]));
  double y = add(p.x0[1][q], mul(p.dtf[1], u[1]));
  double z = add(p.x0[2][q], mul(p.dtf[2], u[2]));
  const double fnx = (double)p.nx, fny = (double)p.ny, fnz = (double)p.nz;
  x = wrap(wrap(x, fnx), fnx);
  y = wrap(wrap(y, fny), fny);
  z = wrap(z, fnz);
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
  constexpr long long per_block = kStageBlock / kStageLanes;
  hipLaunchKernelGGL(k_tracer_stage, dim3((unsigned)((p.np + per_block - 1) / per_block)), dim3(kStageBlock), 0, s, p);
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
