// ghx_common.h -- shared device helpers for the ghx HIP kernels (gfx950, FP64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define GHX_DEV __device__ __forceinline__

typedef double2 cplx;

GHX_DEV cplx c_make(double x, double y) { cplx r; r.x = x; r.y = y; return r; }
// Streaming store of one complex value (the nontemporal hint on a vector
// store): for outputs of many GB that no cache holds until they are read back.
typedef double ghx_d2 __attribute__((ext_vector_type(2)));
GHX_DEV void c_store_nt(cplx *p, cplx v) {
  ghx_d2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<ghx_d2 *>(p));
}
template <bool NT> GHX_DEV void c_put(cplx *p, cplx v) {
  if constexpr (NT) c_store_nt(p, v);
  else *p = v;
}
template <bool NT> GHX_DEV void d_put(double *p, double v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}
GHX_DEV cplx c_add(cplx a, cplx b) { return c_make(a.x + b.x, a.y + b.y); }
GHX_DEV cplx c_sub(cplx a, cplx b) { return c_make(a.x - b.x, a.y - b.y); }
GHX_DEV cplx c_mul(cplx a, cplx b) { return c_make(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
GHX_DEV cplx c_scale(cplx a, double s) { return c_make(a.x * s, a.y * s); }
GHX_DEV cplx c_conj(cplx a) { return c_make(a.x, -a.y); }
// multiply by i*k as GHOST's im*k*a: (-k Im a, k Re a)
GHX_DEV cplx c_ik(double k, cplx a) { return c_make(-k * a.y, k * a.x); }
// multiply by sign*i  (sign = -1 or +1)
template <int S> GHX_DEV cplx c_mul_si(cplx z) {
  return S < 0 ? c_make(z.y, -z.x) : c_make(-z.y, z.x);
}
// twiddle from a forward table tw[e] = exp(-2 pi i e/N); inverse uses the conjugate
template <int DIR> GHX_DEV cplx c_tw(const cplx *tw, int e) {
  cplx w = tw[e];
  return DIR < 0 ? w : c_make(w.x, -w.y);
}

// Wavenumber of index q on an axis of n points (GHOST main3D.fpp:1746-1763),
// scaled by the box factor dk: k(q) = q for q < n/2, q - n otherwise.
GHX_DEV double wavenum(int q, int n, double dk) {
  int half = n / 2;
  double k = (double)(q < half ? q : q - n);
  if (n == 1) k = 0.0;
  return k * dk;
}

// Grid constants shared by the spectral kernels.
struct GridK {
  int nx, ny, nz, nxh;
  int anis;
  double dkx, dky, dkz;   // box factors 1/L (1 when not an arbitrary box)
  double rmp, rmq, rms;   // kn2 weights (1/n^2 when anisotropic, else 1)
  double kmax, tiny;
  double inv_n6;          // 1/(nx ny nz)^2
};

// kn2 and kk2 exactly as GHOST forms them (no fused multiply-add, so the
// dealias mask matches bit for bit): main3D.fpp:1764-1800.
GHX_DEV double kn2_of(const GridK &g, int i, int j, int k) {
  double kx = wavenum(i, g.nx, 1.0), ky = wavenum(j, g.ny, 1.0), kz = wavenum(k, g.nz, 1.0);
  double a = __dmul_rn(g.rmp, __dmul_rn(kx, kx));
  double b = __dmul_rn(g.rmq, __dmul_rn(ky, ky));
  double c = __dmul_rn(g.rms, __dmul_rn(kz, kz));
  return __dadd_rn(__dadd_rn(a, b), c);
}
GHX_DEV double kk2_of(const GridK &g, double kx, double ky, double kz) {
  return __dadd_rn(__dadd_rn(__dmul_rn(kx, kx), __dmul_rn(ky, ky)), __dmul_rn(kz, kz));
}
