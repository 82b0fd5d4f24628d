// fft_passes.h -- host-side interface of the FFT pass kernels.
//
// Layouts (complex double unless noted):
//   R  real    r[z][y][x]          GHOST R(nx,ny,nz)
//   S  spectral s[kx][ky][kz]      GHOST C(nz,ny,nx/2+1), kz fastest
//   Q  intermediate q[z][kx][ky]   ky fastest
//   P  intermediate p[z][ky][kx]   kx fastest, rows of p_row() >= nx/2+1 points
// r2c = X (R->P, contiguous rows) . Y (P->Q, transposed load) . Z (Q->S, transposed load)
// c2r = Z^-1 (S->Q, transposed store) . Y^-1 (Q->P, transposed store) . X^-1 (P->R)
#pragma once
#include "ghx_common.h"

namespace ghx {

struct FFTCtx {
  GridK gk;
  const cplx *tw_xh;  // nx/2 table
  const cplx *tw_x;   // nx table (real/half-complex packing)
  const cplx *tw_y;   // ny table
  const cplx *tw_z;   // nz table
  hipStream_t stream;
  // Batched launches: the fused-substep passes (zinv, yinv, xcross, yfwd,
  // zfwd_rk_hd) run `batch` independent fields of this grid in one launch;
  // member b's pointers are the FieldSet's plus b * bstride.  Other passes
  // take batch == 1.
  int batch = 1;
  // tile widths (lines per block) of the z and y passes; 0 = built-in choice
  int tile_z = 0, tile_y = 0;
  // radix schedules of the fused-substep passes: bit 0 the z passes (curl
  // inverse, forward update), bit 1 the y passes, on four points per lane
  // (FFTCfgLow) where that differs from the default schedule
  int sched = 0;
};

constexpr int kMaxFields = 12;
struct FieldSet {
  cplx *p[kMaxFields];
  size_t bstride = 0;  // elements between batch members
};

// Fused-epilogue parameters for the Z forward pass of an RK substep.
struct RKParams {
  double dt, nu, rmp;
  const cplx *v[3];    // state the nonlinear term was computed from (laplacian)
  const cplx *c[3];    // state at the start of the step
  cplx *out[3];        // updated state (may alias c or v element-wise)
  const cplx *f[3];    // forcing (may be null)
  const uint8_t *fline;  // per (kx,ky) line: 1 if the forcing is non-zero there
  size_t bstride = 0;       // elements between batch members (v, c, out, f)
  const double *nu_b = nullptr;  // per-member viscosity of a batched launch (else nu)
  double om2 = 0.0;  // 2 omegaz: Coriolis force of a frame rotating about z (0: none)
};

struct MHDRKParams {
  double dt, nu, mu, rmp;
  const cplx *v[6], *c[6];
  cplx *out[6];
  const cplx *f[6];
  const uint8_t *fline;
};

// Boussinesq update parameters: fields 0..2 velocity, 3 temperature.
struct BoussRKParams {
  double dt, nu, kappa, xmom, xtemp, rmp;
  const cplx *v[4], *c[4];
  cplx *out[4];
  const cplx *f[4];
  const uint8_t *fline;
  double om2 = 0.0;  // 2 omegaz (rotating frame), 0: none
};

bool fft_size_supported(int n);

// row length of the P layout: nx/2+1 rounded up to 8 points (128-byte rows)
inline int p_row(const GridK &g) { return (g.nxh + 7) & ~7; }
// complex points of one scratch field slot, large enough for Q or P
inline size_t slot_points(const GridK &g) { return (size_t)g.nz * g.ny * p_row(g); }

// Z^-1 pass, S -> Q.  mode 0: nf fields in, nf out.  mode 1 (HD): 3 in
// (v) -> 6 out (v, curl v).  mode 2 (MHD): 6 in (v, A) -> 12 out
// (v, curl v, b = curl A, j = curl b).  mode 4 (advect3): 4 in (a, b, c, d)
// -> 6 out (a, b, c, grad d).  mode 5 (Boussinesq): 4 in (v, th) -> 9 out
// (v, curl v, grad th).
// prune=1: the input is dealiased (zero outside tiny <= kn2 <= kmax): only
// lines that carry kept modes are read and written (every pass below takes
// the same flag; see the kernels for exactly what each one skips).
// mode 3: nf fields in, nf out, each mode scaled by aux[kx] aux[nxh+ky]
// aux[nxh+ny+kz] (the periodic cubic-spline prefilter of the tracers).
hipError_t zinv(const FFTCtx &c, int mode, int nf, const FieldSet &in, const FieldSet &out, int prune = 0,
                const double *aux = nullptr);
// Y^-1: Q -> P.  Y forward: P -> Q.
hipError_t yinv(const FFTCtx &c, int nf, const FieldSet &in, const FieldSet &out, int prune = 0);
hipError_t yfwd(const FFTCtx &c, int nf, const FieldSet &in, const FieldSet &out, int prune = 0);
// X^-1: P -> real rows (prune: P holds data only for kx < active_kx).
hipError_t xinv(const FFTCtx &c, int nf, const FieldSet &in, double *const *out, int prune = 0);
// X^-1 of three P fields written interleaved: out[(row * nx + x) * 3 + f].
hipError_t xinv3i(const FFTCtx &c, const FieldSet &in, double *out, int prune = 0);
// X forward: real rows -> P.
hipError_t xfwd(const FFTCtx &c, int nf, const double *const *in, const FieldSet &out);
// Fused X^-1 . cross product . X on P rows: in.p[0..2] = A, in.p[3..5] = B
// -> out.p[0..2] = FFT[(A x B)/N^6] in GHOST's operand order.
hipError_t xcross(const FFTCtx &c, int mode, const FieldSet &in, const FieldSet &out, int prune = 0);
// Z forward, Q -> S (mode 0), or fused with projection + RK update (HD/MHD).
hipError_t zfwd(const FFTCtx &c, int nf, const FieldSet &in, const FieldSet &out, int prune = 0);
hipError_t zfwd_rk_hd(const FFTCtx &c, const FieldSet &in, const RKParams &rk, int prune = 0);
hipError_t zfwd_rk_mhd(const FFTCtx &c, const FieldSet &in, const MHDRKParams &rk, int prune = 0);
// Fused X^-1 . dot product . X for advect3: in.p[0..2] = A, in.p[3..5] = B
// (P rows) -> out.p[0] = FFT[-(A . B)/N^6].
hipError_t xdot(const FFTCtx &c, const FieldSet &in, const FieldSet &out, int prune = 0);
// Z forward fused with the Boussinesq projection and update: in.p[0..2] =
// products w x u, in.p[3] = -(u . grad th), Q layout.
hipError_t zfwd_rk_bouss(const FFTCtx &c, const FieldSet &in, const BoussRKParams &rk, int prune = 0);
// number of kx indices (from 0) whose kn2(kx,0,0) <= kmax
int active_kx(const GridK &g);

}  // namespace ghx
