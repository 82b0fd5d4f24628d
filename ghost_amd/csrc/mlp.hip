// mlp.hip -- the learned subgrid closure's network on gfx950 MFMA.
//
// A 10 -> 64 -> 64 -> 6 perceptron with GELU between the layers maps the
// resolved velocity-gradient tensor (9 inputs) plus the filter width to the
// six independent components of the subgrid stress.  Weights are kept in
// FP32 (the master copy the optimiser updates) and enter the matrix cores as
// BF16; every product accumulates in FP32 (v_mfma_f32_16x16x32_bf16 for the
// layers, v_mfma_f32_16x16x16_bf16 for the weight gradients).
//
// Layout.  A wavefront takes 16 points at a time and works on transposed
// activations: a layer is H^T = W X^T with the weights as the A operand
// (rows = units of the layer) and the activations as the B operand (columns =
// points).  The accumulator of a 16-unit block holds, in lane l, units
// 4(l/16) .. 4(l/16)+3 of point l%16; the B operand of the next layer wants 8
// consecutive k-slots per lane.  Rather than moving data between lanes, the
// k-slots of the next layer are ordered to match what each lane holds: slot
// (s, q, r) of k-step s is unit 32s + 4q + r for r < 4 and 32s + 16 + 4q + r-4
// for r >= 4 (q = l/16), and the weights are loaded in that order.  The
// backward pass reuses the same ordering for W2^T and W3^T; only the weight
// gradients (contractions over points) go through an LDS transpose.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/ghx.h"

int ghx_internal_set_err(int code, const char *msg);

namespace {

constexpr int NI = 10, NH = 64, NO = 6;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NO * NH;
constexpr int NPAR = OB3 + NO;  // 5254
constexpr int NNORM = 2 * NI + 2 * NO;
static_assert(NNORM == 32, "the C ABI passes the normalisation as 32 floats: mx, sx, my, sy");
constexpr int WAVES = 4, BLOCK = 64 * WAVES, GRID = 256, NWAVES = GRID * WAVES;

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

#define DEV __device__ __forceinline__

DEV f4 mma32(bf8 a, bf8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
DEV f4 mma16(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
DEV short bfbits(float x) { return __builtin_bit_cast(short, (__bf16)x); }

// GELU(x) = x Phi(x) with Phi the standard normal distribution.  Phi(x) =
// 1/2 + x P(min(x^2, 16)), clamped to [0, 1], with P the degree-7 minimax fit
// (Lawson's iteration, weighted by x) of (Phi(x) - 1/2)/x in x^2 on [0, 16].
// Inside |x| < 4 the fit's error is below 3.2e-5; beyond, the clamped line
// reaches 0 or 1 within |x| < 4.01 and Phi differs from them by less than
// 3.2e-5, so |Phi error| < 3.2e-5 and |GELU error| < 1.3e-4 everywhere.
// Both are far below the bfloat16 rounding the activations then receive,
// and the form needs no transcendental and no branch (the clamp is the
// output modifier of the last fma): the activations are the inference
// kernel's VALU cost, 11 operations each.
DEV float phi_poly(float x) {
  const float t = fminf(x * x, 16.0f);
  float p = -1.58079339e-09f;
  p = fmaf(p, t, 1.21711508e-07f);
  p = fmaf(p, t, -4.10087569e-06f);
  p = fmaf(p, t, 8.06675016e-05f);
  p = fmaf(p, t, -1.04820507e-03f);
  p = fmaf(p, t, 9.66487639e-03f);
  p = fmaf(p, t, -6.61753789e-02f);
  p = fmaf(p, t, 3.98847520e-01f);
  return __builtin_amdgcn_fmed3f(fmaf(x, p, 0.5f), 0.0f, 1.0f);
}
DEV float gelu(float x) { return x * phi_poly(x); }
DEV float gelu_d(float x) { return phi_poly(x) + x * 0.39894228040143268f * __expf(-0.5f * x * x); }

// unit held in k-slot (s, q, r) of the 64-wide layers
DEV int uslot(int s, int q, int r) { return 32 * s + (r < 4 ? 4 * q + r : 16 + 4 * q + (r - 4)); }

struct FwdW {
  bf8 w1[4], w2[4][2], w3[2];
  f4 b1[4], b2[4], b3;
};
struct BwdW {
  bf8 w3t[4], w2t[4][2];
};

DEV void load_fwd(FwdW &W, const float *__restrict__ P, int l) {
  const int c = l & 15, q = l >> 4;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int k = 8 * q + r;
      W.w1[b][r] = (__bf16)(k < NI ? P[OW1 + (16 * b + c) * NI + k] : 0.0f);
      W.w2[b][0][r] = (__bf16)P[OW2 + (16 * b + c) * NH + uslot(0, q, r)];
      W.w2[b][1][r] = (__bf16)P[OW2 + (16 * b + c) * NH + uslot(1, q, r)];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      W.b1[b][r] = P[OB1 + 16 * b + 4 * q + r];
      W.b2[b][r] = P[OB2 + 16 * b + 4 * q + r];
    }
  }
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int r = 0; r < 8; ++r) W.w3[s][r] = (__bf16)(c < NO ? P[OW3 + c * NH + uslot(s, q, r)] : 0.0f);
#pragma unroll
  for (int r = 0; r < 4; ++r) W.b3[r] = 4 * q + r < NO ? P[OB3 + 4 * q + r] : 0.0f;
}

DEV void load_bwd(BwdW &W, const float *__restrict__ P, int l) {
  const int c = l & 15, q = l >> 4;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int m = 16 * b + c;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int j = 4 * q + r;
      W.w3t[b][r] = (__bf16)(r < 4 && j < NO ? P[OW3 + j * NH + m] : 0.0f);
      W.w2t[b][0][r] = (__bf16)P[OW2 + uslot(0, q, r) * NH + m];
      W.w2t[b][1][r] = (__bf16)P[OW2 + uslot(1, q, r) * NH + m];
    }
  }
}

DEV bf8 pack2(f4 lo, f4 hi) {
  bf8 o;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    o[r] = (__bf16)lo[r];
    o[4 + r] = (__bf16)hi[r];
  }
  return o;
}

struct Acts {
  f4 z1[4], z2[4], y;
  bf8 h1b[2], h2b[2];
};

DEV void forward(const FwdW &W, bf8 xb, Acts &a) {
#pragma unroll
  for (int b = 0; b < 4; ++b) a.z1[b] = mma32(W.w1[b], xb, W.b1[b]);
  f4 h[4];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) h[b][r] = gelu(a.z1[b][r]);
  a.h1b[0] = pack2(h[0], h[1]);
  a.h1b[1] = pack2(h[2], h[3]);
#pragma unroll
  for (int b = 0; b < 4; ++b) a.z2[b] = mma32(W.w2[b][1], a.h1b[1], mma32(W.w2[b][0], a.h1b[0], W.b2[b]));
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int r = 0; r < 4; ++r) h[b][r] = gelu(a.z2[b][r]);
  a.h2b[0] = pack2(h[0], h[1]);
  a.h2b[1] = pack2(h[2], h[3]);
  a.y = mma32(W.w3[1], a.h2b[1], mma32(W.w3[0], a.h2b[0], W.b3));
}

struct Norm {
  float mx[NI], sx[NI], my[NO], sy[NO];
};

// ---- inference on fields -------------------------------------------------
// grad9: the nine real fields d u_i / d x_j (i major), tau6: outputs in the
// order xx, yy, zz, xy, xz, yz.  Inputs are normalised with (mx, sx), outputs
// mapped back with (my, sy).
struct Ptr9 {
  const double *p[9];
};
struct Ptr6 {
  double *p[6];
};

// One wave takes 64 points per trip, one point per lane for the memory side:
// lane p loads point p's nine gradients (nine 512-byte coalesced loads; the
// next trip's loads are issued before this trip's work), normalises them and
// writes them as bf16 to its LDS feature row.  The four 16-point MFMA tiles
// then read their B operands from those rows and their weights (bf16, in
// the A-operand order of load_fwd) and biases from the block's LDS copy, and
// leave the outputs in LDS rows from which lane p takes point p's six values
// for six coalesced 512-byte stores.  Weights in LDS rather than registers
// leave room for several waves per SIMD.
constexpr int IW = 4, IBLOCK = 64 * IW;
struct InferW {
  bf8 w1[4][64], w2[4][2][64], w3[2][64];
  f4 b1[4][64], b2[4][64], b3[64];
};
struct InferLDS {
  InferW W;
  bf8 feat[IW][64][2];  // features 0..15 of each point (10 used) as bf16
  f4 out[IW][64][2];    // outputs 0..7 of each point (6 used)
};

DEV void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(IBLOCK) void k_mlp_fields(Ptr9 g, double delta, const float *__restrict__ P, Norm nm,
                                                       Ptr6 tau, long long n) {
  __shared__ InferLDS S;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, c = lane & 15, q = lane >> 4;
  {
    // the block's weights in load_fwd's operand order; wave w fills its share
    const int cc = lane & 15, qq = lane >> 4;
    auto row8 = [&](int off, int row, int ld, int kind, int s) {
      bf8 o;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int k = kind == 0 ? 8 * qq + r : uslot(s, qq, r);
        const bool ok = kind == 0 ? k < NI : (kind == 2 ? row < NO : true);
        o[r] = (__bf16)(ok ? P[off + row * ld + k] : 0.0f);
      }
      return o;
    };
    auto bias4 = [&](int off, int base, int lim) {
      f4 o;
#pragma unroll
      for (int r = 0; r < 4; ++r) o[r] = base + r < lim ? P[off + base + r] : 0.0f;
      return o;
    };
    if (w == 0) {
      for (int b = 0; b < 4; ++b) S.W.w1[b][lane] = row8(OW1, 16 * b + cc, NI, 0, 0);
      for (int s2 = 0; s2 < 2; ++s2) S.W.w3[s2][lane] = row8(OW3, cc, NH, 2, s2);
      S.W.b3[lane] = bias4(OB3, 4 * qq, NO);
    } else if (w == 1 || w == 2) {
      for (int b = 0; b < 4; ++b) S.W.w2[b][w - 1][lane] = row8(OW2, 16 * b + cc, NH, 1, w - 1);
    } else {
      for (int b = 0; b < 4; ++b) {
        S.W.b1[b][lane] = bias4(OB1, 16 * b + 4 * qq, NH);
        S.W.b2[b][lane] = bias4(OB2, 16 * b + 4 * qq, NH);
      }
    }
  }
  __syncthreads();
  const InferW &W = S.W;
  const long long groups = (n + 63) / 64;
  const long long first = (long long)blockIdx.x * IW + w, stride = (long long)gridDim.x * IW;
  const float dn = (float)(delta - (double)nm.mx[9]) / nm.sx[9];
  // the features' scales as reciprocals: one multiply per feature and trip
  // instead of a division (the product differs from the quotient by at most
  // an ulp of float, far below the bfloat16 rounding that follows)
  float isx[9];
#pragma unroll
  for (int f = 0; f < 9; ++f) isx[f] = 1.0f / nm.sx[f];
  auto fetch = [&](long long gi, double (&raw)[9]) {
    const long long p = gi * 64 + lane;
    const long long pp = p < n ? p : n - 1;
#pragma unroll
    for (int f = 0; f < 9; ++f) raw[f] = g.p[f][pp];
  };
  double cur[9], nxt[9];
  if (first < groups) fetch(first, cur);
  for (long long gi = first; gi < groups; gi += stride) {
    if (gi + stride < groups) fetch(gi + stride, nxt);
    bf8 f0, f1;
#pragma unroll
    for (int f = 0; f < 8; ++f) f0[f] = (__bf16)((float)(cur[f] - (double)nm.mx[f]) * isx[f]);
    f1[0] = (__bf16)((float)(cur[8] - (double)nm.mx[8]) * isx[8]);
    f1[1] = (__bf16)dn;
#pragma unroll
    for (int f = 2; f < 8; ++f) f1[f] = (__bf16)0.0f;
    S.feat[w][lane][0] = f0;
    S.feat[w][lane][1] = f1;
    wave_sync();
#pragma unroll 1
    for (int t = 0; t < 4; ++t) {
      // an index the compiler cannot prove invariant keeps the weight reads
      // in the loop (hoisted, they would hold ~90 registers per lane)
      int l = lane;
      asm volatile("" : "+v"(l));
      bf8 xb;
      if (q < 2) {
        xb = S.feat[w][16 * t + c][q];
      } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) xb[r] = (__bf16)0.0f;
      }
      f4 z[4], h[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) z[b] = mma32(W.w1[b][l], xb, W.b1[b][l]);
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) h[b][r] = gelu(z[b][r]);
      bf8 h0 = pack2(h[0], h[1]), h1 = pack2(h[2], h[3]);
#pragma unroll
      for (int b = 0; b < 4; ++b) z[b] = mma32(W.w2[b][1][l], h1, mma32(W.w2[b][0][l], h0, W.b2[b][l]));
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) h[b][r] = gelu(z[b][r]);
      h0 = pack2(h[0], h[1]);
      h1 = pack2(h[2], h[3]);
      const f4 y = mma32(W.w3[1][l], h1, mma32(W.w3[0][l], h0, W.b3[l]));
      if (q < 2) S.out[w][16 * t + c][q] = y;
    }
    wave_sync();
    const f4 o0 = S.out[w][lane][0], o1 = S.out[w][lane][1];
    const long long p = gi * 64 + lane;
    if (p < n) {
#pragma unroll
      for (int j = 0; j < NO; ++j) {
        const float v = j < 4 ? o0[j] : o1[j - 4];
        tau.p[j][p] = (double)v * (double)nm.sy[j] + (double)nm.my[j];
      }
    }
#pragma unroll
    for (int f = 0; f < 9; ++f) cur[f] = nxt[f];
    wave_sync();  // this trip's LDS reads are done before the next trip's writes
  }
}

// ---- dense forward on normalised arrays (x: n x NI, y: n x NO) -------------
__global__ __launch_bounds__(BLOCK) void k_mlp_forward(const float *__restrict__ X, const float *__restrict__ P,
                                                       float *__restrict__ Y, long long n) {
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4;
  FwdW W;
  load_fwd(W, P, l);
  const long long wave = (long long)blockIdx.x * WAVES + (threadIdx.x >> 6), nw = (long long)gridDim.x * WAVES;
  const long long tiles = (n + 15) / 16;
  for (long long t = wave; t < tiles; t += nw) {
    const long long p = t * 16 + c;
    const bool ok = p < n;
    bf8 xb;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int f = 8 * q + r;
      xb[r] = (__bf16)(ok && f < NI ? X[p * NI + f] : 0.0f);
    }
    Acts a;
    forward(W, xb, a);
    if (ok) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = 4 * q + r;
        if (j < NO) Y[p * NO + j] = a.y[r];
      }
    }
  }
}

// ---- training: forward, backward and per-wave gradient partials ------------
// Point i of the batch is sample (mul * (base + i) + add) mod n of the data
// set, a permutation when mul is coprime to n.  loss = mean over the batch
// and the NO outputs of (y - t)^2.  partial[w * (NPAR + 1) + k]: gradient
// element k summed over the tiles of wave w; element NPAR is the loss.
struct Perm {
  unsigned long long mul, add, base;
};

__global__ __launch_bounds__(BLOCK) void k_mlp_train(const float *__restrict__ X, const float *__restrict__ T,
                                                     long long n, long long batch, Perm pm,
                                                     const float *__restrict__ P, float *__restrict__ partial) {
  __shared__ __bf16 sh[WAVES][4 * 16 * NH + 2 * 16 * 16];
  const int l = threadIdx.x & 63, c = l & 15, q = l >> 4, wv = threadIdx.x >> 6;
  __bf16 *Tdz2 = sh[wv], *Th1 = Tdz2 + 16 * NH, *Tdz1 = Th1 + 16 * NH, *Th2 = Tdz1 + 16 * NH;
  __bf16 *Tx = Th2 + 16 * NH, *Tdy = Tx + 256;
  FwdW W;
  BwdW B;
  load_fwd(W, P, l);
  load_bwd(B, P, l);
  f4 gW2[4][4], gW3[4], gW1[4];
  f4 gb1[4], gb2[4], gb3;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    gW3[i] = f4{0, 0, 0, 0};
    gW1[i] = f4{0, 0, 0, 0};
    gb1[i] = f4{0, 0, 0, 0};
    gb2[i] = f4{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j) gW2[i][j] = f4{0, 0, 0, 0};
  }
  gb3 = f4{0, 0, 0, 0};
  float loss = 0.0f;
  const float dscale = 2.0f / ((float)batch * (float)NO);
  const long long wave = (long long)blockIdx.x * WAVES + wv, nw = (long long)gridDim.x * WAVES;
  const long long tiles = (batch + 15) / 16;
  for (long long t = wave; t < tiles; t += nw) {
    const long long i = t * 16 + c;
    const bool ok = i < batch;
    const long long p = ok ? (long long)((pm.mul * (unsigned long long)(pm.base + i) + pm.add) % (unsigned long long)n) : 0;
    bf8 xb;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int f = 8 * q + r;
      xb[r] = (__bf16)(ok && f < NI ? X[p * NI + f] : 0.0f);
    }
    Acts a;
    forward(W, xb, a);
    f4 dy;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = 4 * q + r;
      const bool v = ok && j < NO;
      const float e = v ? a.y[r] - T[p * NO + j] : 0.0f;
      loss += e * e;
      dy[r] = dscale * e;
    }
    bf8 dyb;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      dyb[r] = (__bf16)dy[r];
      dyb[4 + r] = (__bf16)0.0f;
    }
    f4 dz2[4], dz1[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const f4 dh = mma32(B.w3t[b], dyb, f4{0, 0, 0, 0});
#pragma unroll
      for (int r = 0; r < 4; ++r) dz2[b][r] = dh[r] * gelu_d(a.z2[b][r]);
    }
    const bf8 dz2b0 = pack2(dz2[0], dz2[1]), dz2b1 = pack2(dz2[2], dz2[3]);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const f4 dh = mma32(B.w2t[b][1], dz2b1, mma32(B.w2t[b][0], dz2b0, f4{0, 0, 0, 0}));
#pragma unroll
      for (int r = 0; r < 4; ++r) dz1[b][r] = dh[r] * gelu_d(a.z1[b][r]);
    }
    // bias gradients: sums over points, per lane now and across lanes at the end
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      gb1[b] += dz1[b];
      gb2[b] += dz2[b];
    }
    gb3 += dy;
    // weight gradients: transposes of this tile through LDS, point-major
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int u = 16 * b + 4 * q + r;
        Tdz2[c * NH + u] = (__bf16)dz2[b][r];
        Tdz1[c * NH + u] = (__bf16)dz1[b][r];
      }
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        Th1[c * NH + uslot(s, q, r)] = a.h1b[s][r];
        Th2[c * NH + uslot(s, q, r)] = a.h2b[s][r];
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) Tdy[c * 16 + 4 * q + r] = (__bf16)dy[r];
    // features of point c: lanes q = 0, 1 hold features 0..15 as xb[0..7]
    if (q < 2) {
#pragma unroll
      for (int r = 0; r < 8; ++r) Tx[c * 16 + 8 * q + r] = xb[r];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // A (rows m, k = points 4q + r) and B (k = points, columns n) operands
    auto opA = [&](const __bf16 *Tm, int stride, int col) {
      s4 o;
#pragma unroll
      for (int r = 0; r < 4; ++r) o[r] = __builtin_bit_cast(short, Tm[(4 * q + r) * stride + col]);
      return o;
    };
    s4 h1B[4], xB;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) h1B[nb] = opA(Th1, NH, 16 * nb + c);
    xB = opA(Tx, 16, c);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      const s4 dzA = opA(Tdz2, NH, 16 * mb + c);
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) gW2[mb][nb] = mma16(dzA, h1B[nb], gW2[mb][nb]);
      gW1[mb] = mma16(opA(Tdz1, NH, 16 * mb + c), xB, gW1[mb]);
    }
    const s4 dyA = opA(Tdy, 16, c);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) gW3[nb] = mma16(dyA, opA(Th2, NH, 16 * nb + c), gW3[nb]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  // reduce bias sums and the loss over the 16 points of each lane group
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        gb1[b][r] += __shfl_xor(gb1[b][r], off);
        gb2[b][r] += __shfl_xor(gb2[b][r], off);
      }
#pragma unroll
    for (int r = 0; r < 4; ++r) gb3[r] += __shfl_xor(gb3[r], off);
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) loss += __shfl_xor(loss, off);
  float *out = partial + (size_t)wave * (NPAR + 1);
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * mb + 4 * q + r;
      if (c < NI) out[OW1 + row * NI + c] = gW1[mb][r];
#pragma unroll
      for (int nb = 0; nb < 4; ++nb) out[OW2 + row * NH + 16 * nb + c] = gW2[mb][nb][r];
    }
#pragma unroll
  for (int nb = 0; nb < 4; ++nb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int j = 4 * q + r;
      if (j < NO) out[OW3 + j * NH + 16 * nb + c] = gW3[nb][r];
    }
  if (c == 0) {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        out[OB1 + 16 * b + 4 * q + r] = gb1[b][r];
        out[OB2 + 16 * b + 4 * q + r] = gb2[b][r];
      }
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (4 * q + r < NO) out[OB3 + 4 * q + r] = gb3[r];
  }
  if (l == 0) out[NPAR] = loss / ((float)batch * (float)NO);
}

// gradient = sum of the wave partials in wave order (deterministic)
__global__ void k_mlp_reduce(const float *__restrict__ partial, int nwaves, float *__restrict__ grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > NPAR) return;
  double s = 0.0;
  for (int w = 0; w < nwaves; ++w) s += (double)partial[(size_t)w * (NPAR + 1) + k];
  grad[k] = (float)s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters
__global__ void k_mlp_adam(float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m1,
                           float *__restrict__ m2, float lr, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NPAR) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gk = g[k];
  const float a = b1 * m1[k] + (1.0f - b1) * gk, v = b2 * m2[k] + (1.0f - b2) * gk * gk;
  m1[k] = a;
  m2[k] = v;
  P[k] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// features and targets of the data set: X[p] = normalised (grad9, delta),
// T[p] = normalised tau6 (skipped when tau is null)
__global__ void k_mlp_pack(Ptr9 g, double delta, Ptr6 tau, Norm nm, long long n, float *__restrict__ X,
                           float *__restrict__ T) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
#pragma unroll
  for (int f = 0; f < 9; ++f) X[p * NI + f] = (float)(g.p[f][p] - (double)nm.mx[f]) / nm.sx[f];
  X[p * NI + 9] = (float)(delta - (double)nm.mx[9]) / nm.sx[9];
  if (T && tau.p[0]) {
#pragma unroll
    for (int j = 0; j < NO; ++j) T[p * NO + j] = (float)(tau.p[j][p] - (double)nm.my[j]) / nm.sy[j];
  }
}

// per-block sums of x and x^2 for the 15 field columns (9 inputs + 6 outputs)
__global__ void k_mlp_moments(Ptr9 g, Ptr6 tau, long long n, double *__restrict__ part) {
  __shared__ double sh[2 * 15][256 / 64];
  double s[15], s2[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) s[k] = s2[k] = 0.0;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      const double v = k < 9 ? g.p[k][p] : tau.p[k - 9][p];
      s[k] += v;
      s2[k] += v * v;
    }
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    for (int off = 32; off > 0; off >>= 1) {
      s[k] += __shfl_xor(s[k], off);
      s2[k] += __shfl_xor(s2[k], off);
    }
    if (lane == 0) {
      sh[k][w] = s[k];
      sh[15 + k][w] = s2[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 30) {
    double t = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += sh[threadIdx.x][i];
    part[(size_t)blockIdx.x * 30 + threadIdx.x] = t;
  }
}

int fail(int code, const std::string &m) { return ghx_internal_set_err(code, m.c_str()); }
#define MHIP(expr)                                                                                      \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess)                                                                               \
      return fail(e_ == hipErrorOutOfMemory ? GHX_ERR_NOMEM : GHX_ERR_HIP, std::string(#expr) + ": " + \
                                                                                hipGetErrorString(e_));  \
  } while (0)

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t next_prime_above(uint64_t n) {
  auto prime = [](uint64_t v) {
    if (v < 2) return false;
    for (uint64_t d = 2; d * d <= v; ++d)
      if (v % d == 0) return false;
    return true;
  };
  uint64_t v = n + 1;
  while (!prime(v)) ++v;
  return v;
}

}  // namespace

struct ghx_mlp {
  float *P = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *partial = nullptr;
  double *mom = nullptr;
  Norm norm;
  long long step = 0;
  hipStream_t stream = nullptr;
};

// inference grids: enough waves to hide the feature loads (several per SIMD),
// each wave taking at least two tiles so its weight loads are amortised
static unsigned infer_blocks(long long n) {
  const long long tiles = (n + 15) / 16, b = (tiles + 2 * WAVES - 1) / (2 * WAVES);
  return (unsigned)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}
// k_mlp_fields: one 64-point group per wave and trip, two trips per wave at
// least (the weight copy into LDS is amortised), at most 8 blocks per CU's worth
static unsigned field_blocks(long long n) {
  const long long groups = (n + 63) / 64, b = (groups + 2 * IW - 1) / (2 * IW);
  return (unsigned)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

static int check_ptr(const void *p, const char *what) {
  return p ? GHX_OK : fail(GHX_ERR_ARG, std::string("mlp: null ") + what);
}

extern "C" int ghx_mlp_nparams(void) { return NPAR; }

extern "C" int ghx_mlp_create(ghx_mlp **mp, unsigned long long seed) {
  if (!mp) return fail(GHX_ERR_ARG, "mlp_create: null handle");
  ghx_mlp *m = new (std::nothrow) ghx_mlp;
  if (!m) return fail(GHX_ERR_NOMEM, "mlp_create: host allocation");
  for (int i = 0; i < NI; ++i) { m->norm.mx[i] = 0.0f; m->norm.sx[i] = 1.0f; }
  for (int j = 0; j < NO; ++j) { m->norm.my[j] = 0.0f; m->norm.sy[j] = 1.0f; }
  // Glorot-uniform weights, zero biases
  std::vector<float> h(NPAR, 0.0f);
  uint64_t s = seed;
  auto fill = [&](int off, int rows, int cols) {
    const double a = std::sqrt(6.0 / (double)(rows + cols));
    for (int k = 0; k < rows * cols; ++k) h[off + k] = (float)(a * (2.0 * ((splitmix(s) >> 11) * 0x1.0p-53) - 1.0));
  };
  fill(OW1, NH, NI);
  fill(OW2, NH, NH);
  fill(OW3, NO, NH);
  hipError_t e = hipMalloc(&m->P, 4 * (NPAR + 1) * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->partial, (size_t)NWAVES * (NPAR + 1) * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&m->mom, (size_t)GRID * 30 * sizeof(double));
  if (e != hipSuccess) {
    if (m->P) (void)hipFree(m->P);
    if (m->partial) (void)hipFree(m->partial);
    delete m;
    return fail(GHX_ERR_NOMEM, "mlp_create: device allocation");
  }
  m->m1 = m->P + (NPAR + 1);
  m->m2 = m->m1 + (NPAR + 1);
  m->grad = m->m2 + (NPAR + 1);
  MHIP(hipMemset(m->P, 0, 4 * (NPAR + 1) * sizeof(float)));
  MHIP(hipMemcpy(m->P, h.data(), NPAR * sizeof(float), hipMemcpyHostToDevice));
  *mp = m;
  return GHX_OK;
}

extern "C" int ghx_mlp_destroy(ghx_mlp *m) {
  if (!m) return GHX_OK;
  (void)hipFree(m->P);
  (void)hipFree(m->partial);
  (void)hipFree(m->mom);
  delete m;
  return GHX_OK;
}

extern "C" int ghx_mlp_set_stream(ghx_mlp *m, void *stream) {
  if (!m) return fail(GHX_ERR_ARG, "mlp: null handle");
  m->stream = (hipStream_t)stream;
  return GHX_OK;
}

extern "C" int ghx_mlp_params(ghx_mlp *m, float **params) {
  if (!m || !params) return fail(GHX_ERR_ARG, "mlp_params: null argument");
  *params = m->P;
  return GHX_OK;
}

extern "C" int ghx_mlp_set_norm(ghx_mlp *m, const float *norm) {
  if (!m || !norm) return fail(GHX_ERR_ARG, "mlp_set_norm: null argument");
  std::memcpy(m->norm.mx, norm, NI * sizeof(float));
  std::memcpy(m->norm.sx, norm + NI, NI * sizeof(float));
  std::memcpy(m->norm.my, norm + 2 * NI, NO * sizeof(float));
  std::memcpy(m->norm.sy, norm + 2 * NI + NO, NO * sizeof(float));
  for (int i = 0; i < NI; ++i)
    if (!(m->norm.sx[i] > 0.0f)) return fail(GHX_ERR_ARG, "mlp_set_norm: input scales must be positive");
  for (int j = 0; j < NO; ++j)
    if (!(m->norm.sy[j] > 0.0f)) return fail(GHX_ERR_ARG, "mlp_set_norm: output scales must be positive");
  return GHX_OK;
}

extern "C" int ghx_mlp_get_norm(ghx_mlp *m, float *norm) {
  if (!m || !norm) return fail(GHX_ERR_ARG, "mlp_get_norm: null argument");
  std::memcpy(norm, m->norm.mx, NI * sizeof(float));
  std::memcpy(norm + NI, m->norm.sx, NI * sizeof(float));
  std::memcpy(norm + 2 * NI, m->norm.my, NO * sizeof(float));
  std::memcpy(norm + 2 * NI + NO, m->norm.sy, NO * sizeof(float));
  return GHX_OK;
}

static Ptr9 ptr9(const double *const *g) {
  Ptr9 p;
  for (int i = 0; i < 9; ++i) p.p[i] = g[i];
  return p;
}
static Ptr6 ptr6(double *const *t) {
  Ptr6 p;
  for (int i = 0; i < 6; ++i) p.p[i] = t ? t[i] : nullptr;
  return p;
}

extern "C" int ghx_mlp_fit_norm(ghx_mlp *m, long long n, const double *const *grad9, double delta,
                                const double *const *tau6) {
  if (!m || !grad9 || !tau6 || n < 2) return fail(GHX_ERR_ARG, "mlp_fit_norm: bad arguments");
  for (int i = 0; i < 9; ++i) if (!grad9[i]) return fail(GHX_ERR_ARG, "mlp_fit_norm: null gradient field");
  for (int j = 0; j < 6; ++j) if (!tau6[j]) return fail(GHX_ERR_ARG, "mlp_fit_norm: null stress field");
  hipLaunchKernelGGL(k_mlp_moments, dim3(GRID), dim3(256), 0, m->stream, ptr9(grad9),
                     ptr6(const_cast<double *const *>(tau6)), n, m->mom);
  MHIP(hipGetLastError());
  std::vector<double> h((size_t)GRID * 30);
  MHIP(hipMemcpyAsync(h.data(), m->mom, h.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  MHIP(hipStreamSynchronize(m->stream));
  for (int k = 0; k < 15; ++k) {
    double s = 0, s2 = 0;
    for (int b = 0; b < GRID; ++b) {
      s += h[(size_t)b * 30 + k];
      s2 += h[(size_t)b * 30 + 15 + k];
    }
    const double mu = s / (double)n, var = s2 / (double)n - mu * mu;
    const double sd = var > 0 ? std::sqrt(var) : 1.0;
    if (k < 9) { m->norm.mx[k] = (float)mu; m->norm.sx[k] = (float)sd; }
    else { m->norm.my[k - 9] = (float)mu; m->norm.sy[k - 9] = (float)sd; }
  }
  m->norm.mx[9] = (float)delta;  // one filter width: the input is centred to zero
  m->norm.sx[9] = 1.0f;
  return GHX_OK;
}

extern "C" int ghx_mlp_pack(ghx_mlp *m, long long n, const double *const *grad9, double delta,
                            const double *const *tau6, float *X, float *T) {
  if (!m || !grad9 || !X || n < 1) return fail(GHX_ERR_ARG, "mlp_pack: bad arguments");
  for (int i = 0; i < 9; ++i) if (!grad9[i]) return fail(GHX_ERR_ARG, "mlp_pack: null gradient field");
  if (T && !tau6) return fail(GHX_ERR_ARG, "mlp_pack: targets need the stress fields");
  hipLaunchKernelGGL(k_mlp_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, m->stream, ptr9(grad9), delta,
                     ptr6(const_cast<double *const *>(tau6)), m->norm, n, X, T);
  MHIP(hipGetLastError());
  return GHX_OK;
}

extern "C" int ghx_mlp_forward(ghx_mlp *m, long long n, const float *X, float *Y) {
  if (!m || !X || !Y || n < 0) return fail(GHX_ERR_ARG, "mlp_forward: bad arguments");
  if (n == 0) return GHX_OK;
  hipLaunchKernelGGL(k_mlp_forward, dim3(infer_blocks(n)), dim3(BLOCK), 0, m->stream, X, m->P, Y, n);
  MHIP(hipGetLastError());
  return GHX_OK;
}

// gradient of the batch loss into m->grad (element NPAR = loss)
static int grad_batch(ghx_mlp *m, long long n, const float *X, const float *T, long long batch, Perm pm) {
  hipLaunchKernelGGL(k_mlp_train, dim3(GRID), dim3(BLOCK), 0, m->stream, X, T, n, batch, pm, m->P, m->partial);
  MHIP(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((NPAR + 1 + 255) / 256), dim3(256), 0, m->stream, m->partial, NWAVES,
                     m->grad);
  MHIP(hipGetLastError());
  return GHX_OK;
}

extern "C" int ghx_mlp_grad(ghx_mlp *m, long long n, const float *X, const float *T, float *grad, double *loss) {
  if (!m || !X || !T || n < 1) return fail(GHX_ERR_ARG, "mlp_grad: bad arguments");
  Perm pm{1, 0, 0};
  int rc = grad_batch(m, n, X, T, n, pm);
  if (rc) return rc;
  float l = 0.0f;
  if (grad) MHIP(hipMemcpyAsync(grad, m->grad, NPAR * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
  MHIP(hipMemcpyAsync(&l, m->grad + NPAR, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MHIP(hipStreamSynchronize(m->stream));
  if (loss) *loss = l;
  return GHX_OK;
}

extern "C" int ghx_mlp_train(ghx_mlp *m, long long n, const float *X, const float *T, long long batch, int steps,
                             float lr, unsigned long long seed, float *losses) {
  if (!m || !X || !T || n < 1 || batch < 1 || steps < 0 || !(lr > 0.0f)) return fail(GHX_ERR_ARG, "mlp_train: bad arguments");
  if (batch > n) batch = n;
  uint64_t s = seed;
  // a prime multiplier between n/2 and n never divides n, so i -> mul i + add
  // (mod n) is a permutation; near the golden section of n it scatters
  // consecutive batch entries over the whole data set
  uint64_t mul = n > 2 ? next_prime_above((uint64_t)(0.6180339887498949 * (double)n)) : 1;
  if (mul >= (uint64_t)n) mul = 1;
  std::vector<float> lh(steps > 0 ? steps : 1);
  unsigned long long base = 0;
  const unsigned long long add = splitmix(s) % (uint64_t)n;
  for (int it = 0; it < steps; ++it) {
    Perm pm{mul, add, base};
    int rc = grad_batch(m, n, X, T, batch, pm);
    if (rc) return rc;
    ++m->step;
    const float c1 = 1.0f - std::pow(0.9f, (float)m->step), c2 = 1.0f - std::pow(0.999f, (float)m->step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPAR + 255) / 256), dim3(256), 0, m->stream, m->P, m->grad, m->m1, m->m2,
                       lr, c1, c2);
    MHIP(hipGetLastError());
    if (losses) MHIP(hipMemcpyAsync(&lh[it], m->grad + NPAR, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    base = (base + (unsigned long long)batch) % (unsigned long long)n;
  }
  MHIP(hipStreamSynchronize(m->stream));
  if (losses) std::memcpy(losses, lh.data(), (size_t)steps * sizeof(float));
  return GHX_OK;
}

extern "C" int ghx_mlp_stress(ghx_mlp *m, long long n, const double *const *grad9, double delta, double *const *tau6) {
  if (!m || !grad9 || !tau6 || n < 0) return fail(GHX_ERR_ARG, "mlp_stress: bad arguments");
  for (int i = 0; i < 9; ++i) if (!grad9[i]) return fail(GHX_ERR_ARG, "mlp_stress: null gradient field");
  for (int j = 0; j < 6; ++j) if (!tau6[j]) return fail(GHX_ERR_ARG, "mlp_stress: null stress field");
  if (n == 0) return GHX_OK;
  hipLaunchKernelGGL(k_mlp_fields, dim3(field_blocks(n)), dim3(IBLOCK), 0, m->stream, ptr9(grad9), delta, m->P,
                     m->norm, ptr6(tau6), n);
  MHIP(hipGetLastError());
  return GHX_OK;
}
