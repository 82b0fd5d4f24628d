// fft_passes.hip -- FFT pass kernels for GHOST's fftp layouts on gfx950.
//
// Every pass streams whole lines through registers (LineFFT) and uses LDS only
// for the Stockham exchanges and for the tile transposes that connect the
// x-contiguous real layout, the ky-contiguous intermediate Q and the
// kz-contiguous spectral layout S.  The transposes are fused into the first
// (load) or last (store) step of the x and z passes; global accesses are
// either whole contiguous lines or L-point chunks along ky (L*16 bytes).
#include "fft_engine.h"
#include "fft_passes.h"

// The launchers are split into parts (GHX_PART=1..8) so the build compiles
// them in parallel; each part instantiates only the kernels it launches.
#ifndef GHX_PART
#error "compile with -DGHX_PART=1..8"
#endif

namespace ghx {

// Sizes with a compiled schedule (complex line lengths).
#define GHX_SIZES(X) \
  X(2) X(3) X(4) X(5) X(6) X(8) X(10) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(1024)

#if GHX_PART == 1
bool fft_size_supported(int n) {
  switch (n) {
#define GHX_CASE(n_) case n_:
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
    return true;
  default:
    return false;
  }
}

#endif

static size_t g_max_lds = 0;
static size_t max_lds() {
  if (!g_max_lds) {
    int dev = 0, v = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = 65536;
    g_max_lds = (size_t)v;
  }
  return g_max_lds;
}

// ---------------------------------------------------------------------------
// Launch shapes.  The z and y passes run tiles of L lines x T lanes
// (__launch_bounds__(kTileBound)); the x passes run blocks of whole rows
// (__launch_bounds__(kRowBound)).  Both shapes are computed here, for the
// launchers and for ghx_pass_shape, and every launch is checked against them
// before it is issued (launch_check).
// ---------------------------------------------------------------------------
constexpr unsigned kTileBound = 512;
constexpr unsigned kRowBound = 256;
constexpr int kMaxTile = 32;  // the tile passes keep one bit per line in a 32-bit mask

struct PassShape {
  int lines;       // tile width L, or rows per block
  unsigned block;  // lanes per block
  size_t lds;      // dynamic LDS bytes
};

// Tile width: a power of two no wider than `want`, kMaxTile and the lines
// there are, with L * T <= kTileBound and L rows of `rowslots` points in LDS.
static int pick_L(int T, int rowslots, int nlines, int want, size_t ldsmax) {
  int L = 1;
  while (L * 2 <= want && L * 2 <= kMaxTile) L *= 2;
  while (L > 1 && ((size_t)L * rowslots * sizeof(cplx) > ldsmax || L * T > (int)kTileBound)) L >>= 1;
  if (L > nlines) {
    L = 1;
    while (L * 2 <= nlines) L *= 2;
  }
  return L < 1 ? 1 : L;
}
static PassShape tile_shape(int T, int rowslots, int nlines, int want, size_t ldsmax) {
  const int L = pick_L(T, rowslots, nlines, want, ldsmax);
  return PassShape{L, (unsigned)(L * T), (size_t)L * rowslots * sizeof(cplx)};
}
// The wide y inverse: 8-line tiles (the kernel stores 128-byte chunks); the
// launcher takes it only where their LDS fits, else the tile pass.
static PassShape wide_shape(int TW, int rowslots, int nlines) {
  (void)nlines;
  return PassShape{8, (unsigned)(8 * TW), (size_t)8 * rowslots * sizeof(cplx)};
}
// Rows per block of an x pass: kRowBound lanes, fewer rows when their LDS
// lines would not fit (the kernels read the row count from blockDim.x / T).
static PassShape row_shape(int T, int rowslots, size_t ldsmax) {
  int r = (int)kRowBound / T;
  if (r < 1) r = 1;
  while (r > 1 && (size_t)r * rowslots * sizeof(cplx) > ldsmax) r >>= 1;
  return PassShape{r, (unsigned)(r * T), (size_t)r * rowslots * sizeof(cplx)};
}

// Refuses, before it is issued, a launch that its kernel was not built for:
// more lanes than its __launch_bounds__ or than the code object's maximum
// work-group size, a block that is not whole lines of T lanes, a tile wider
// than the line masks, or more LDS than the device gives one block.
static hipError_t launch_check(const void *kern, const char *name, const PassShape &s, unsigned bound, int T,
                               bool tile) {
  const char *why = nullptr;
  if (s.block < 1 || s.block > bound) why = "more lanes than the kernel's launch bounds";
  else if (T < 1 || s.block % (unsigned)T) why = "a block that is not whole lines";
  else if (tile && (s.lines < 1 || s.lines > kMaxTile || (s.lines & (s.lines - 1)))) why = "a tile width that is not a power of two up to 32";
  else if (s.lds > max_lds()) why = "more LDS than one block may hold";
  if (!why) {
    hipFuncAttributes a;
    if (hipFuncGetAttributes(&a, kern) == hipSuccess && a.maxThreadsPerBlock > 0 &&
        s.block > (unsigned)a.maxThreadsPerBlock)
      why = "more lanes than the code object allows";
  }
  if (!why) return hipSuccess;
  fprintf(stderr, "ghx: refused launch of %s: %s (block %u, %d lines, T %d, LDS %zu bytes, bound %u)\n", name, why,
          s.block, s.lines, T, s.lds, bound);
  return hipErrorInvalidConfiguration;
}
#define GHX_LAUNCH_CHECK(kern, s, bound, T, tile)                                              \
  do {                                                                                         \
    const hipError_t ce_ = launch_check((const void *)(kern), #kern, (s), (bound), (T), (tile)); \
    if (ce_ != hipSuccess) return ce_;                                                         \
  } while (0)

#define GHX_X2_SIZES(X) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)
static bool x2_size(int nx) {
  switch (nx) {
#define GHX_CASE(n_) case n_:
    GHX_X2_SIZES(GHX_CASE)
#undef GHX_CASE
    return true;
  default:
    return false;
  }
}

#if GHX_PART == 1
// The launch shape a pass of line length n takes (kind 0: z and y tile
// passes; 1: the Boussinesq z forward, whose LDS rows hold n more points; 2:
// the wide y inverse; 3: two-for-one x passes, n = nx; 4: half-length x
// passes, n = nx/2), for `nlines` lines along the tiled axis, a requested
// tile width `want` (as the GHX_LZ/GHX_LY knobs and the autotuner give it)
// and a device LDS limit; then the check the launchers apply.  Host-only.
extern "C" int ghx_pass_shape(int kind, int n, int nlines, int want, long long ldsmax, int *lines, int *block,
                              long long *lds, int *lanes) {
  int T = 0, rows = 0;
  switch (n) {
#define GHX_CASE(n_)                                                                   \
  case n_:                                                                             \
    if (kind == 0 || kind == 1) T = FFTCfg<n_>::T, rows = lds_line_stride(n_) + (kind == 1 ? n_ : 0); \
    if (kind == 2 && FFTCfgWide<n_>::E != FFTCfg<n_>::E) T = FFTCfgWide<n_>::T, rows = lds_line_stride(n_); \
    if (kind == 3 && x2_size(n_)) T = FFTCfg<n_>::T, rows = lds_line_stride(n_);      \
    if (kind == 4) T = FFTCfgLow<n_>::T, rows = lds_line_stride(n_);                 \
    break;
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    break;
  }
  if (T <= 0 || nlines < 1 || ldsmax <= 0) return -1;
  PassShape s;
  unsigned bound = kTileBound;
  if (kind == 0 || kind == 1) s = tile_shape(T, rows, nlines, want, (size_t)ldsmax);
  else if (kind == 2) s = wide_shape(T, rows, nlines);
  else s = row_shape(T, rows, (size_t)ldsmax), bound = kRowBound;
  if (lines) *lines = s.lines;
  if (block) *block = (int)s.block;
  if (lds) *lds = (long long)s.lds;
  if (lanes) *lanes = T;
  const bool tile = kind <= 2;
  if (s.block < 1 || s.block > bound || s.block % (unsigned)T || s.lds > (size_t)ldsmax) return -2;
  if (tile && (s.lines > kMaxTile || (s.lines & (s.lines - 1)))) return -2;
  return 0;
}
#endif

// Tile widths (lines per block along ky) for the z and x passes; overridable
// with GHX_LZ / GHX_LX for tuning sweeps.
static int tile_want(const char *env, int dflt) {
  const char *e = getenv(env);
  int v = e ? atoi(e) : 0;
  return v > 0 ? v : dflt;
}

// An on/off switch for measurements: its value when set (0 included), else dflt.
static int env_switch(const char *env, int dflt) {
  const char *e = getenv(env);
  return e && *e ? atoi(e) : dflt;
}

// tile width: the environment override (tuning sweeps), else the context's
// choice, else the built-in default
static int tile_pick(const char *env, int ctx, int dflt) {
  const int e = tile_want(env, 0);
  return e > 0 ? e : (ctx > 0 ? ctx : dflt);
}

// Row lengths with a two-for-one x kernel (a full-length schedule of 6 or 8
// points per lane and at most 64 lanes; ghx_pass_shape takes the same list);
// GHX_XC2=0 selects the half-length kernels instead, for comparisons.
static bool use_x2(int nx) {
  if (env_switch("GHX_XC2", 1) != 1) return false;
  return x2_size(nx);
}

template <typename K> static void allow_lds(K kern, size_t bytes) {
  if (bytes > 49152) (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// ---------------------------------------------------------------------------
// XCD-aware tile order.  Workgroups are dealt round-robin to the 8 XCDs by
// linear id and each XCD has its own L2.  A tile narrower than a 128-byte
// line shares its lines with the neighbouring tile along the chunked axis, so
// within each group of 16 linear ids the two ids that land on the same XCD
// (b and b+8) are given neighbouring tiles.  Tiles stay spread over the XCDs
// in their natural order, which keeps pruned passes load-balanced.
// ---------------------------------------------------------------------------
// Member b of a batched launch: every field pointer moves by b * bstride.
// The view reads the pointers from the kernel arguments on each use and adds
// the offset; a shifted copy of the argument that is indexed by a run-time
// field number would be placed in scratch memory.
struct FieldRef {
  const FieldSet &fs;
  size_t off;
  GHX_DEV cplx *operator[](int f) const { return fs.p[f] + off; }
};
struct FieldView {
  FieldRef p;
};
GHX_DEV FieldView batch_view(const FieldSet &fs, unsigned b) { return FieldView{FieldRef{fs, (size_t)b * fs.bstride}}; }

GHX_DEV void xcd_tile(int &bx, int &by) {
  const int nb = gridDim.x * gridDim.y;
  int b = blockIdx.y * gridDim.x + blockIdx.x;
  if (b < (nb & ~15)) b = (b & ~15) | ((b & 7) << 1) | ((b >> 3) & 1);
  bx = b % gridDim.x;
  by = b / gridDim.x;
}

// ---------------------------------------------------------------------------
// Tile helpers (block = L lines x T lanes; line lt = threadIdx.x / T)
// ---------------------------------------------------------------------------

// Store the L lines of a block (kept in LDS rows) with the line index
// fastest: point p of line l goes to dst[p * pstride + l] when keep(l) (the
// dealias pruning and the tile's edge; the loader zero-fills the rest without
// touching memory).  The block is exactly L x T lanes with L a power of two
// and E * T points per row, so every lane moves E points: line
// l = idx & (L - 1), point idx >> log2 L.  The loop stays rolled (G points
// per trip): holding all E points at once costs most y and z kernels a wave
// per SIMD or spills; a caller with registers to spare sets G = E through GW
// (the 512-point curl z inverse, 8 points).  NT selects streaming
// stores: the y inverse (128-byte chunks) gains 10-14% with them, the z
// inverse (64-byte chunks at L = 4) loses 5-11%.
//
// The block is L x T lanes with L a power of two, so l is the same on every
// trip: the keep test is made once per lane, and the LDS reads go in groups
// of G (up to 4; GW > 0 sets it) ahead of their stores, so that a group's
// reads are in flight together instead of each store waiting on its own read.
template <int E, bool NT, int GW = 0, typename Keep>
GHX_DEV void tile_store_rows(const cplx *lds, int stride, int L, cplx *dst, size_t pstride, Keep keep) {
  constexpr int G = GW > 0 ? GW : E % 4 == 0 ? 4 : E % 3 == 0 ? 3 : E % 2 == 0 ? 2 : 1;
  static_assert(E % G == 0, "the group width divides the points per lane");
  const int lsh = 31 - __builtin_clz((unsigned)L);
  if constexpr (G == 1) {
    // one point per trip, the keep test in the loop: the fewest live
    // registers, for kernels at the register limit
#pragma unroll 2
    for (int q = 0; q < E; ++q) {
      const int idx = threadIdx.x + q * blockDim.x;
      const int l = idx & (L - 1);
      const cplx x = lds[l * stride + lds_pad(idx >> lsh)];
      if (keep(l)) c_put<NT>(dst + (size_t)(idx >> lsh) * pstride + l, x);
    }
    return;
  }
  const int l = threadIdx.x & (L - 1);
  if (!keep(l)) return;
  const int p0 = threadIdx.x >> lsh, dp = blockDim.x >> lsh;
  const cplx *src = lds + l * stride;
  cplx *d = dst + (size_t)p0 * pstride + l;
  const size_t dstep = (size_t)dp * pstride;
#pragma unroll 1
  for (int q0 = 0; q0 < E; q0 += G) {
    cplx x[G];
#pragma unroll
    for (int u = 0; u < G; ++u) x[u] = src[lds_pad(p0 + (q0 + u) * dp)];
#pragma unroll
    for (int u = 0; u < G; ++u) c_put<NT>(d + (size_t)(q0 + u) * dstep, x[u]);
  }
}
// Block = L lines x T lanes, E = N / T points per thread.  The E loads of a
// thread issue back to back (lines outside the mask or past ny read line j0,
// inside the tile, and are zeroed by a select), then go to LDS.
template <int E>
GHX_DEV void tile_load_q(cplx *lds, int stride, int L, const cplx *src, size_t pstride, size_t base, int j0, int ny,
                         unsigned lmask = ~0u) {
  cplx w[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int idx = threadIdx.x + q * blockDim.x;
    const int l = idx % L, p = idx / L;
    const bool inside = j0 + l < ny && ((lmask >> l) & 1u);
    const cplx x = src[base + (size_t)p * pstride + j0 + (inside ? l : 0)];
    w[q] = inside ? x : c_make(0.0, 0.0);
  }
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int idx = threadIdx.x + q * blockDim.x;
    lds[(idx % L) * stride + lds_pad(idx / L)] = w[q];
  }
}

// tile_load_q in two halves, so that a tile's loads can be in flight while
// the previous tile is transformed: the E loads of a thread into registers,
// then the registers into the LDS rows.
template <int E>
GHX_DEV void tile_fetch_q(cplx (&w)[E], int L, const cplx *src, size_t pstride, size_t base, int j0, int ny,
                          unsigned lmask) {
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int idx = threadIdx.x + q * blockDim.x;
    const int l = idx % L, p = idx / L;
    const bool inside = j0 + l < ny && ((lmask >> l) & 1u);
    const cplx x = src[base + (size_t)p * pstride + j0 + (inside ? l : 0)];
    w[q] = inside ? x : c_make(0.0, 0.0);
  }
}
template <int E>
GHX_DEV void tile_put_q(cplx *lds, int stride, int L, const cplx (&w)[E]) {
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int idx = threadIdx.x + q * blockDim.x;
    lds[(idx % L) * stride + lds_pad(idx / L)] = w[q];
  }
}

// A (kx, ky) line carries dealiased modes iff kn2(kx, ky, 0) <= kmax.
GHX_DEV bool line_active(const GridK &g, int i, int j) { return kn2_of(g, i, j, 0) <= g.kmax; }
GHX_DEV unsigned tile_lines_active(const GridK &g, int i, int j0, int L) {
  unsigned m = 0;
  for (int l = 0; l < L; ++l)
    if (j0 + l < g.ny && line_active(g, i, j0 + l)) m |= 1u << l;
  return m;
}

// ---------------------------------------------------------------------------
// Z^-1: S -> Q.  Block = (kx index i, ky chunk j0..j0+L-1).
// ---------------------------------------------------------------------------
// The store half of zinv_emit: the transformed line v (natural order, lane t
// holding points t + T m) into the block's LDS rows, then the tile to Q.
template <int N, int E, typename Cfg = FFTCfg<N>, int GW = 0>
GHX_DEV void zinv_store(const cplx (&v)[E], cplx *lds, cplx *row, int stride, int L, int t, cplx *dst,
                        const GridK &g, int i, int j0, unsigned lmask) {
  constexpr int T = LineFFT<N, +1, Cfg>::T;
#pragma unroll
  for (int m = 0; m < E; ++m) row[lds_pad(t + T * m)] = v[m];
  __syncthreads();
  const int ny = g.ny;
  tile_store_rows<E, false, GW>(lds, stride, L, dst + (size_t)i * ny + j0, (size_t)g.nxh * ny,
                     [=](int l) { return j0 + l < ny && ((lmask >> l) & 1u); });
  __syncthreads();
}

template <int N, int E, typename Cfg = FFTCfg<N>, int GW = 0>
GHX_DEV void zinv_emit(cplx (&v)[E], cplx *lds, cplx *row, int stride, int L, int t, const TwCache<N, Cfg> &tw,
                       cplx *dst, const GridK &g, int i, int j0, unsigned lmask) {
  LineFFT<N, +1, Cfg>::run(v, row, t, tw);
  zinv_store<N, E, Cfg, GW>(v, lds, row, stride, L, t, dst, g, i, j0, lmask);
}

// A vector field's z inverse and its curl's with five transforms instead of
// six.  kx and ky are constant along a z line, so Z^-1(i ky a) = i ky Z^-1(a);
// only the kz terms of rotor3 (pseudospec3D_hd.f90:124-200) need transforms of
// their own:
//   curl_x = i ky Z^-1(a_z) - Z^-1(i kz a_y)
//   curl_y = Z^-1(i kz a_x) - i kx Z^-1(a_z)
//   curl_z = i kx Z^-1(a_y) - i ky Z^-1(a_x)
// ax, ay, az (the spectral lines, lane t holding points t + T m) are
// transformed in place; the six results go to o[0..2] (a) and o[3..5] (curl a).
template <int N, int E, typename Cfg, int GW = 0>
GHX_DEV void zinv_curl5(cplx (&ax)[E], cplx (&ay)[E], cplx (&az)[E], double kx, double ky, cplx *lds, cplx *row,
                        int stride, int L, int t, const TwCache<N, Cfg> &tw, cplx *const *o, const GridK &g, int i,
                        int j0, unsigned lmask) {
  using F = LineFFT<N, +1, Cfg>;
  constexpr int T = F::T;
  zinv_emit<N, E, Cfg, GW>(az, lds, row, stride, L, t, tw, o[2], g, i, j0, lmask);  // az := Z^-1(a_z)
  {
    cplx w[E];
#pragma unroll
    for (int m = 0; m < E; ++m) w[m] = c_ik(wavenum(t + T * m, g.nz, g.dkz), ay[m]);
    F::run(w, row, t, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) w[m] = c_sub(c_ik(ky, az[m]), w[m]);
    zinv_store<N, E, Cfg, GW>(w, lds, row, stride, L, t, o[3], g, i, j0, lmask);
  }
  {
    cplx w[E];
#pragma unroll
    for (int m = 0; m < E; ++m) w[m] = c_ik(wavenum(t + T * m, g.nz, g.dkz), ax[m]);
    F::run(w, row, t, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) w[m] = c_sub(w[m], c_ik(kx, az[m]));
    zinv_store<N, E, Cfg, GW>(w, lds, row, stride, L, t, o[4], g, i, j0, lmask);
  }
  zinv_emit<N, E, Cfg, GW>(ax, lds, row, stride, L, t, tw, o[0], g, i, j0, lmask);  // ax := Z^-1(a_x)
  zinv_emit<N, E, Cfg, GW>(ay, lds, row, stride, L, t, tw, o[1], g, i, j0, lmask);  // ay := Z^-1(a_y)
  {
    cplx w[E];
#pragma unroll
    for (int m = 0; m < E; ++m) w[m] = c_sub(c_ik(kx, ay[m]), c_ik(ky, ax[m]));
    zinv_store<N, E, Cfg, GW>(w, lds, row, stride, L, t, o[5], g, i, j0, lmask);
  }
}

template <int N, int MODE, typename Cfg = FFTCfg<N>>
__global__ __launch_bounds__(kTileBound) void k_zinv(FieldSet in_, FieldSet out_, int nf, GridK g, const cplx *__restrict__ twt,
                                             int prune, const double *__restrict__ aux) {
  using F = LineFFT<N, +1, Cfg>;
  constexpr int T = F::T, E = F::E;
  // store groups: the MHD and Boussinesq passes (MODE 2, 5) are at the
  // register limit and spill more with grouped reads; they store one point
  // per trip.  The curl pass takes all 8 points of a lane in one group
  // (0.853 -> 0.812 ms at 512^3, same register count)
  constexpr int SG = MODE == 2 || MODE == 5 ? 1 : (MODE == 1 && E == 8) ? 8 : 0;
  extern __shared__ cplx lds[];
  const int L = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  TwCache<N, Cfg> tw;  // stage twiddle roots of this lane, loaded once
  tw.load(twt, t);
  const FieldView in = batch_view(in_, blockIdx.z);
  const FieldView out = batch_view(out_, blockIdx.z);
  int bx, by;
  xcd_tile(bx, by);
  const int i = by, j0 = bx * L, j = j0 + lt;
  // prune: the input is dealiased, so only lines inside the dealias disc
  // carry data; only their kept modes are read and only they are written.
  const unsigned lmask = prune ? tile_lines_active(g, i, j0, L) : ~0u;
  if (lmask == 0u) return;  // uniform over the block
  const bool ok = j < g.ny && ((lmask >> lt) & 1u);
  const int stride = lds_line_stride(N);
  cplx *row = lds + lt * stride;
  const size_t sline = ((size_t)i * g.ny + (ok ? j : 0)) * N;
  auto need = [&](int k) { return ok && (!prune || kn2_of(g, i, j, k) <= g.kmax); };
  // every lane loads from inside its line (skipped modes read the line's
  // first point, which the line's own loads fetch anyway) and a select zeroes
  // what is not needed, so the loads of a field issue back to back
  auto load = [&](const cplx *p, int k) {
    const bool nd = need(k);
    const cplx x = p[sline + (nd ? k : 0)];
    return nd ? x : c_make(0.0, 0.0);
  };
  if constexpr (MODE == 0 || MODE == 3) {
    // MODE 3: periodic cubic-spline prefilter.  aux = [sx (nx/2+1) | sy (ny) |
    // sz (nz)] and every mode is scaled by sx[kx] sy[ky] sz[kz] on load.
    double sxy = 1.0;
    if constexpr (MODE == 3) sxy = aux[i] * aux[g.nxh + (ok ? j : 0)];
    for (int f = 0; f < nf; ++f) {
      cplx v[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        v[m] = load(in.p[f], t + T * m);
        if constexpr (MODE == 3) v[m] = c_scale(v[m], sxy * aux[g.nxh + g.ny + t + T * m]);
      }
      zinv_emit<N, E, Cfg, SG>(v, lds, row, stride, L, t, tw, out.p[f], g, i, j0, lmask);
    }
  } else if constexpr (MODE == 1) {
    // v -> (v, curl v)
    const double kx = wavenum(i, g.nx, g.dkx), ky = wavenum(j, g.ny, g.dky);
    cplx vx[E], vy[E], vz[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      vx[m] = load(in.p[0], t + T * m);
      vy[m] = load(in.p[1], t + T * m);
      vz[m] = load(in.p[2], t + T * m);
    }
    cplx *const o[6] = {out.p[0], out.p[1], out.p[2], out.p[3], out.p[4], out.p[5]};
    zinv_curl5<N, E, Cfg, SG>(vx, vy, vz, kx, ky, lds, row, stride, L, t, tw, o, g, i, j0, lmask);
  } else if constexpr (MODE == 2) {
    // (v, A) -> (v, curl v, b = curl A, curl b)
    const double kx = wavenum(i, g.nx, g.dkx), ky = wavenum(j, g.ny, g.dky);
    {
      cplx vx[E], vy[E], vz[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        vx[m] = load(in.p[0], t + T * m);
        vy[m] = load(in.p[1], t + T * m);
        vz[m] = load(in.p[2], t + T * m);
      }
      cplx *const o[6] = {out.p[0], out.p[1], out.p[2], out.p[3], out.p[4], out.p[5]};
      zinv_curl5<N, E, Cfg, SG>(vx, vy, vz, kx, ky, lds, row, stride, L, t, tw, o, g, i, j0, lmask);
    }
    // b = curl A per mode (rotor3), then b and curl b
    cplx bx[E], by[E], bz[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double kz = wavenum(t + T * m, g.nz, g.dkz);
      const cplx ax = load(in.p[3], t + T * m), ay = load(in.p[4], t + T * m), az = load(in.p[5], t + T * m);
      bx[m] = c_sub(c_ik(ky, az), c_ik(kz, ay));
      by[m] = c_sub(c_ik(kz, ax), c_ik(kx, az));
      bz[m] = c_sub(c_ik(kx, ay), c_ik(ky, ax));
    }
    cplx *const o[6] = {out.p[6], out.p[7], out.p[8], out.p[9], out.p[10], out.p[11]};
    zinv_curl5<N, E, Cfg, SG>(bx, by, bz, kx, ky, lds, row, stride, L, t, tw, o, g, i, j0, lmask);
  } else if constexpr (MODE == 6) {
    // velocity gradients: v -> d v_i / d x_j for i, j = 1..3 (out.p[3 i + j],
    // i major), derivk3 (pseudospec3D_hd.f90:22-88) scaled by 1/(nx ny nz)
    const double kx = wavenum(i, g.nx, g.dkx), ky = wavenum(j, g.ny, g.dky);
    const double sc = 1.0 / ((double)g.nx * (double)g.ny * (double)g.nz);
    cplx u[3][E];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int m = 0; m < E; ++m) u[c][m] = load(in.p[c], t + T * m);
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int dir = 0; dir < 3; ++dir) {
        cplx v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const double kd = dir == 0 ? kx : (dir == 1 ? ky : wavenum(t + T * m, g.nz, g.dkz));
          v[m] = c_scale(c_ik(kd, u[c][m]), sc);
        }
        zinv_emit<N, E, Cfg, SG>(v, lds, row, stride, L, t, tw, out.p[3 * c + dir], g, i, j0, lmask);
      }
  } else {
    // MODE 4: (a, b, c, d) -> (a, b, c, grad d).  MODE 5: (v, th) -> (v, curl v, grad th).
    using F = LineFFT<N, +1, Cfg>;
    const double kx = wavenum(i, g.nx, g.dkx), ky = wavenum(j, g.ny, g.dky);
    cplx ux[E], uy[E], uz[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      ux[m] = load(in.p[0], t + T * m);
      uy[m] = load(in.p[1], t + T * m);
      uz[m] = load(in.p[2], t + T * m);
    }
    if constexpr (MODE == 5) {
      cplx *const o[6] = {out.p[0], out.p[1], out.p[2], out.p[3], out.p[4], out.p[5]};
      zinv_curl5<N, E, Cfg, SG>(ux, uy, uz, kx, ky, lds, row, stride, L, t, tw, o, g, i, j0, lmask);
    } else {
      zinv_emit<N, E, Cfg, SG>(ux, lds, row, stride, L, t, tw, out.p[0], g, i, j0, lmask);
      zinv_emit<N, E, Cfg, SG>(uy, lds, row, stride, L, t, tw, out.p[1], g, i, j0, lmask);
      zinv_emit<N, E, Cfg, SG>(uz, lds, row, stride, L, t, tw, out.p[2], g, i, j0, lmask);
    }
    // derivk3 (pseudospec3D_hd.f90:22-88) of the scalar along x, y, z:
    // Z^-1(i kx d) = i kx Z^-1(d) and likewise for ky, so two transforms
    constexpr int base = MODE == 4 ? 3 : 6;
    cplx d[E];
#pragma unroll
    for (int m = 0; m < E; ++m) d[m] = load(in.p[3], t + T * m);
    {
      cplx w[E];
#pragma unroll
      for (int m = 0; m < E; ++m) w[m] = c_ik(wavenum(t + T * m, g.nz, g.dkz), d[m]);
      zinv_emit<N, E, Cfg, SG>(w, lds, row, stride, L, t, tw, out.p[base + 2], g, i, j0, lmask);
    }
    F::run(d, row, t, tw);
#pragma unroll
    for (int dir = 0; dir < 2; ++dir) {
      cplx w[E];
      const double kd = dir == 0 ? kx : ky;
#pragma unroll
      for (int m = 0; m < E; ++m) w[m] = c_ik(kd, d[m]);
      zinv_store<N, E, Cfg, SG>(w, lds, row, stride, L, t, out.p[base + dir], g, i, j0, lmask);
    }
  }
}

// ---------------------------------------------------------------------------
// Z forward: Q -> S.  Block = (kx index i, ky chunk).  Transposed load.
// ---------------------------------------------------------------------------
template <int N, int E, typename Cfg = FFTCfg<N>>
GHX_DEV void zfwd_load(cplx (&v)[E], cplx *lds, cplx *row, int stride, int L, int t, const TwCache<N, Cfg> &tw,
                       const cplx *src, const GridK &g, int i, int j0, unsigned lmask = ~0u) {
  using F = LineFFT<N, -1, Cfg>;
  constexpr int T = F::T;
  tile_load_q<E>(lds, stride, L, src, (size_t)g.nxh * g.ny, (size_t)i * g.ny, j0, g.ny, lmask);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = row[lds_pad(t + T * m)];
  __syncthreads();
  F::run(v, row, t, tw);
}

// The transform of a difference of two Q fields, FFT(a - b) = FFT(a) - FFT(b):
// the lines are subtracted before one transform.
template <int N, int E>
GHX_DEV void zfwd_load_diff(cplx (&v)[E], cplx *lds, cplx *row, int stride, int L, int t, const TwCache<N> &tw,
                            const cplx *srca, const cplx *srcb, const GridK &g, int i, int j0, unsigned lmask) {
  using F = LineFFT<N, -1>;
  constexpr int T = F::T;
  tile_load_q<E>(lds, stride, L, srca, (size_t)g.nxh * g.ny, (size_t)i * g.ny, j0, g.ny, lmask);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = row[lds_pad(t + T * m)];
  __syncthreads();
  tile_load_q<E>(lds, stride, L, srcb, (size_t)g.nxh * g.ny, (size_t)i * g.ny, j0, g.ny, lmask);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = c_sub(v[m], row[lds_pad(t + T * m)]);
  __syncthreads();
  F::run(v, row, t, tw);
}

template <int N>
// prune: only the kept modes of the output are wanted (the caller reads no
// other), so only the dealias disc's lines are loaded (the y forward wrote no
// others) and only their kept modes are written; the rest of `out` keeps
// whatever it held.
__global__ __launch_bounds__(kTileBound) void k_zfwd(FieldSet in, FieldSet out, int nf, GridK g, const cplx *__restrict__ twt,
                                             int prune) {
  using F = LineFFT<N, -1>;
  constexpr int T = F::T, E = F::E;
  extern __shared__ cplx lds[];
  const int L = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  TwCache<N> tw;  // stage twiddle roots of this lane, loaded once
  tw.load(twt, t);
  int bx, by;
  xcd_tile(bx, by);
  const int i = by, j0 = bx * L, j = j0 + lt;
  const unsigned lmask = prune ? tile_lines_active(g, i, j0, L) : ~0u;
  if (lmask == 0u) return;  // uniform over the block
  const bool ok = j < g.ny && ((lmask >> lt) & 1u);
  const int stride = lds_line_stride(N);
  cplx *row = lds + lt * stride;
  const size_t sline = ((size_t)i * g.ny + (ok ? j : 0)) * N;
  for (int f = 0; f < nf; ++f) {
    cplx v[E];
    zfwd_load<N, E>(v, lds, row, stride, L, t, tw, in.p[f], g, i, j0, lmask);
    if (ok) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + T * m;
        if (!prune || kn2_of(g, i, j, k) <= g.kmax) out.p[f][sline + k] = v[m];
      }
    }
    __syncthreads();
  }
}

// Projection of pseudospec3D_hd.f90:399-509 (nonlhd3) for one mode.
GHX_DEV void project_hd(const cplx &a, const cplx &b, const cplx &c, double kx, double ky, double kz,
                        double kk2, int i, int j, int k, cplx &gx, cplx &gy, cplx &gz) {
  cplx dot = c_add(c_add(c_scale(a, kx), c_scale(b, ky)), c_scale(c, kz));
  gx = i == 0 ? c_make(-a.x, -a.y) : c_add(c_make(-a.x, -a.y), c_make(kx * dot.x / kk2, kx * dot.y / kk2));
  gy = j == 0 ? c_make(-b.x, -b.y) : c_add(c_make(-b.x, -b.y), c_make(ky * dot.x / kk2, ky * dot.y / kk2));
  gz = k == 0 ? c_make(-c.x, -c.y) : c_add(c_make(-c.x, -c.y), c_make(kz * dot.x / kk2, kz * dot.y / kk2));
}

// Coriolis force of include/roth_rkstep2.f90:6-14 added to the products
// (a, b) = FFT of (w x u)_x,y before the projection: a - 2 omegaz v_y,
// b + 2 omegaz v_x, with om2 = 2 omegaz.
GHX_DEV void coriolis(cplx &a, cplx &b, const cplx &vx, const cplx &vy, double om2) {
  a = c_make(a.x - om2 * vy.x, a.y - om2 * vy.y);
  b = c_make(b.x + om2 * vx.x, b.y + om2 * vx.y);
}

// RK update of include/hd_rkstep2.f90:12-32 for one component:
// v = C + dt*(nu*(-kk2 v) + N + f)*rmp
GHX_DEV cplx rk_update(cplx cc, cplx vv, cplx nl, cplx ff, double kk2, double nu, double dt, double rmp) {
  double lx = -kk2 * vv.x, ly = -kk2 * vv.y;
  double sx = nu * lx + nl.x + ff.x, sy = nu * ly + nl.y + ff.y;
  return c_make(cc.x + dt * sx * rmp, cc.y + dt * sy * rmp);
}

template <int N, typename Cfg = FFTCfg<N>, int PF = 0>
__global__ __launch_bounds__(kTileBound) void k_zfwd_rk_hd(FieldSet in_, GridK g, const cplx *__restrict__ twt, RKParams rk,
                                                   int prune) {
  using F = LineFFT<N, -1, Cfg>;
  constexpr int T = F::T, E = F::E;
  extern __shared__ cplx lds[];
  const int L = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  TwCache<N, Cfg> tw;  // stage twiddle roots of this lane, loaded once
  tw.load(twt, t);
  const FieldView in = batch_view(in_, blockIdx.z);
  if (blockIdx.z && rk.bstride) {
    const size_t o = (size_t)blockIdx.z * rk.bstride;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rk.v[c] += o;
      rk.c[c] += o;
      rk.out[c] += o;
      if (rk.f[c]) rk.f[c] += o;
    }
  }
  if (rk.nu_b) rk.nu = rk.nu_b[blockIdx.z];
  int bx, by;
  xcd_tile(bx, by);
  const int i = by, j0 = bx * L, j = j0 + lt;
  // prune: all state buffers are dealiased (zero outside the kept modes), so
  // only lines inside the dealias disc are transformed and only kept modes
  // are read and written.
  const unsigned lmask = prune ? tile_lines_active(g, i, j0, L) : ~0u;
  if (lmask == 0u) return;  // uniform over the block
  const bool ok = j < g.ny && ((lmask >> lt) & 1u);
  const int stride = lds_line_stride(N);
  cplx *row = lds + lt * stride;
  cplx P[3][E];
  if constexpr (PF) {
    // the next field's tile loads are in flight while this field is transformed
    const size_t qs = (size_t)g.nxh * g.ny, qb = (size_t)i * g.ny;
    cplx w[E];
    tile_fetch_q<E>(w, L, in.p[0], qs, qb, j0, g.ny, lmask);
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      tile_put_q<E>(lds, stride, L, w);
      __syncthreads();
#pragma unroll
      for (int m = 0; m < E; ++m) P[f][m] = row[lds_pad(t + T * m)];
      __syncthreads();
      if (f < 2) tile_fetch_q<E>(w, L, in.p[f + 1], qs, qb, j0, g.ny, lmask);
      F::run(P[f], row, t, tw);
      __syncthreads();
    }
  } else {
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      zfwd_load<N, E, Cfg>(P[f], lds, row, stride, L, t, tw, in.p[f], g, i, j0, lmask);
      __syncthreads();
    }
  }
  if (!ok) return;
  const double kx = wavenum(i, g.nx, g.dkx), ky = wavenum(j, g.ny, g.dky);
  const size_t sline = ((size_t)i * g.ny + j) * N;
  const bool forced = rk.fline && rk.fline[(size_t)i * g.ny + j];
  const cplx z = c_make(0.0, 0.0);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    const size_t s = sline + k;
    const double kz = wavenum(k, g.nz, g.dkz);
    const double kn2 = kn2_of(g, i, j, k);
    if (kn2 <= g.kmax && kn2 >= g.tiny) {
      const double kk2 = g.anis ? kk2_of(g, kx, ky, kz) : kn2;
      cplx n0, n1, n2;
      if (rk.om2 != 0.0) coriolis(P[0][m], P[1][m], rk.v[0][s], rk.v[1][s], rk.om2);
      project_hd(P[0][m], P[1][m], P[2][m], kx, ky, kz, kk2, i, j, k, n0, n1, n2);
      c_put<true>(&rk.out[0][s], rk_update(rk.c[0][s], rk.v[0][s], n0, forced ? rk.f[0][s] : z, kk2, rk.nu, rk.dt, rk.rmp));
      c_put<true>(&rk.out[1][s], rk_update(rk.c[1][s], rk.v[1][s], n1, forced ? rk.f[1][s] : z, kk2, rk.nu, rk.dt, rk.rmp));
      c_put<true>(&rk.out[2][s], rk_update(rk.c[2][s], rk.v[2][s], n2, forced ? rk.f[2][s] : z, kk2, rk.nu, rk.dt, rk.rmp));
    } else if (!prune || kn2 < g.tiny) {
      // outside the kept modes; the k = 0 mode is written in pruned passes
      // too, since the next substep's inverse reads it from this output
      c_put<true>(&rk.out[0][s], z);
      c_put<true>(&rk.out[1][s], z);
      c_put<true>(&rk.out[2][s], z);
    }
    // keep the loads of one mode from being hoisted over the others (register pressure)
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Boussinesq (include/bouss_rkstep2.f90:1-71): in.p[0..2] = FFT of w x u,
// in.p[3] = FFT of -(u . grad th).  Per kept mode: buoyancy xmom th added to
// the z product, nonlhd3 projection, heat current xtemp v_z added to the
// advection, then the updates of v (nu) and th (kappa).  The k = 0 mode keeps
// the step-start mean of th; v is zero there and outside the dealias set.
template <int N>
__global__ __launch_bounds__(kTileBound) void k_zfwd_rk_bouss(FieldSet in, GridK g, const cplx *__restrict__ twt,
                                                      BoussRKParams rk, int prune) {
  using F = LineFFT<N, -1>;
  constexpr int T = F::T, E = F::E;
  extern __shared__ cplx lds[];
  const int L = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  TwCache<N> tw;  // stage twiddle roots of this lane, loaded once
  tw.load(twt, t);
  int bx, by;
  xcd_tile(bx, by);
  const int i = by, j0 = bx * L, j = j0 + lt;
  const unsigned lmask = prune ? tile_lines_active(g, i, j0, L) : ~0u;
  if (lmask == 0u) return;  // uniform over the block
  const bool ok = j < g.ny && ((lmask >> lt) & 1u);
  const int stride = lds_line_stride(N);
  cplx *row = lds + lt * stride;
  const double kx = wavenum(i, g.nx, g.dkx), ky = wavenum(j < g.ny ? j : 0, g.ny, g.dky);
  const size_t sline = ((size_t)i * g.ny + (ok ? j : 0)) * N;
  const bool forced = ok && rk.fline && rk.fline[(size_t)i * g.ny + j];
  const cplx z = c_make(0.0, 0.0);
  // v_z as the nonlinear term saw it, for the heat current: with ord = 1 the
  // update writes over the input (out == v) before the temperature is done.
  // Kept in LDS past the line tiles (each lane rereads only its own modes).
  cplx *vz_in = lds + L * stride + lt * N;
  {
    cplx P[3][E];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      zfwd_load<N, E>(P[f], lds, row, stride, L, t, tw, in.p[f], g, i, j0, lmask);
      __syncthreads();
    }
    if (ok) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = t + T * m;
        const size_t s = sline + k;
        const double kz = wavenum(k, g.nz, g.dkz);
        const double kn2 = kn2_of(g, i, j, k);
        if (kn2 <= g.kmax && kn2 >= g.tiny) {
          const double kk2 = g.anis ? kk2_of(g, kx, ky, kz) : kn2;
          vz_in[k] = rk.v[2][s];
          const cplx th = rk.v[3][s];
          const cplx nzb = c_make(P[2][m].x + rk.xmom * th.x, P[2][m].y + rk.xmom * th.y);
          // rotbouss_rkstep2.f90:6-15: Coriolis on x, y beside the gravity on z
          if (rk.om2 != 0.0) coriolis(P[0][m], P[1][m], rk.v[0][s], rk.v[1][s], rk.om2);
          cplx n0, n1, n2;
          project_hd(P[0][m], P[1][m], nzb, kx, ky, kz, kk2, i, j, k, n0, n1, n2);
          rk.out[0][s] = rk_update(rk.c[0][s], rk.v[0][s], n0, forced ? rk.f[0][s] : z, kk2, rk.nu, rk.dt, rk.rmp);
          rk.out[1][s] = rk_update(rk.c[1][s], rk.v[1][s], n1, forced ? rk.f[1][s] : z, kk2, rk.nu, rk.dt, rk.rmp);
          rk.out[2][s] = rk_update(rk.c[2][s], rk.v[2][s], n2, forced ? rk.f[2][s] : z, kk2, rk.nu, rk.dt, rk.rmp);
        } else if (!prune || kn2 < g.tiny) {
          rk.out[0][s] = z;
          rk.out[1][s] = z;
          rk.out[2][s] = z;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  cplx A[E];
  zfwd_load<N, E>(A, lds, row, stride, L, t, tw, in.p[3], g, i, j0, lmask);
  if (!ok) return;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    const size_t s = sline + k;
    const double kz = wavenum(k, g.nz, g.dkz);
    const double kn2 = kn2_of(g, i, j, k);
    if (kn2 <= g.kmax && kn2 >= g.tiny) {
      const double kk2 = g.anis ? kk2_of(g, kx, ky, kz) : kn2;
      const cplx vz = vz_in[k];
      const cplx a = c_make(A[m].x + rk.xtemp * vz.x, A[m].y + rk.xtemp * vz.y);
      rk.out[3][s] = rk_update(rk.c[3][s], rk.v[3][s], a, forced && rk.f[3] ? rk.f[3][s] : z, kk2, rk.kappa, rk.dt,
                               rk.rmp);
    } else if (kn2 < g.tiny) {
      rk.out[3][s] = rk.c[3][s];
    } else if (!prune) {
      rk.out[3][s] = z;
    }
  }
}

// MHD: nonlin3 (pseudospec3D_mhd.f90:108-227) for the velocity from
// (w x u, j x b) (PART 0) and gauge3 (:230-336) for the potential from u x b
// (PART 1).  The two parts are separate launches so that each keeps its
// fields in registers.
template <int N, int PART, typename Cfg = FFTCfg<N>>
__global__ __launch_bounds__(kTileBound) void k_zfwd_rk_mhd(FieldSet in, GridK g, const cplx *__restrict__ twt, MHDRKParams rk,
                                                    int prune) {
  using F = LineFFT<N, -1, Cfg>;
  constexpr int T = F::T, E = F::E;
  extern __shared__ cplx lds[];
  const int L = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  TwCache<N, Cfg> tw;  // stage twiddle roots of this lane, loaded once
  tw.load(twt, t);
  int bx, by;
  xcd_tile(bx, by);
  const int i = by, j0 = bx * L, j = j0 + lt;
  const unsigned lmask = prune ? tile_lines_active(g, i, j0, L) : ~0u;
  if (lmask == 0u) return;  // uniform over the block
  const bool ok = j < g.ny && ((lmask >> lt) & 1u);
  const int stride = lds_line_stride(N);
  cplx *row = lds + lt * stride;
  // PART 0, velocity: fields 0..2 = w x u, 3..5 = j x b; nonlin3 needs only
  // the difference of their transforms, D = FFT(w x u - j x b), and the
  // difference is taken on the tiles before one transform.  PART 1, vector
  // potential: fields 6..8 = u x b.  The next field's tiles are in flight
  // while this one is transformed.
  const size_t qs = (size_t)g.nxh * g.ny, qb = (size_t)i * g.ny;
  cplx wa[E], wb[E];
  auto fetch = [&](int f) {
    tile_fetch_q<E>(wa, L, in.p[PART == 0 ? f : 6 + f], qs, qb, j0, g.ny, lmask);
    if constexpr (PART == 0) tile_fetch_q<E>(wb, L, in.p[3 + f], qs, qb, j0, g.ny, lmask);
  };
  cplx D[3][E];
  fetch(0);
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if constexpr (PART == 0) {
#pragma unroll
      for (int q = 0; q < E; ++q) wa[q] = c_sub(wa[q], wb[q]);
    }
    tile_put_q<E>(lds, stride, L, wa);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; ++m) D[f][m] = row[lds_pad(t + T * m)];
    __syncthreads();
    if (f < 2) fetch(f + 1);
    F::run(D[f], row, t, tw);
    __syncthreads();
  }
  if (!ok) return;
  constexpr int c0 = PART == 0 ? 0 : 3;
  const double coef = PART == 0 ? rk.nu : rk.mu;
  const double kx = wavenum(i, g.nx, g.dkx), ky = wavenum(j, g.ny, g.dky);
  const size_t sline = ((size_t)i * g.ny + j) * N;
  const bool forced = rk.fline && rk.fline[(size_t)i * g.ny + j];
  const cplx z = c_make(0.0, 0.0);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    const size_t s = sline + k;
    const double kz = wavenum(k, g.nz, g.dkz);
    const double kn2 = kn2_of(g, i, j, k);
    if (kn2 <= g.kmax && kn2 >= g.tiny) {
      const double kk2 = g.anis ? kk2_of(g, kx, ky, kz) : kn2;
      cplx n0, n1, n2;
      // nonlin3 is the nonlhd3 projection of D; gauge3 is its negative
      // applied to u x b (negation is exact, so both match the routines)
      project_hd(D[0][m], D[1][m], D[2][m], kx, ky, kz, kk2, i, j, k, n0, n1, n2);
      auto upd = [&](int c, cplx nl) {
        if constexpr (PART == 1) nl = c_make(-nl.x, -nl.y);
        const cplx ff = forced && rk.f[c0 + c] ? rk.f[c0 + c][s] : z;
        rk.out[c0 + c][s] = rk_update(rk.c[c0 + c][s], rk.v[c0 + c][s], nl, ff, kk2, coef, rk.dt, rk.rmp);
      };
      upd(0, n0);
      upd(1, n1);
      upd(2, n2);
    } else if (!prune || kn2 < g.tiny) {  // the k = 0 mode: as in k_zfwd_rk_hd
      rk.out[c0][s] = z;
      rk.out[c0 + 1][s] = z;
      rk.out[c0 + 2][s] = z;
    }
    __builtin_amdgcn_sched_barrier(0);  // one mode's loads at a time (register pressure)
  }
}

// ---------------------------------------------------------------------------
// Y passes between Q[z][kx][ky] (ky fastest) and P[z][ky][kx] (kx fastest,
// rows of nxp >= nx/2+1 points).  Block = (z, kx chunk kx0..kx0+L-1); the
// ky lines are contiguous in Q and the transpose to or from P goes through
// an LDS tile with L-point (L*16-byte) chunks along kx.  Only kx < kxlim is
// touched (kxlim < nx/2+1 when the fields are dealiased).
// ---------------------------------------------------------------------------
template <int N, typename Cfg = FFTCfg<N>>
__global__ __launch_bounds__(kTileBound) void k_yinv(FieldSet in_, FieldSet out_, int nf, GridK g, int nxp,
                                             const cplx *__restrict__ twt, int kxlim, int prune) {
  using F = LineFFT<N, +1, Cfg>;
  constexpr int T = F::T, E = F::E;
  extern __shared__ cplx lds[];
  const int L = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  TwCache<N, Cfg> tw;  // stage twiddle roots of this lane, loaded once
  tw.load(twt, t);
  const FieldView in = batch_view(in_, blockIdx.z);
  const FieldView out = batch_view(out_, blockIdx.z);
  int bx, by;
  xcd_tile(bx, by);
  const int z = by, kx0 = bx * L, kx = kx0 + lt;
  if (kx0 >= kxlim) return;  // uniform over the block
  const bool ok = kx < kxlim;
  const int stride = lds_line_stride(N);
  cplx *row = lds + lt * stride;
  // dealiased input: Q holds data only on the lines (kx,ky) of the dealias
  // disc; every lane loads inside its line and a select drops the rest
  auto load = [&](int f, cplx (&v)[E]) {
    const cplx *src = in.p[f] + ((size_t)z * g.nxh + (ok ? kx : 0)) * N;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int j = t + T * m;
      const bool need = ok && (!prune || kn2_of(g, kx, j, 0) <= g.kmax);
      const cplx x = src[need ? j : 0];
      v[m] = need ? x : c_make(0.0, 0.0);
    }
  };
  // E <= 8: the next field's line is loaded while this one is transformed;
  // wider schedules have no registers for it and rely on a second block per CU
  constexpr bool prefetch = E <= 8;
  cplx nxt[prefetch ? E : 1];
  if constexpr (prefetch) load(0, nxt);
  for (int f = 0; f < nf; ++f) {
    cplx v[E];
    if constexpr (prefetch) {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = nxt[m];
      if (f + 1 < nf) load(f + 1, nxt);  // the next field's line in flight during this one
    } else {
      load(f, v);
    }
    F::run(v, row, t, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) row[lds_pad(t + T * m)] = v[m];
    __syncthreads();
    tile_store_rows<E, true>(lds, stride, L, out.p[f] + (size_t)z * g.ny * nxp + kx0, (size_t)nxp,
                       [=](int l) { return kx0 + l < kxlim; });
    __syncthreads();
  }
}

template <int N, typename Cfg = FFTCfg<N>>
__global__ __launch_bounds__(kTileBound) void k_yfwd(FieldSet in_, FieldSet out_, int nf, GridK g, int nxp,
                                             const cplx *__restrict__ twt, int kxlim, int prune) {
  using F = LineFFT<N, -1, Cfg>;
  constexpr int T = F::T, E = F::E;
  extern __shared__ cplx lds[];
  const int L = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  TwCache<N, Cfg> tw;  // stage twiddle roots of this lane, loaded once
  tw.load(twt, t);
  const FieldView in = batch_view(in_, blockIdx.z);
  const FieldView out = batch_view(out_, blockIdx.z);
  int bx, by;
  xcd_tile(bx, by);
  const int z = by, kx0 = bx * L, kx = kx0 + lt;
  if (kx0 >= kxlim) return;  // uniform over the block
  const bool ok = kx < kxlim;
  const int stride = lds_line_stride(N);
  cplx *row = lds + lt * stride;
  // The block's tile of a field: E points per thread (L*N points over L*T
  // threads), loaded into registers with in-tile addresses (columns at or
  // beyond kxlim read column kx0 and are zeroed), so the loads issue back to
  // back and the next field's tile is in flight while this one is transformed.
  auto tload = [&](int f, cplx (&w)[E]) {
    const cplx *src = in.p[f] + (size_t)z * g.ny * nxp;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = threadIdx.x + q * blockDim.x;
      const int l = idx % L, p = idx / L;
      const bool inside = kx0 + l < kxlim;
      // a 32-bit offset from the block's uniform base (ny * nxp < 2^31)
      const cplx x = src[p * nxp + kx0 + (inside ? l : 0)];
      w[q] = inside ? x : c_make(0.0, 0.0);
    }
  };
  cplx nxt[E];
  tload(0, nxt);
  for (int f = 0; f < nf; ++f) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = threadIdx.x + q * blockDim.x;
      lds[(idx % L) * stride + lds_pad(idx / L)] = nxt[q];
    }
    __syncthreads();
    if (f + 1 < nf) tload(f + 1, nxt);
    cplx v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = row[lds_pad(t + T * m)];
    __syncthreads();
    F::run(v, row, t, tw);
    if (ok) {
      cplx *dst = out.p[f] + ((size_t)z * g.nxh + kx) * N;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int j = t + T * m;
        // dealiased output: only the ky of the dealias disc are used downstream
        if (!prune || kn2_of(g, kx, j, 0) <= g.kmax) c_put<true>(dst + j, v[m]);
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// X passes on contiguous rows: P rows (nx/2+1 points, kx fastest) <-> real
// rows of nx points held as N = nx/2 packed complex points.  A row is one
// line of T <= 64 lanes (one wavefront for nx = 512), so these kernels use
// only wave-level synchronisation.  Block = RPB rows.
// ---------------------------------------------------------------------------

// Twiddles of the x pass held by one lane for the whole kernel: the stage
// roots of the half-length transform and the packing twiddles w^k of the
// real <-> half-complex step (k = t + T m).
template <int N, typename Cfg> struct XTw {
  static constexpr int E = Cfg::E, T = Cfg::T;
  TwCache<N, Cfg> stage;
  cplx pack[E];
  GHX_DEV void load(const cplx *twh, const cplx *twx, int t) {
    stage.load(twh, t);
#pragma unroll
    for (int m = 0; m < E; ++m) pack[m] = twx[t + T * m];
  }
};

// The loads of one P row for X^-1: points k = t + T m and the Nyquist point.
// Every lane loads from inside the row (points at kx >= kxlim read point 0
// and are zeroed by a select), so the loads issue back to back and a caller
// can start the next row's loads before transforming this one.
template <int E> struct XRowIn {
  cplx a[E];
  cplx nyq;
};

template <int N, int E, typename Cfg>
GHX_DEV void xrow_load(XRowIn<E> &r, const cplx *src, int t, int kxlim) {
  constexpr int T = Cfg::T;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    r.a[m] = src[k < kxlim ? k : 0];
  }
  r.nyq = src[N < kxlim ? N : 0];
}

// Half-spectrum row -> unnormalised c2r of the row, packed: v[m] holds
// (x[2k], x[2k+1]) * nx for k = t + T m.  GHOST's c2r ignores the imaginary
// parts of the kx = 0 and Nyquist entries.  The mirrored points X[N-k] come
// through the line's LDS buffer (slot N holds the Nyquist point).
template <int N, int E, typename Cfg>
GHX_DEV void xrow_inv(cplx (&v)[E], const XRowIn<E> &r, cplx *buf, int t, const XTw<N, Cfg> &tw, int kxlim) {
  using F = LineFFT<N, +1, Cfg>;
  constexpr int T = F::T;
  cplx a[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    a[m] = k < kxlim ? r.a[m] : c_make(0.0, 0.0);
    buf[lds_pad(k)] = a[m];
  }
  if (t == 0) buf[lds_pad(N)] = N < kxlim ? r.nyq : c_make(0.0, 0.0);
  line_sync<T>();
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    cplx A = a[m];
    cplx Bs = buf[lds_pad(N - k)];
    if (k == 0) { A.y = 0.0; Bs.y = 0.0; }
    cplx B = c_conj(Bs);
    cplx Xe = c_add(A, B);
    cplx Xo = c_mul(c_sub(A, B), c_conj(tw.pack[m]));
    v[m] = c_add(Xe, c_make(-Xo.y, Xo.x));
  }
  line_sync<T>();
  F::run(v, buf, t, tw.stage);
}

// Packed real row -> r2c half spectrum, stored to a P row (kx < kxlim only).
template <int N, int E, typename Cfg>
GHX_DEV void xrow_fwd(cplx (&v)[E], cplx *buf, int t, const XTw<N, Cfg> &tw, cplx *dst, int kxlim) {
  using F = LineFFT<N, -1, Cfg>;
  constexpr int T = F::T;
  F::run(v, buf, t, tw.stage);
#pragma unroll
  for (int m = 0; m < E; ++m) buf[lds_pad(t + T * m)] = v[m];
  line_sync<T>();
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    cplx A = v[m];
    cplx B = c_conj(buf[lds_pad((N - k) % N)]);
    cplx Xe = c_scale(c_add(A, B), 0.5);
    cplx d = c_sub(A, B);
    cplx Xo = c_make(0.5 * d.y, -0.5 * d.x);  // -0.5 i (A - B)
    if (k < kxlim) dst[k] = c_add(Xe, c_mul(tw.pack[m], Xo));
  }
  if (t == 0 && N < kxlim) {
    cplx z0 = v[0];
    dst[N] = c_make(z0.x - z0.y, 0.0);
  }
  line_sync<T>();
}

struct RealSet {
  double *p[kMaxFields];
};

template <int N>
__global__ __launch_bounds__(kRowBound) void k_xinv(FieldSet in, RealSet out, int nf, GridK g, int nxp,
                                             const cplx *__restrict__ twh, const cplx *__restrict__ twx,
                                             int kxlim) {
  using Cfg = FFTCfgLow<N>;
  constexpr int T = Cfg::T, E = Cfg::E;
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long nrows = (long)g.nz * g.ny;
  const long r = (long)blockIdx.x * rpb + lt;
  const bool ok = r < nrows;
  if constexpr (T <= 64) {
    if (!ok) return;  // only wave-level syncs
  }
  const int lim = ok ? kxlim : 0;
  const size_t off = (size_t)(ok ? r : 0) * nxp;
  cplx *buf = lds + lt * lds_line_stride(N);
  XTw<N, Cfg> tw;
  tw.load(twh, twx, t);
  XRowIn<E> cur, nxt;
  xrow_load<N, E, Cfg>(cur, in.p[0] + off, t, lim);
  for (int f = 0; f < nf; ++f) {
    if (f + 1 < nf) xrow_load<N, E, Cfg>(nxt, in.p[f + 1] + off, t, lim);  // next field in flight
    cplx v[E];
    xrow_inv<N, E, Cfg>(v, cur, buf, t, tw, lim);
    if (ok) {
      cplx *dst = reinterpret_cast<cplx *>(out.p[f] + r * g.nx);
#pragma unroll
      for (int m = 0; m < E; ++m) dst[t + T * m] = v[m];
    }
    line_sync<T>();
    cur = nxt;
  }
}

// X^-1 of three fields written interleaved, out[(row * nx + x) * 3 + f]: the
// layout the tracer gather reads (one 4-point stencil row of all three
// components is 96 contiguous bytes).
template <int N>
__global__ __launch_bounds__(kRowBound) void k_xinv3i(FieldSet in, double *out, GridK g, int nxp,
                                               const cplx *__restrict__ twh, const cplx *__restrict__ twx,
                                               int kxlim) {
  using Cfg = FFTCfgLow<N>;
  constexpr int T = Cfg::T, E = Cfg::E;
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long nrows = (long)g.nz * g.ny;
  const long r = (long)blockIdx.x * rpb + lt;
  const bool ok = r < nrows;
  if constexpr (T <= 64) {
    if (!ok) return;
  }
  const int lim = ok ? kxlim : 0;
  const size_t off = (size_t)(ok ? r : 0) * nxp;
  cplx *buf = lds + lt * lds_line_stride(N);
  XTw<N, Cfg> tw;
  tw.load(twh, twx, t);
  XRowIn<E> r0, r1;
  xrow_load<N, E, Cfg>(r0, in.p[0] + off, t, lim);
  xrow_load<N, E, Cfg>(r1, in.p[1] + off, t, lim);
  cplx v[3][E];
  xrow_inv<N, E, Cfg>(v[0], r0, buf, t, tw, lim);
  xrow_load<N, E, Cfg>(r0, in.p[2] + off, t, lim);
  xrow_inv<N, E, Cfg>(v[1], r1, buf, t, tw, lim);
  xrow_inv<N, E, Cfg>(v[2], r0, buf, t, tw, lim);
  if (!ok) return;
  // lane point k holds x = 2k and 2k+1: six consecutive doubles, 16-byte aligned
  double2 *dst = reinterpret_cast<double2 *>(out + (size_t)r * g.nx * 3);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    dst[3 * k + 0] = make_double2(v[0][m].x, v[1][m].x);
    dst[3 * k + 1] = make_double2(v[2][m].x, v[0][m].y);
    dst[3 * k + 2] = make_double2(v[1][m].y, v[2][m].y);
  }
}

template <int N>
__global__ __launch_bounds__(kRowBound) void k_xfwd(RealSet in, FieldSet out, int nf, GridK g, int nxp,
                                             const cplx *__restrict__ twh, const cplx *__restrict__ twx) {
  using Cfg = FFTCfgLow<N>;
  constexpr int T = Cfg::T, E = Cfg::E;
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long nrows = (long)g.nz * g.ny;
  const long r = (long)blockIdx.x * rpb + lt;
  const bool ok = r < nrows;
  if constexpr (T <= 64) {
    if (!ok) return;
  }
  cplx *buf = lds + lt * lds_line_stride(N);
  XTw<N, Cfg> tw;
  tw.load(twh, twx, t);
  const size_t roff = (size_t)(ok ? r : 0) * g.nx;
  cplx cur[E], nxt[E];
  {
    const cplx *src = reinterpret_cast<const cplx *>(in.p[0] + roff);
#pragma unroll
    for (int m = 0; m < E; ++m) cur[m] = src[t + T * m];
  }
  for (int f = 0; f < nf; ++f) {
    if (f + 1 < nf) {
      const cplx *src = reinterpret_cast<const cplx *>(in.p[f + 1] + roff);
#pragma unroll
      for (int m = 0; m < E; ++m) nxt[m] = src[t + T * m];
    }
    cplx v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = ok ? cur[m] : c_make(0.0, 0.0);
    xrow_fwd<N, E, Cfg>(v, buf, t, tw, out.p[f] + (ok ? r : 0) * nxp, ok ? g.nxh : 0);
#pragma unroll
    for (int m = 0; m < E; ++m) cur[m] = nxt[m];
  }
}

// ---------------------------------------------------------------------------
// Two-for-one x passes.  Two Hermitian half rows A, B (kx = 0..NX/2) make one
// complex line of NX points, z[k] = A[k] + i B[k] with A[NX-k] = conj A[k];
// its inverse transform is x_a + i x_b, the two real rows at once.  Forward,
// z = x_a + i x_b gives A[k] = (Z[k] + conj Z[NX-k])/2 and
// B[k] = (Z[k] - conj Z[NX-k])/(2i).  Against one half-length transform per
// row plus its packing step, a pair costs one full-length transform (two LDS
// exchanges for NX = 512 with radix 8^3) and no mirror exchange on the
// inverse side: the mirrored points are read from the row itself.
// ---------------------------------------------------------------------------
// Loads of one pair: lane t takes the points k = t + T m of the full line,
// the upper half from the mirrored kx.  Points with kx >= kxlim read kx = 0
// (kept in-row so the loads issue back to back) and are zeroed when the line
// is formed.
template <int NX, int E, int T>
GHX_DEV void xpair_load(cplx (&A)[E], cplx (&B)[E], const cplx *sa, const cplx *sb, int t, int kxlim) {
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    const int kk = k <= NX / 2 ? k : NX - k;
    const int i = kk < kxlim ? kk : 0;
    A[m] = sa[i];
    B[m] = sb[i];
  }
}

// z[m] = A[k] + i B[k] over the full line.  GHOST's c2r ignores the
// imaginary parts of the kx = 0 and Nyquist entries.
// With hasb false the line carries A alone (B taken as zero).
template <int NX, int E, int T>
GHX_DEV void xpair_make(cplx (&z)[E], const cplx (&A)[E], const cplx (&B)[E], int t, int kxlim, bool hasb = true) {
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + T * m;
    const int kk = k <= NX / 2 ? k : NX - k;
    const bool live = kk < kxlim;
    const bool realpt = (kk == 0) || (kk == NX / 2);
    const double s = k > NX / 2 ? -1.0 : 1.0;  // upper half: conjugate
    const bool lb = live && hasb;
    const double ar = live ? A[m].x : 0.0, br = lb ? B[m].x : 0.0;
    const double ai = (live && !realpt) ? s * A[m].y : 0.0, bi = (lb && !realpt) ? s * B[m].y : 0.0;
    z[m] = c_make(ar - bi, ai + br);
  }
}

// Z = forward transform of x_a + i x_b in natural order (z[m] = Z[t + T m]);
// stores A[k], B[k] for k < kxlim (k <= NX/2) to the rows da, db.  The
// mirrored Z[NX-k] come through the line's LDS buffer.
template <int NX, int E, int T>
GHX_DEV void xpair_store2(const cplx (&z)[E], cplx *buf, int t, cplx *da, cplx *db, int kxlim) {
#pragma unroll
  for (int m = 0; m < E; ++m) buf[lds_pad(t + T * m)] = z[m];
  line_sync<T>();
#pragma unroll
  for (int m = 0; m <= E / 2; ++m) {
    const int k = t + T * m;
    if (k <= NX / 2) {
      const cplx Zc = c_conj(buf[lds_pad((NX - k) % NX)]);
      const cplx a = c_scale(c_add(z[m], Zc), 0.5);
      const cplx d = c_sub(z[m], Zc);
      const cplx b = c_make(0.5 * d.y, -0.5 * d.x);  // -0.5 i (Z - conj Z')
      if (k < kxlim) {
        da[k] = a;
        db[k] = b;
      }
    }
  }
  line_sync<T>();
}

// single row: z = forward transform of x_a (imaginary input zero) is A itself
template <int NX, int E, int T>
GHX_DEV void xpair_store1(const cplx (&z)[E], int t, cplx *da, int kxlim) {
#pragma unroll
  for (int m = 0; m <= E / 2; ++m) {
    const int k = t + T * m;
    if (k <= NX / 2 && k < kxlim) da[k] = z[m];
  }
}

// x_a + i x_b -> the line's samples n = t + T m; the single-row form leaves
// the imaginary input zero
template <int E>
GHX_DEV void xpair_real(cplx (&z)[E], const double *sa, const double *sb, int t, int T) {
#pragma unroll
  for (int m = 0; m < E; ++m) z[m] = c_make(sa[t + T * m], sb ? sb[t + T * m] : 0.0);
}

// X^-1 of nf fields, two rows per transform (plain c2r; the last field of an
// odd count goes alone).
template <int NX>
__global__ __launch_bounds__(kRowBound) void k_xinv2(FieldSet in, RealSet out, int nf, GridK g, int nxp,
                                              const cplx *__restrict__ twt, int kxlim) {
  using Cfg = FFTCfg<NX>;
  using FI = LineFFT<NX, +1>;
  constexpr int T = Cfg::T, E = Cfg::E;
  static_assert(T <= 64, "one row per wavefront or less");
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long r = (long)blockIdx.x * rpb + lt;
  if (r >= (long)g.nz * g.ny) return;
  const size_t off = (size_t)r * nxp, roff = (size_t)r * NX;
  cplx *buf = lds + lt * lds_line_stride(NX);
  TwCache<NX> tw;
  tw.load(twt, t);
  cplx a[E], b[E], z[E];
  const int np = (nf + 1) / 2;
  xpair_load<NX, E, T>(a, b, in.p[0] + off, in.p[nf > 1 ? 1 : 0] + off, t, kxlim);
  for (int p = 0; p < np; ++p) {
    const bool two = 2 * p + 1 < nf;
    xpair_make<NX, E, T>(z, a, b, t, kxlim, two);
    if (p + 1 < np) {  // next pair in flight
      const int fa = 2 * p + 2, fb = fa + 1 < nf ? fa + 1 : fa;
      xpair_load<NX, E, T>(a, b, in.p[fa] + off, in.p[fb] + off, t, kxlim);
    }
    FI::run(z, buf, t, tw);
    double *da = out.p[2 * p] + roff;
#pragma unroll
    for (int m = 0; m < E; ++m) da[t + T * m] = z[m].x;
    if (two) {
      double *db = out.p[2 * p + 1] + roff;
#pragma unroll
      for (int m = 0; m < E; ++m) db[t + T * m] = z[m].y;
    }
  }
}

// X^-1 of three fields written interleaved (the tracer layout of k_xinv3i),
// pairs (0, 1) and (2, alone).
template <int NX>
__global__ __launch_bounds__(kRowBound) void k_xinv3i2(FieldSet in, double *out, GridK g, int nxp,
                                                const cplx *__restrict__ twt, int kxlim) {
  using Cfg = FFTCfg<NX>;
  using FI = LineFFT<NX, +1>;
  constexpr int T = Cfg::T, E = Cfg::E;
  static_assert(T <= 64, "one row per wavefront or less");
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long r = (long)blockIdx.x * rpb + lt;
  if (r >= (long)g.nz * g.ny) return;
  const size_t off = (size_t)r * nxp;
  cplx *buf = lds + lt * lds_line_stride(NX);
  TwCache<NX> tw;
  tw.load(twt, t);
  cplx a[E], b[E], z0[E], z1[E];
  xpair_load<NX, E, T>(a, b, in.p[0] + off, in.p[1] + off, t, kxlim);
  xpair_make<NX, E, T>(z0, a, b, t, kxlim);
  xpair_load<NX, E, T>(a, b, in.p[2] + off, in.p[2] + off, t, kxlim);
  FI::run(z0, buf, t, tw);
  xpair_make<NX, E, T>(z1, a, b, t, kxlim, false);
  FI::run(z1, buf, t, tw);
  double *dst = out + (size_t)r * NX * 3;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = t + T * m;
    d_put<true>(dst + 3 * n + 0, z0[m].x);
    d_put<true>(dst + 3 * n + 1, z0[m].y);
    d_put<true>(dst + 3 * n + 2, z1[m].x);
  }
}

// X of nf real rows, two rows per transform.
template <int NX>
__global__ __launch_bounds__(kRowBound) void k_xfwd2(RealSet in, FieldSet out, int nf, GridK g, int nxp,
                                              const cplx *__restrict__ twt) {
  using Cfg = FFTCfg<NX>;
  using FF = LineFFT<NX, -1>;
  constexpr int T = Cfg::T, E = Cfg::E;
  static_assert(T <= 64, "one row per wavefront or less");
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long r = (long)blockIdx.x * rpb + lt;
  if (r >= (long)g.nz * g.ny) return;
  const size_t off = (size_t)r * nxp, roff = (size_t)r * NX;
  cplx *buf = lds + lt * lds_line_stride(NX);
  TwCache<NX> tw;
  tw.load(twt, t);
  cplx z[E], nxt[E];
  const int np = (nf + 1) / 2;
  xpair_real<E>(z, in.p[0] + roff, nf > 1 ? in.p[1] + roff : nullptr, t, T);
  for (int p = 0; p < np; ++p) {
    const bool two = 2 * p + 1 < nf;
    if (p + 1 < np) {
      const int fa = 2 * p + 2;
      xpair_real<E>(nxt, in.p[fa] + roff, fa + 1 < nf ? in.p[fa + 1] + roff : nullptr, t, T);
    }
    FF::run(z, buf, t, tw);
    if (two)
      xpair_store2<NX, E, T>(z, buf, t, out.p[2 * p] + off, out.p[2 * p + 1] + off, NX / 2 + 1);
    else
      xpair_store1<NX, E, T>(z, t, out.p[2 * p] + off, NX / 2 + 1);
#pragma unroll
    for (int m = 0; m < E; ++m) z[m] = nxt[m];
  }
}

// Fused X^-1 . ((A_x B_x + A_y B_y) + A_z B_z) tmp . X (advect3) with pairs
// (A_x, A_y), (A_z, B_x), (B_y, B_z); fields as k_xdot.
template <int NX>
__global__ __launch_bounds__(kRowBound) void k_xdot2(FieldSet in, FieldSet out, GridK g, int nxp,
                                              const cplx *__restrict__ twt, int kxlim) {
  using Cfg = FFTCfg<NX>;
  using FI = LineFFT<NX, +1>;
  using FF = LineFFT<NX, -1>;
  constexpr int T = Cfg::T, E = Cfg::E;
  static_assert(T <= 64, "one row per wavefront or less");
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long r = (long)blockIdx.x * rpb + lt;
  if (r >= (long)g.nz * g.ny) return;
  const size_t off = (size_t)r * nxp;
  cplx *buf = lds + lt * lds_line_stride(NX);
  const double tmp = -g.inv_n6;
  TwCache<NX> tw;
  tw.load(twt, t);
  cplx a[E], b[E], z0[E], z1[E], z2[E];
  xpair_load<NX, E, T>(a, b, in.p[0] + off, in.p[1] + off, t, kxlim);
  xpair_make<NX, E, T>(z0, a, b, t, kxlim);
  xpair_load<NX, E, T>(a, b, in.p[2] + off, in.p[3] + off, t, kxlim);
  FI::run(z0, buf, t, tw);
  xpair_make<NX, E, T>(z1, a, b, t, kxlim);
  xpair_load<NX, E, T>(a, b, in.p[4] + off, in.p[5] + off, t, kxlim);
  FI::run(z1, buf, t, tw);
  xpair_make<NX, E, T>(z2, a, b, t, kxlim);
  FI::run(z2, buf, t, tw);
  // samples: A_x = z0.x, A_y = z0.y, A_z = z1.x, B_x = z1.y, B_y = z2.x, B_z = z2.y
#pragma unroll
  for (int m = 0; m < E; ++m)
    z0[m] = c_make(((z0[m].x * z1[m].y + z0[m].y * z2[m].x) + z1[m].x * z2[m].y) * tmp, 0.0);
  FF::run(z0, buf, t, tw);
  xpair_store1<NX, E, T>(z0, t, out.p[0] + off, kxlim);
}

// Fused X^-1 . (P x Q)/N^6 . X with two-for-one transforms: pairs (P_x, P_y),
// (P_z, Q_x), (Q_y, Q_z) inverse, (out_x, out_y) and out_z forward.  Same
// fields, products and operand order as k_xcross.  One row per line of
// T <= 64 lanes.
template <int NX>
__global__ __launch_bounds__(kRowBound) void k_xcross2(FieldSet in_, FieldSet out_, GridK g, int nxp,
                                                const cplx *__restrict__ twt, int kxlim) {
  using Cfg = FFTCfg<NX>;
  using FI = LineFFT<NX, +1>;
  using FF = LineFFT<NX, -1>;
  constexpr int T = Cfg::T, E = Cfg::E;
  static_assert(T <= 64, "one row per wavefront or less");
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long nrows = (long)g.nz * g.ny;
  const long r = (long)blockIdx.x * rpb + lt;
  if (r >= nrows) return;  // whole lines only: wave-level syncs
  const size_t off = (size_t)r * nxp;
  cplx *buf = lds + lt * lds_line_stride(NX);
  const double tmp = g.inv_n6;
  const FieldView in = batch_view(in_, blockIdx.y);
  const FieldView out = batch_view(out_, blockIdx.y);
  TwCache<NX> tw;
  tw.load(twt, t);
  cplx a[E], b[E], z0[E], z1[E], z2[E];
  xpair_load<NX, E, T>(a, b, in.p[0] + off, in.p[1] + off, t, kxlim);
  xpair_make<NX, E, T>(z0, a, b, t, kxlim);
  xpair_load<NX, E, T>(a, b, in.p[2] + off, in.p[3] + off, t, kxlim);
  FI::run(z0, buf, t, tw);
  xpair_make<NX, E, T>(z1, a, b, t, kxlim);
  xpair_load<NX, E, T>(a, b, in.p[4] + off, in.p[5] + off, t, kxlim);
  FI::run(z1, buf, t, tw);
  xpair_make<NX, E, T>(z2, a, b, t, kxlim);
  FI::run(z2, buf, t, tw);
  // samples: P_x = z0.x, P_y = z0.y, P_z = z1.x, Q_x = z1.y, Q_y = z2.x, Q_z = z2.y
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const double px = z0[m].x, py = z0[m].y, pz = z1[m].x, qx = z1[m].y, qy = z2[m].x, qz = z2[m].y;
    const double ox = (py * qz - qy * pz) * tmp;
    const double oy = (pz * qx - qz * px) * tmp;
    const double oz = (px * qy - qx * py) * tmp;
    z0[m] = c_make(ox, oy);
    z1[m] = c_make(oz, 0.0);
  }
  FF::run(z0, buf, t, tw);
  xpair_store2<NX, E, T>(z0, buf, t, out.p[0] + off, out.p[1] + off, kxlim);
  FF::run(z1, buf, t, tw);
  xpair_store1<NX, E, T>(z1, t, out.p[2] + off, kxlim);
}

// Fused X^-1 . (P x Q)/N^6 . X on P rows with GHOST's operand order
// (pseudospec3D_hd.f90:377-389): out_x = (P_y Q_z - Q_y P_z) tmp,
// out_y = (P_z Q_x - Q_z P_x) tmp, out_z = (P_x Q_y - Q_x P_y) tmp.
// Fields: in.p[0..2] = P, in.p[3..5] = Q; out.p[0..2].  Only kx < kxlim is
// read and written (dealiased inputs, dealiased use of the outputs).  Each
// row's next input is loaded before the current one is transformed.
template <int N, int W, int SB>
__global__ __launch_bounds__(kRowBound) __attribute__((amdgpu_waves_per_eu(W))) void k_xcross(
    FieldSet in_, FieldSet out_, GridK g, int nxp, const cplx *__restrict__ twh, const cplx *__restrict__ twx,
    int kxlim) {
  using Cfg = FFTCfgLow<N>;
  constexpr int T = Cfg::T, E = Cfg::E;
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long nrows = (long)g.nz * g.ny;
  const long r = (long)blockIdx.x * rpb + lt;
  const bool ok = r < nrows;
  if constexpr (T <= 64) {
    if (!ok) return;
  }
  const int lim = ok ? kxlim : 0;
  const size_t off = (size_t)(ok ? r : 0) * nxp;
  cplx *buf = lds + lt * lds_line_stride(N);
  const double tmp = g.inv_n6;
  const FieldView in = batch_view(in_, blockIdx.y);
  const FieldView out = batch_view(out_, blockIdx.y);
  XTw<N, Cfg> tw;
  tw.load(twh, twx, t);
  cplx qx[E], qy[E], qz[E], ay[E], az[E];
  XRowIn<E> ra, rb;
#define GHX_PHASE() do { if constexpr (SB) __builtin_amdgcn_sched_barrier(0); } while (0)
  xrow_load<N, E, Cfg>(ra, in.p[3] + off, t, lim);
  xrow_load<N, E, Cfg>(rb, in.p[4] + off, t, lim);
  xrow_inv<N, E, Cfg>(qx, ra, buf, t, tw, lim);
  GHX_PHASE();
  xrow_load<N, E, Cfg>(ra, in.p[5] + off, t, lim);
  xrow_inv<N, E, Cfg>(qy, rb, buf, t, tw, lim);
  GHX_PHASE();
  xrow_load<N, E, Cfg>(rb, in.p[0] + off, t, lim);
  xrow_inv<N, E, Cfg>(qz, ra, buf, t, tw, lim);
  GHX_PHASE();
  xrow_load<N, E, Cfg>(ra, in.p[1] + off, t, lim);
  {
    cplx px[E];
    xrow_inv<N, E, Cfg>(px, rb, buf, t, tw, lim);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      ay[m] = c_make(qz[m].x * px[m].x, qz[m].y * px[m].y);  // Q_z P_x
      az[m] = c_make(px[m].x * qy[m].x, px[m].y * qy[m].y);  // P_x Q_y
    }
  }
  GHX_PHASE();
  xrow_load<N, E, Cfg>(rb, in.p[2] + off, t, lim);
  {
    cplx py[E];
    xrow_inv<N, E, Cfg>(py, ra, buf, t, tw, lim);
    cplx oz[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      oz[m] = c_make((az[m].x - qx[m].x * py[m].x) * tmp, (az[m].y - qx[m].y * py[m].y) * tmp);
      az[m] = c_make(py[m].x * qz[m].x, py[m].y * qz[m].y);  // now P_y Q_z
    }
    xrow_fwd<N, E, Cfg>(oz, buf, t, tw, out.p[2] + off, lim);
  }
  GHX_PHASE();
  {
    cplx pz[E];
    xrow_inv<N, E, Cfg>(pz, rb, buf, t, tw, lim);
    cplx ox[E], oy[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      ox[m] = c_make((az[m].x - qy[m].x * pz[m].x) * tmp, (az[m].y - qy[m].y * pz[m].y) * tmp);
      oy[m] = c_make((pz[m].x * qx[m].x - ay[m].x) * tmp, (pz[m].y * qx[m].y - ay[m].y) * tmp);
    }
    GHX_PHASE();
    xrow_fwd<N, E, Cfg>(ox, buf, t, tw, out.p[0] + off, lim);
    GHX_PHASE();
    xrow_fwd<N, E, Cfg>(oy, buf, t, tw, out.p[1] + off, lim);
  }
#undef GHX_PHASE
}

// Fused X^-1 . (A . B) tmp . X on P rows for advect3 (pseudospec3D_phd.f90:
// 23-115): in.p[0..2] = A, in.p[3..5] = B (grad of the scalar), out.p[0] =
// FFT[((A_x B_x + A_y B_y) + A_z B_z) tmp] with tmp = -1/N^6.
template <int N>
__global__ __launch_bounds__(kRowBound) void k_xdot(FieldSet in, FieldSet out, GridK g, int nxp,
                                             const cplx *__restrict__ twh, const cplx *__restrict__ twx, int kxlim) {
  using Cfg = FFTCfgLow<N>;
  constexpr int T = Cfg::T, E = Cfg::E;
  extern __shared__ cplx lds[];
  const int rpb = blockDim.x / T;
  const int lt = threadIdx.x / T, t = threadIdx.x % T;
  const long nrows = (long)g.nz * g.ny;
  const long r = (long)blockIdx.x * rpb + lt;
  const bool ok = r < nrows;
  if constexpr (T <= 64) {
    if (!ok) return;
  }
  const int lim = ok ? kxlim : 0;
  const size_t off = (size_t)(ok ? r : 0) * nxp;
  cplx *buf = lds + lt * lds_line_stride(N);
  const double tmp = -g.inv_n6;
  XTw<N, Cfg> tw;
  tw.load(twh, twx, t);
  XRowIn<E> ra, rb;
  cplx acc[E];
  xrow_load<N, E, Cfg>(ra, in.p[0] + off, t, lim);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    cplx a[E], b[E];
    xrow_load<N, E, Cfg>(rb, in.p[3 + d] + off, t, lim);
    xrow_inv<N, E, Cfg>(a, ra, buf, t, tw, lim);
    if (d < 2) xrow_load<N, E, Cfg>(ra, in.p[d + 1] + off, t, lim);
    xrow_inv<N, E, Cfg>(b, rb, buf, t, tw, lim);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const cplx p = c_make(a[m].x * b[m].x, a[m].y * b[m].y);
      if (d == 0) acc[m] = p;
      else if (d == 1) acc[m] = c_make(acc[m].x + p.x, acc[m].y + p.y);
      else acc[m] = c_make((acc[m].x + p.x) * tmp, (acc[m].y + p.y) * tmp);
    }
  }
  xrow_fwd<N, E, Cfg>(acc, buf, t, tw, out.p[0] + off, lim);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------


#if GHX_PART == 1
int active_kx(const GridK &g) {
  int n = 0;
  for (int q = 0; q < g.nxh; ++q) {
    double kx = (double)(q < g.nx / 2 ? q : q - g.nx);
    if (g.rmp * (kx * kx) <= g.kmax) n = q + 1;
  }
  return n;
}

#endif

#if GHX_PART == 2
// 512-point z lines on the 16-points-per-lane schedule: 8-line tiles (128-byte
// chunks of Q) from 256-lane blocks, for a pass that keeps one field live
template <int MODE>
static hipError_t launch_zinv_wide(const FFTCtx &c, int nf, const FieldSet &in, const FieldSet &out, int prune,
                                   const double *aux) {
  const GridK &g = c.gk;
  using Cfg = FFTCfgWide<512>;
  constexpr int T = Cfg::T;
  const PassShape s = wide_shape(T, lds_line_stride(512), g.ny);
  const int L = s.lines;
  dim3 grid((g.ny + L - 1) / L, g.nxh, c.batch);
  allow_lds(k_zinv<512, MODE, Cfg>, s.lds);
  GHX_LAUNCH_CHECK((k_zinv<512, MODE, Cfg>), s, kTileBound, T, true);
  hipLaunchKernelGGL((k_zinv<512, MODE, Cfg>), grid, dim3(L * T), s.lds, c.stream, in, out, nf, g, c.tw_z, prune, aux);
  return hipGetLastError();
}

// the curl Z^-1 on the four-points-per-lane schedule (FFTCtx::sched bit 0)
template <int N>
static hipError_t launch_zinv_low(const FFTCtx &c, int nf, const FieldSet &in, const FieldSet &out, int prune,
                                  const double *aux) {
  const GridK &g = c.gk;
  using Cfg = FFTCfgLow<N>;
  constexpr int T = Cfg::T;
  const PassShape s = tile_shape(T, lds_line_stride(N), g.ny, tile_pick("GHX_LZ", c.tile_z, 4), max_lds());
  const int L = s.lines;
  dim3 grid((g.ny + L - 1) / L, g.nxh, c.batch);
  allow_lds(k_zinv<N, 1, Cfg>, s.lds);
  GHX_LAUNCH_CHECK((k_zinv<N, 1, Cfg>), s, kTileBound, T, true);
  hipLaunchKernelGGL((k_zinv<N, 1, Cfg>), grid, dim3(L * T), s.lds, c.stream, in, out, nf, g, c.tw_z, prune, aux);
  return hipGetLastError();
}

hipError_t zinv(const FFTCtx &c, int mode, int nf, const FieldSet &in, const FieldSet &out, int prune,
                const double *aux) {
  if (mode == 3 && !aux) return hipErrorInvalidValue;
  const GridK &g = c.gk;
  if (mode == 1 && (c.sched & 1)) {
    switch (g.nz) {
    case 32: return launch_zinv_low<32>(c, nf, in, out, prune, aux);
    case 64: return launch_zinv_low<64>(c, nf, in, out, prune, aux);
    case 128: return launch_zinv_low<128>(c, nf, in, out, prune, aux);
    case 256: return launch_zinv_low<256>(c, nf, in, out, prune, aux);
    default: break;  // the same schedule either way
    }
  }
  // the spline prefilter pass at 512 points runs the wide schedule (GHX_ZWIDE=0
  // selects the default one); the plain and curl passes are slower on it
  if (mode == 3 && g.nz == 512 && env_switch("GHX_ZWIDE", 1) != 0 &&
      wide_shape(FFTCfgWide<512>::T, lds_line_stride(512), g.ny).lds <= max_lds())
    return launch_zinv_wide<3>(c, nf, in, out, prune, aux);
  switch (g.nz) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = tile_shape(T, lds_line_stride(n_), g.ny, tile_pick("GHX_LZ", c.tile_z, mode == 0 ? 8 : 4), max_lds());              \
    const int L = s.lines; const size_t lds = s.lds;                            \
    dim3 grid((g.ny + L - 1) / L, g.nxh, c.batch);                                          \
    if (mode == 0) {                                                                        \
      allow_lds(k_zinv<n_, 0>, lds);                                                        \
      GHX_LAUNCH_CHECK((k_zinv<n_, 0>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zinv<n_, 0>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, c.tw_z, prune, aux); \
    } else if (mode == 1) {                                                                 \
      allow_lds(k_zinv<n_, 1>, lds);                                                        \
      GHX_LAUNCH_CHECK((k_zinv<n_, 1>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zinv<n_, 1>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, c.tw_z, prune, aux); \
    } else if (mode == 2) {                                                                 \
      allow_lds(k_zinv<n_, 2>, lds);                                                        \
      GHX_LAUNCH_CHECK((k_zinv<n_, 2>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zinv<n_, 2>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, c.tw_z, prune, aux); \
    } else if (mode == 4) {                                                                 \
      allow_lds(k_zinv<n_, 4>, lds);                                                        \
      GHX_LAUNCH_CHECK((k_zinv<n_, 4>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zinv<n_, 4>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, c.tw_z, prune, aux); \
    } else if (mode == 5) {                                                                 \
      allow_lds(k_zinv<n_, 5>, lds);                                                        \
      GHX_LAUNCH_CHECK((k_zinv<n_, 5>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zinv<n_, 5>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, c.tw_z, prune, aux); \
    } else if (mode == 6) {                                                                 \
      allow_lds(k_zinv<n_, 6>, lds);                                                        \
      GHX_LAUNCH_CHECK((k_zinv<n_, 6>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zinv<n_, 6>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, c.tw_z, prune, aux); \
    } else {                                                                                \
      allow_lds(k_zinv<n_, 3>, lds);                                                        \
      GHX_LAUNCH_CHECK((k_zinv<n_, 3>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zinv<n_, 3>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, c.tw_z, prune, aux); \
    }                                                                                       \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#endif

#if GHX_PART == 3
hipError_t zfwd(const FFTCtx &c, int nf, const FieldSet &in, const FieldSet &out, int prune) {
  const GridK &g = c.gk;
  switch (g.nz) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = tile_shape(T, lds_line_stride(n_), g.ny, tile_want("GHX_LZ", 8), max_lds());                                        \
    const int L = s.lines; const size_t lds = s.lds;                            \
    dim3 grid((g.ny + L - 1) / L, g.nxh);                                                   \
    allow_lds(k_zfwd<n_>, lds);                                                             \
    GHX_LAUNCH_CHECK((k_zfwd<n_>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zfwd<n_>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, c.tw_z, prune); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#endif

#if GHX_PART == 4
hipError_t zfwd_rk_hd(const FFTCtx &c, const FieldSet &in, const RKParams &rk, int prune) {
  const GridK &g = c.gk;
  // 512-point lines take the four-points-per-lane schedule by default (128
  // lanes per line, four lines per 512-lane block): 124 registers per lane
  // against 224, two blocks per CU, so the update's loads of v and C overlap
  // more (0.71 -> 0.70 ms at 512^3; GHX_ZLOW=0 selects the default schedule).
  // Other sizes take it on request (FFTCtx::sched bit 0).
  const bool low = (c.sched & 1) || (g.nz == 512 && env_switch("GHX_ZLOW", 1) == 1);
  // On the four-points-per-lane schedule the next field's Q tile is loaded
  // while this one is transformed (138 registers against 124; 0.698 -> 0.671
  // ms at 512^3; GHX_ZPF=0 loads one tile at a time)
  const int pf = env_switch("GHX_ZPF", 1) != 0 ? 1 : 0;
  if (low) {
    switch (g.nz) {
#define GHX_LOW(n_)                                                                                     \
  case n_: {                                                                                            \
    using Cfg = FFTCfgLow<n_>;                                                                          \
    constexpr int T = Cfg::T;                                                                           \
    const PassShape s = tile_shape(T, lds_line_stride(n_), g.ny, tile_pick("GHX_LZ", c.tile_z, 4), max_lds()); \
    const int L = s.lines;                                                                              \
    dim3 grid((g.ny + L - 1) / L, g.nxh, c.batch);                                                      \
    if (pf) {                                                                                           \
      allow_lds(k_zfwd_rk_hd<n_, Cfg, 1>, s.lds);                                                       \
      GHX_LAUNCH_CHECK((k_zfwd_rk_hd<n_, Cfg, 1>), s, kTileBound, T, true);                             \
      hipLaunchKernelGGL((k_zfwd_rk_hd<n_, Cfg, 1>), grid, dim3(L * T), s.lds, c.stream, in, g, c.tw_z, rk, prune); \
    } else {                                                                                            \
      allow_lds(k_zfwd_rk_hd<n_, Cfg>, s.lds);                                                          \
      GHX_LAUNCH_CHECK((k_zfwd_rk_hd<n_, Cfg>), s, kTileBound, T, true);                                \
      hipLaunchKernelGGL((k_zfwd_rk_hd<n_, Cfg>), grid, dim3(L * T), s.lds, c.stream, in, g, c.tw_z, rk, prune); \
    }                                                                                                   \
    return hipGetLastError();                                                                           \
  }
      GHX_LOW(32) GHX_LOW(64) GHX_LOW(128) GHX_LOW(256) GHX_LOW(512)
#undef GHX_LOW
    default:
      break;  // the same schedule either way
    }
  }
  switch (g.nz) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = tile_shape(T, lds_line_stride(n_), g.ny, tile_pick("GHX_LZ", c.tile_z, 4), max_lds());         \
    const int L = s.lines; const size_t lds = s.lds;                            \
    dim3 grid((g.ny + L - 1) / L, g.nxh, c.batch);                                          \
    allow_lds(k_zfwd_rk_hd<n_>, lds);                                                       \
    GHX_LAUNCH_CHECK((k_zfwd_rk_hd<n_>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zfwd_rk_hd<n_>), grid, dim3(L * T), lds, c.stream, in, g, c.tw_z, rk, prune); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#endif

#if GHX_PART == 5
hipError_t zfwd_rk_mhd(const FFTCtx &c, const FieldSet &in, const MHDRKParams &rk, int prune) {
  const GridK &g = c.gk;
  // the four-points-per-lane schedule where a size has one (GHX_ZLOW=0: the
  // default schedule)
  const bool low = env_switch("GHX_ZLOW", 1) == 1;
  switch (g.nz) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    if constexpr (FFTCfgLow<n_>::E != FFTCfg<n_>::E) if (low) {                              \
      using Cl = FFTCfgLow<n_>;                                                              \
      constexpr int T = Cl::T;                                                               \
      const PassShape s = tile_shape(T, lds_line_stride(n_), g.ny, tile_want("GHX_LZ", 4), max_lds()); \
      const int L = s.lines; const size_t lds = s.lds;                                       \
      dim3 grid((g.ny + L - 1) / L, g.nxh);                                                  \
      allow_lds(k_zfwd_rk_mhd<n_, 0, Cl>, lds);                                              \
      allow_lds(k_zfwd_rk_mhd<n_, 1, Cl>, lds);                                              \
      GHX_LAUNCH_CHECK((k_zfwd_rk_mhd<n_, 0, Cl>), s, kTileBound, T, true);                  \
      hipLaunchKernelGGL((k_zfwd_rk_mhd<n_, 0, Cl>), grid, dim3(L * T), lds, c.stream, in, g, c.tw_z, rk, prune); \
      GHX_LAUNCH_CHECK((k_zfwd_rk_mhd<n_, 1, Cl>), s, kTileBound, T, true);                  \
      hipLaunchKernelGGL((k_zfwd_rk_mhd<n_, 1, Cl>), grid, dim3(L * T), lds, c.stream, in, g, c.tw_z, rk, prune); \
      break;                                                                                 \
    }                                                                                        \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = tile_shape(T, lds_line_stride(n_), g.ny, tile_want("GHX_LZ", 4), max_lds());                                        \
    const int L = s.lines; const size_t lds = s.lds;                            \
    dim3 grid((g.ny + L - 1) / L, g.nxh);                                                   \
    allow_lds(k_zfwd_rk_mhd<n_, 0>, lds);                                                   \
    allow_lds(k_zfwd_rk_mhd<n_, 1>, lds);                                                   \
    GHX_LAUNCH_CHECK((k_zfwd_rk_mhd<n_, 0>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zfwd_rk_mhd<n_, 0>), grid, dim3(L * T), lds, c.stream, in, g, c.tw_z, rk, prune); \
    GHX_LAUNCH_CHECK((k_zfwd_rk_mhd<n_, 1>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zfwd_rk_mhd<n_, 1>), grid, dim3(L * T), lds, c.stream, in, g, c.tw_z, rk, prune); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

#endif

#if GHX_PART == 1
// the y passes on the four-points-per-lane schedule (FFTCtx::sched bit 1)
template <int N>
static hipError_t launch_y_low(const FFTCtx &c, bool inv, int nf, const FieldSet &in, const FieldSet &out, int kxlim,
                               int prune) {
  const GridK &g = c.gk;
  using Cfg = FFTCfgLow<N>;
  constexpr int T = Cfg::T;
  const int nxp = p_row(g);
  const PassShape s = tile_shape(T, lds_line_stride(N), g.nxh, tile_pick("GHX_LY", c.tile_y, inv ? 8 : 4), max_lds());
  const int L = s.lines;
  dim3 grid((g.nxh + L - 1) / L, g.nz, c.batch);
  if (inv) {
    allow_lds(k_yinv<N, Cfg>, s.lds);
    GHX_LAUNCH_CHECK((k_yinv<N, Cfg>), s, kTileBound, T, true);
    hipLaunchKernelGGL((k_yinv<N, Cfg>), grid, dim3(L * T), s.lds, c.stream, in, out, nf, g, nxp, c.tw_y, kxlim, prune);
  } else {
    allow_lds(k_yfwd<N, Cfg>, s.lds);
    GHX_LAUNCH_CHECK((k_yfwd<N, Cfg>), s, kTileBound, T, true);
    hipLaunchKernelGGL((k_yfwd<N, Cfg>), grid, dim3(L * T), s.lds, c.stream, in, out, nf, g, nxp, c.tw_y, kxlim, prune);
  }
  return hipGetLastError();
}
static bool y_low(const FFTCtx &c, bool inv, int nf, const FieldSet &in, const FieldSet &out, int kxlim, int prune,
                  hipError_t *err) {
  if (!(c.sched & 2)) return false;
  switch (c.gk.ny) {
  case 32: *err = launch_y_low<32>(c, inv, nf, in, out, kxlim, prune); return true;
  case 64: *err = launch_y_low<64>(c, inv, nf, in, out, kxlim, prune); return true;
  case 128: *err = launch_y_low<128>(c, inv, nf, in, out, kxlim, prune); return true;
  case 256: *err = launch_y_low<256>(c, inv, nf, in, out, kxlim, prune); return true;
  default: return false;  // the same schedule either way, or the wide 512-point pass
  }
}

hipError_t yinv(const FFTCtx &c, int nf, const FieldSet &in, const FieldSet &out, int prune) {
  const GridK &g = c.gk;
  const int kxlim = prune ? active_kx(g) : g.nxh;
  const int nxp = p_row(g);
  hipError_t lerr;
  if (y_low(c, true, nf, in, out, kxlim, prune, &lerr)) return lerr;
  switch (g.ny) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    if constexpr (FFTCfgWide<n_>::E != FFTCfg<n_>::E)                                        \
    if (env_switch("GHX_YWIDE", 1) == 1 &&                                                   \
        wide_shape(FFTCfgWide<n_>::T, lds_line_stride(n_), g.nxh).lds <= max_lds()) {         \
      /* 16 points per lane: 8-line tiles from 256-lane blocks, two blocks per CU */       \
      constexpr int TW = FFTCfgWide<n_>::T;                                                  \
      const PassShape s = wide_shape(TW, lds_line_stride(n_), g.nxh);                        \
      const int L = s.lines; const size_t lds = s.lds;                                       \
      dim3 grid((g.nxh + L - 1) / L, g.nz, c.batch);                                         \
      GHX_LAUNCH_CHECK((k_yinv<n_, FFTCfgWide<n_>>), s, kTileBound, TW, true);               \
      allow_lds(k_yinv<n_, FFTCfgWide<n_>>, lds);                                            \
      hipLaunchKernelGGL((k_yinv<n_, FFTCfgWide<n_>>), grid, dim3(L * TW), lds, c.stream, in, out, nf, g, nxp, \
                         c.tw_y, kxlim, prune);                                              \
      break;                                                                                 \
    }                                                                                        \
    const PassShape s = tile_shape(T, lds_line_stride(n_), g.nxh, tile_pick("GHX_LY", c.tile_y, 8), max_lds());        \
    const int L = s.lines; const size_t lds = s.lds;                            \
    dim3 grid((g.nxh + L - 1) / L, g.nz, c.batch);                                          \
    allow_lds(k_yinv<n_>, lds);                                                             \
    GHX_LAUNCH_CHECK((k_yinv<n_>), s, kTileBound, T, true); hipLaunchKernelGGL((k_yinv<n_>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, nxp, c.tw_y, kxlim, prune); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t yfwd(const FFTCtx &c, int nf, const FieldSet &in, const FieldSet &out, int prune) {
  const GridK &g = c.gk;
  const int kxlim = prune ? active_kx(g) : g.nxh;
  const int nxp = p_row(g);
  hipError_t lerr;
  if (y_low(c, false, nf, in, out, kxlim, prune, &lerr)) return lerr;
  switch (g.ny) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = tile_shape(T, lds_line_stride(n_), g.nxh, tile_pick("GHX_LY", c.tile_y, 4), max_lds());        \
    const int L = s.lines; const size_t lds = s.lds;                            \
    dim3 grid((g.nxh + L - 1) / L, g.nz, c.batch);                                          \
    allow_lds(k_yfwd<n_>, lds);                                                             \
    GHX_LAUNCH_CHECK((k_yfwd<n_>), s, kTileBound, T, true); hipLaunchKernelGGL((k_yfwd<n_>), grid, dim3(L * T), lds, c.stream, in, out, nf, g, nxp, c.tw_y, kxlim, prune); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
#endif

#if GHX_PART == 6

hipError_t xinv(const FFTCtx &c, int nf, const FieldSet &in, double *const *out, int prune) {
  const GridK &g = c.gk;
  if (g.nx % 2) return hipErrorInvalidValue;
  const int kxlim = prune ? active_kx(g) : g.nxh;
  RealSet rs;
  for (int f = 0; f < nf; ++f) rs.p[f] = out[f];
  const long nrows = (long)g.nz * g.ny;
  const int nxp = p_row(g);
  if (use_x2(g.nx)) {
    switch (g.nx) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    const dim3 gr((unsigned)((nrows + rpb - 1) / rpb)), bl(rpb * T);                      \
    GHX_LAUNCH_CHECK((k_xinv2<n_>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xinv2<n_>), gr, bl, lds, c.stream, in, rs, nf, g, nxp, c.tw_x, kxlim);                                                                 \
    return hipGetLastError();                                                               \
  }
      GHX_X2_SIZES(GHX_CASE)
#undef GHX_CASE
      default:
        break;
    }
  }
  switch (g.nx / 2) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfgLow<n_>::T;                                                     \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    GHX_LAUNCH_CHECK((k_xinv<n_>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xinv<n_>), dim3((unsigned)((nrows + rpb - 1) / rpb)), dim3(rpb * T), lds, c.stream, in, rs, nf, g, nxp, c.tw_xh, c.tw_x, kxlim); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t xinv3i(const FFTCtx &c, const FieldSet &in, double *out, int prune) {
  const GridK &g = c.gk;
  if (g.nx % 2) return hipErrorInvalidValue;
  const int kxlim = prune ? active_kx(g) : g.nxh;
  const long nrows = (long)g.nz * g.ny;
  const int nxp = p_row(g);
  if (use_x2(g.nx)) {
    switch (g.nx) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    const dim3 gr((unsigned)((nrows + rpb - 1) / rpb)), bl(rpb * T);                      \
    GHX_LAUNCH_CHECK((k_xinv3i2<n_>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xinv3i2<n_>), gr, bl, lds, c.stream, in, out, g, nxp, c.tw_x, kxlim);                                                                 \
    return hipGetLastError();                                                               \
  }
      GHX_X2_SIZES(GHX_CASE)
#undef GHX_CASE
      default:
        break;
    }
  }
  switch (g.nx / 2) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfgLow<n_>::T;                                                     \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    GHX_LAUNCH_CHECK((k_xinv3i<n_>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xinv3i<n_>), dim3((unsigned)((nrows + rpb - 1) / rpb)), dim3(rpb * T), lds, c.stream, in, out, g, nxp, c.tw_xh, c.tw_x, kxlim); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t xfwd(const FFTCtx &c, int nf, const double *const *in, const FieldSet &out) {
  const GridK &g = c.gk;
  if (g.nx % 2) return hipErrorInvalidValue;
  RealSet rs;
  for (int f = 0; f < nf; ++f) rs.p[f] = const_cast<double *>(in[f]);
  const long nrows = (long)g.nz * g.ny;
  const int nxp = p_row(g);
  if (use_x2(g.nx)) {
    switch (g.nx) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    const dim3 gr((unsigned)((nrows + rpb - 1) / rpb)), bl(rpb * T);                      \
    GHX_LAUNCH_CHECK((k_xfwd2<n_>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xfwd2<n_>), gr, bl, lds, c.stream, rs, out, nf, g, nxp, c.tw_x);                                                                 \
    return hipGetLastError();                                                               \
  }
      GHX_X2_SIZES(GHX_CASE)
#undef GHX_CASE
      default:
        break;
    }
  }
  switch (g.nx / 2) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfgLow<n_>::T;                                                     \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    GHX_LAUNCH_CHECK((k_xfwd<n_>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xfwd<n_>), dim3((unsigned)((nrows + rpb - 1) / rpb)), dim3(rpb * T), lds, c.stream, rs, out, nf, g, nxp, c.tw_xh, c.tw_x); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
#endif

#if GHX_PART == 7
hipError_t xcross(const FFTCtx &c, int mode, const FieldSet &in, const FieldSet &out, int prune) {
  (void)mode;
  const GridK &g = c.gk;
  if (g.nx % 2) return hipErrorInvalidValue;
  const int kxlim = prune ? active_kx(g) : g.nxh;
  const long nrows = (long)g.nz * g.ny;
  const int nxp = p_row(g);
  if (use_x2(g.nx)) {
    switch (g.nx) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    const dim3 gr((unsigned)((nrows + rpb - 1) / rpb), c.batch), bl(rpb * T);                      \
    GHX_LAUNCH_CHECK((k_xcross2<n_>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xcross2<n_>), gr, bl, lds, c.stream, in, out, g, nxp, c.tw_x, kxlim);                                                                 \
    return hipGetLastError();                                                               \
  }
      GHX_X2_SIZES(GHX_CASE)
#undef GHX_CASE
      default:
        break;
    }
  }
  switch (g.nx / 2) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfgLow<n_>::T;                                                     \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    const dim3 gr((unsigned)((nrows + rpb - 1) / rpb), c.batch), bl(rpb * T);               \
    switch (tile_want("GHX_XV", 1)) {                                                       \
    case 4: GHX_LAUNCH_CHECK((k_xcross<n_, 2, 1>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xcross<n_, 2, 1>), gr, bl, lds, c.stream, in, out, g, nxp, c.tw_xh, c.tw_x, kxlim); break; \
    /* 3 * 2^k rows: one wave per SIMD leaves the registers to hold the six */ \
    /* fields without scratch (7.5% faster on the ensemble's 48^3 and 96^3) */ \
    default: GHX_LAUNCH_CHECK((k_xcross<n_, (n_ % 3 == 0) ? 1 : 2, 0>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xcross<n_, (n_ % 3 == 0) ? 1 : 2, 0>), gr, bl, lds, c.stream, in, out, g, nxp, c.tw_xh, c.tw_x, kxlim); break; \
    }                                                                                       \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
#endif

}  // namespace ghx

#if GHX_PART == 8
namespace ghx {
hipError_t xdot(const FFTCtx &c, const FieldSet &in, const FieldSet &out, int prune) {
  const GridK &g = c.gk;
  if (g.nx % 2) return hipErrorInvalidValue;
  const int kxlim = prune ? active_kx(g) : g.nxh;
  const long nrows = (long)g.nz * g.ny;
  const int nxp = p_row(g);
  if (use_x2(g.nx)) {
    switch (g.nx) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    const dim3 gr((unsigned)((nrows + rpb - 1) / rpb)), bl(rpb * T);                      \
    GHX_LAUNCH_CHECK((k_xdot2<n_>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xdot2<n_>), gr, bl, lds, c.stream, in, out, g, nxp, c.tw_x, kxlim);                                                                 \
    return hipGetLastError();                                                               \
  }
      GHX_X2_SIZES(GHX_CASE)
#undef GHX_CASE
      default:
        break;
    }
  }
  switch (g.nx / 2) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfgLow<n_>::T;                                                     \
    const PassShape s = row_shape(T, lds_line_stride(n_), max_lds());                                                      \
    const int rpb = s.lines; const size_t lds = s.lds;                          \
    GHX_LAUNCH_CHECK((k_xdot<n_>), s, kRowBound, T, false); hipLaunchKernelGGL((k_xdot<n_>), dim3((unsigned)((nrows + rpb - 1) / rpb)), dim3(rpb * T), lds, c.stream, in, out, g, nxp, c.tw_xh, c.tw_x, kxlim); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t zfwd_rk_bouss(const FFTCtx &c, const FieldSet &in, const BoussRKParams &rk, int prune) {
  const GridK &g = c.gk;
  switch (g.nz) {
#define GHX_CASE(n_)                                                                        \
  case n_: {                                                                                \
    constexpr int T = FFTCfg<n_>::T;                                                        \
    const PassShape s = tile_shape(T, lds_line_stride(n_) + n_, g.ny, tile_want("GHX_LZ", 4), max_lds());              \
    const int L = s.lines; const size_t lds = s.lds;                     \
    dim3 grid((g.ny + L - 1) / L, g.nxh);                                                   \
    allow_lds(k_zfwd_rk_bouss<n_>, lds);                                                    \
    GHX_LAUNCH_CHECK((k_zfwd_rk_bouss<n_>), s, kTileBound, T, true); hipLaunchKernelGGL((k_zfwd_rk_bouss<n_>), grid, dim3(L * T), lds, c.stream, in, g, c.tw_z, rk, prune); \
    break;                                                                                  \
  }
    GHX_SIZES(GHX_CASE)
#undef GHX_CASE
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
}  // namespace ghx
#endif
