// fft_engine.h -- register-resident Stockham line FFT for gfx950 (FP64 complex).
//
// A line of N complex points is held by T = N/E lanes, E points per lane:
// lane t holds elements t + T*m (m = 0..E-1).  Every Stockham stage reads
// exactly that set, so a stage is: E/R radix-R butterflies in registers, then
// an exchange through a padded LDS line buffer.  The last stage needs no
// exchange: its outputs land back in the same register slots in natural order.
// For N = 512 (E = 8, radix 8^3) a line is one 64-wide wavefront.
//
// Transform convention (FFTW's, which GHOST uses): DIR = -1 forward,
// DIR = +1 inverse, both unnormalised.
#pragma once
#include "ghx_common.h"

// LDS padding: one spare slot every 8 complex points keeps the stride-R
// butterfly writes of the first stages on distinct banks.
GHX_DEV int lds_pad(int i) { return i + (i >> 3); }
__host__ __device__ constexpr int lds_line_stride(int n) { return n + (n >> 3) + 1; }

// ---------------------------------------------------------------------------
// Per-size schedule: E points per lane and the radix of each stage.
// ---------------------------------------------------------------------------
template <int N> struct FFTCfg;
#define GHX_FFTCFG(N_, E_, NS_, ...)                                    \
  template <> struct FFTCfg<N_> {                                       \
    static constexpr int E = E_;                                        \
    static constexpr int NS = NS_;                                      \
    static constexpr int T = N_ / E_;                                   \
    GHX_DEV static constexpr int radix(int s) {                         \
      constexpr int r[] = {__VA_ARGS__};                                \
      return r[s];                                                      \
    }                                                                   \
  };
GHX_FFTCFG(1, 1, 0, 1)
GHX_FFTCFG(2, 2, 1, 2)
GHX_FFTCFG(3, 3, 1, 3)
GHX_FFTCFG(4, 4, 1, 4)
GHX_FFTCFG(5, 5, 1, 5)
GHX_FFTCFG(6, 6, 2, 3, 2)
GHX_FFTCFG(8, 8, 1, 8)
GHX_FFTCFG(10, 10, 2, 5, 2)
GHX_FFTCFG(12, 12, 2, 3, 4)
GHX_FFTCFG(16, 8, 2, 2, 8)
GHX_FFTCFG(20, 20, 2, 5, 4)
GHX_FFTCFG(24, 6, 4, 3, 2, 2, 2)
GHX_FFTCFG(32, 8, 2, 4, 8)
GHX_FFTCFG(40, 40, 2, 5, 8)
GHX_FFTCFG(48, 6, 5, 3, 2, 2, 2, 2)
GHX_FFTCFG(64, 8, 2, 8, 8)
GHX_FFTCFG(80, 40, 3, 5, 2, 8)
GHX_FFTCFG(96, 6, 6, 3, 2, 2, 2, 2, 2)
GHX_FFTCFG(128, 8, 3, 2, 8, 8)
GHX_FFTCFG(160, 40, 3, 5, 4, 8)
GHX_FFTCFG(192, 6, 7, 3, 2, 2, 2, 2, 2, 2)
GHX_FFTCFG(256, 8, 3, 4, 8, 8)
GHX_FFTCFG(384, 6, 8, 3, 2, 2, 2, 2, 2, 2, 2)
GHX_FFTCFG(512, 8, 3, 8, 8, 8)
GHX_FFTCFG(768, 24, 4, 3, 4, 8, 8)
GHX_FFTCFG(1024, 8, 4, 2, 8, 8, 8)

// Low-register schedules (4 points per lane) for kernels that keep several
// fields live at once (the fused x pass holds six); other sizes fall back to
// the default schedule.
template <int N> struct FFTCfgLow : FFTCfg<N> {};
#define GHX_FFTCFGLOW(N_, E_, NS_, ...)                                 \
  template <> struct FFTCfgLow<N_> {                                    \
    static constexpr int E = E_;                                        \
    static constexpr int NS = NS_;                                      \
    static constexpr int T = N_ / E_;                                   \
    GHX_DEV static constexpr int radix(int s) {                         \
      constexpr int r[] = {__VA_ARGS__};                                \
      return r[s];                                                      \
    }                                                                   \
  };
GHX_FFTCFGLOW(32, 4, 3, 2, 4, 4)
GHX_FFTCFGLOW(64, 4, 3, 4, 4, 4)
GHX_FFTCFGLOW(128, 4, 4, 2, 4, 4, 4)
GHX_FFTCFGLOW(256, 4, 4, 4, 4, 4, 4)
GHX_FFTCFGLOW(512, 4, 5, 2, 4, 4, 4, 4)
#undef GHX_FFTCFGLOW

// Wide schedules (16 points per lane, half the lanes per line) for passes
// whose blocks should hold more lines with fewer lanes.
template <int N> struct FFTCfgWide : FFTCfg<N> {};
template <> struct FFTCfgWide<512> {
  static constexpr int E = 16;
  static constexpr int NS = 3;
  static constexpr int T = 512 / 16;
  GHX_DEV static constexpr int radix(int) { return 8; }
};
#undef GHX_FFTCFG

// ---------------------------------------------------------------------------
// Small DFTs in registers, sign DIR.
// ---------------------------------------------------------------------------
template <int R, int DIR> struct SmallDFT;

template <int DIR> struct SmallDFT<1, DIR> {
  GHX_DEV static void run(cplx *) {}
};
template <int DIR> struct SmallDFT<2, DIR> {
  GHX_DEV static void run(cplx *v) {
    cplx a = c_add(v[0], v[1]), b = c_sub(v[0], v[1]);
    v[0] = a; v[1] = b;
  }
};
template <int DIR> struct SmallDFT<3, DIR> {
  GHX_DEV static void run(cplx *v) {
    const double s3 = 0.86602540378443864676;
    cplx t = c_add(v[1], v[2]), d = c_sub(v[1], v[2]);
    cplx m = c_sub(v[0], c_scale(t, 0.5));
    cplx e = c_mul_si<DIR>(c_scale(d, s3));
    v[0] = c_add(v[0], t);
    v[1] = c_add(m, e);
    v[2] = c_sub(m, e);
  }
};
template <int DIR> struct SmallDFT<4, DIR> {
  GHX_DEV static void run(cplx *v) {
    cplx a0 = c_add(v[0], v[2]), a1 = c_sub(v[0], v[2]);
    cplx a2 = c_add(v[1], v[3]), a3 = c_mul_si<DIR>(c_sub(v[1], v[3]));
    v[0] = c_add(a0, a2); v[2] = c_sub(a0, a2);
    v[1] = c_add(a1, a3); v[3] = c_sub(a1, a3);
  }
};
template <int DIR> struct SmallDFT<5, DIR> {
  GHX_DEV static void run(cplx *v) {
    // Winograd-style radix 5 with c1 = cos(2pi/5), c2 = cos(4pi/5), s1, s2 sines.
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    cplx t1 = c_add(v[1], v[4]), t2 = c_add(v[2], v[3]);
    cplx t3 = c_sub(v[1], v[4]), t4 = c_sub(v[2], v[3]);
    cplx a1 = c_add(v[0], c_add(c_scale(t1, c1), c_scale(t2, c2)));
    cplx a2 = c_add(v[0], c_add(c_scale(t1, c2), c_scale(t2, c1)));
    cplx b1 = c_mul_si<DIR>(c_add(c_scale(t3, s1), c_scale(t4, s2)));
    cplx b2 = c_mul_si<DIR>(c_sub(c_scale(t3, s2), c_scale(t4, s1)));
    v[0] = c_add(v[0], c_add(t1, t2));
    v[1] = c_add(a1, b1); v[4] = c_sub(a1, b1);
    v[2] = c_add(a2, b2); v[3] = c_sub(a2, b2);
  }
};
template <int DIR> struct SmallDFT<8, DIR> {
  GHX_DEV static void run(cplx *v) {
    const double r2 = 0.70710678118654752440;
    cplx e[4] = {v[0], v[2], v[4], v[6]};
    cplx o[4] = {v[1], v[3], v[5], v[7]};
    SmallDFT<4, DIR>::run(e);
    SmallDFT<4, DIR>::run(o);
    // W8 = exp(DIR*2 pi i/8) = (1 + DIR i)/sqrt 2
    cplx o1 = c_scale(c_add(o[1], c_mul_si<DIR>(o[1])), r2);
    cplx o2 = c_mul_si<DIR>(o[2]);
    cplx o3 = c_mul_si<DIR>(c_scale(c_add(o[3], c_mul_si<DIR>(o[3])), r2));
    v[0] = c_add(e[0], o[0]); v[4] = c_sub(e[0], o[0]);
    v[1] = c_add(e[1], o1);   v[5] = c_sub(e[1], o1);
    v[2] = c_add(e[2], o2);   v[6] = c_sub(e[2], o2);
    v[3] = c_add(e[3], o3);   v[7] = c_sub(e[3], o3);
  }
};

// ---------------------------------------------------------------------------
// Twiddle powers w^1..w^(R-1) from one table load by a product tree of depth
// <= 3 (rounding error of a few ulp), instead of R-1 table loads.
// ---------------------------------------------------------------------------
template <int R> GHX_DEV void tw_powers(cplx w1, cplx *w) {
  w[1] = w1;
  if constexpr (R > 2) w[2] = c_mul(w1, w1);
  if constexpr (R > 3) w[3] = c_mul(w[2], w1);
  if constexpr (R > 4) w[4] = c_mul(w[2], w[2]);
  if constexpr (R > 5) w[5] = c_mul(w[4], w1);
  if constexpr (R > 6) w[6] = c_mul(w[3], w[3]);
  if constexpr (R > 7) w[7] = c_mul(w[4], w[3]);
}

// ---------------------------------------------------------------------------
// Synchronisation of one line's LDS exchange.  A line of T <= 64 lanes lives
// inside one wavefront (T is a power of two and blocks are whole lines), so
// ordering the wave's own LDS traffic is enough and the other waves of the
// block run on; longer lines need the block barrier.
// ---------------------------------------------------------------------------
template <int T> GHX_DEV void line_sync() {
  if constexpr (T <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Twiddle sources for the stages.  TwTable reads the root of each butterfly
// from the forward table when the stage runs; TwCache holds the roots of one
// lane for every stage of one schedule in registers, loaded once, for kernels
// that run many transforms of the same length per lane.
// ---------------------------------------------------------------------------
struct TwTable {
  const cplx *p;
  template <int S> GHX_DEV cplx root(int, int idx) const { return p[idx]; }
};

template <int N, typename Cfg = FFTCfg<N>> struct TwCache {
  static constexpr int NS = Cfg::NS, E = Cfg::E, T = Cfg::T;
  cplx w[NS > 0 ? NS : 1][E];
  template <int S, int Ns> GHX_DEV void fill(const cplx *tw, int t) {
    if constexpr (S < NS) {
      constexpr int R = Cfg::radix(S), Q = E / R;
      if constexpr (Ns > 1) {
        constexpr int tws = N / (Ns * R);
#pragma unroll
        for (int q = 0; q < Q; ++q) w[S][q] = tw[((t + q * T) % Ns) * tws];
      }
      fill<S + 1, Ns * R>(tw, t);
    }
  }
  GHX_DEV void load(const cplx *tw, int t) { fill<0, 1>(tw, t); }
  template <int S> GHX_DEV cplx root(int q, int) const { return w[S][q]; }
};

// ---------------------------------------------------------------------------
// The line transform.
// ---------------------------------------------------------------------------
template <int N, int DIR, typename Cfg = FFTCfg<N>> struct LineFFT {
  using C = Cfg;
  static constexpr int E = C::E;
  static constexpr int T = C::T;

  // One Stockham stage (radix R, Ns = product of earlier radices).  If LAST,
  // the outputs stay in v; otherwise they are written to the LDS line buffer
  // and v is reloaded from it.
  template <int S, int NsS, typename TW>
  GHX_DEV static void stages(cplx (&v)[E], cplx *buf, int t, const TW &tw) {
    if constexpr (S < C::NS) {
      constexpr int R = C::radix(S);
      constexpr int Q = E / R;
      constexpr int Ns = NsS;
      constexpr bool LAST = (S == C::NS - 1);
      static_assert(E % R == 0, "radix must divide points per lane");
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int j = t + q * T;
        cplx u[R];
#pragma unroll
        for (int r = 0; r < R; ++r) u[r] = v[q + r * Q];
        if constexpr (Ns > 1) {
          const int k = j % Ns;
          constexpr int tws = N / (Ns * R);
          cplx w[R];
          const cplx w1 = tw.template root<S>(q, k * tws);
          tw_powers<R>(DIR < 0 ? w1 : c_conj(w1), w);
#pragma unroll
          for (int r = 1; r < R; ++r) u[r] = c_mul(u[r], w[r]);
        }
        SmallDFT<R, DIR>::run(u);
        if constexpr (LAST) {
#pragma unroll
          for (int r = 0; r < R; ++r) v[q + r * Q] = u[r];
        } else {
          const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
          for (int r = 0; r < R; ++r) buf[lds_pad(base + r * Ns)] = u[r];
        }
      }
      if constexpr (!LAST) {
        line_sync<T>();
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = buf[lds_pad(t + T * m)];
        line_sync<T>();
        stages<S + 1, Ns * R, TW>(v, buf, t, tw);
      }
    }
  }

  // v holds x[t + T m] on entry and X[t + T m] on exit.  buf is this line's
  // LDS buffer (lds_line_stride(N) slots), used by this line only.  Every
  // thread of the block must call this (lines longer than a wavefront use
  // block barriers).
  GHX_DEV static void run(cplx (&v)[E], cplx *buf, int t, const cplx *tw) {
    stages<0, 1, TwTable>(v, buf, t, TwTable{tw});
  }
  // the same with the stage roots held in registers (TwCache<N, Cfg> of the
  // forward table, loaded by this lane)
  GHX_DEV static void run(cplx (&v)[E], cplx *buf, int t, const TwCache<N, Cfg> &tw) {
    stages<0, 1, TwCache<N, Cfg>>(v, buf, t, tw);
  }
};
