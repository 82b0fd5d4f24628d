// api.hip -- C ABI (include/ghx.h) over the gfx950 kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/ghx.h"
#include "fft_passes.h"
#include "pseudo.h"
#include "tracers.h"

using ghx::FieldSet;
using ghx::FFTCtx;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static int set_err(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return set_err(e_ == hipErrorOutOfMemory ? GHX_ERR_NOMEM : GHX_ERR_HIP, "%s: %s (%s:%d)", #expr, \
                     hipGetErrorString(e_), __FILE__, __LINE__);                              \
  } while (0)

#define ARGCHK(cond, msg)                               \
  do {                                                  \
    if (!(cond)) return set_err(GHX_ERR_ARG, "%s", msg); \
  } while (0)

extern "C" const char *ghx_last_error(void) { return g_err; }

// for the other translation units of the library (binio.hip)
int ghx_internal_set_err(int code, const char *msg) { return set_err(code, "%s", msg); }

extern "C" int ghx_device_ok(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0;
}

// ---------------------------------------------------------------------------
// memory
// ---------------------------------------------------------------------------
extern "C" int ghx_malloc(void **ptr, size_t bytes) {
  ARGCHK(ptr, "null pointer");
  HIPCHK(hipMalloc(ptr, bytes ? bytes : 16));
  return GHX_OK;
}
extern "C" int ghx_free(void *ptr) {
  if (ptr) HIPCHK(hipFree(ptr));
  return GHX_OK;
}
extern "C" int ghx_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GHX_OK;
}
extern "C" int ghx_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GHX_OK;
}
extern "C" int ghx_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return GHX_OK;
}
extern "C" int ghx_memset_zero(void *ptr, size_t bytes) {
  HIPCHK(ghx::op_zero(0, ptr, bytes));
  HIPCHK(hipDeviceSynchronize());
  return GHX_OK;
}
extern "C" int ghx_sync(void) {
  HIPCHK(hipDeviceSynchronize());
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// pass profiler: HIP events around each pass launch on its own stream
// ---------------------------------------------------------------------------
namespace {
enum ProfSlot {
  P_ZINV, P_ZINV_CURL, P_ZINV_MHD, P_YINV, P_YFWD, P_XINV, P_XFWD, P_XCROSS, P_ZFWD, P_ZFWD_RK_HD, P_ZFWD_RK_MHD,
  P_ZINV_SPLINE, P_TRACER_STAGE, P_ZINV_BOUSS, P_XDOT, P_ZFWD_RK_BOUSS, P_YINV_TRACER, P_XINV_TRACER, P_NSLOTS
};
const char *kProfNames[P_NSLOTS] = {"zinv", "zinv_curl", "zinv_mhd", "ypass_inv", "ypass_fwd", "xinv",
                                    "xfwd", "xcross", "zfwd", "zfwd_rk_hd", "zfwd_rk_mhd", "zinv_spline",
                                    "tracer_stage", "zinv_bouss", "xdot", "zfwd_rk_bouss", "ypass_inv_tracer",
                                    "xinv_tracer"};
struct ProfRec {
  int slot;
  double bytes;
  hipEvent_t a, b;
};
struct Prof {
  bool on = false;
  std::vector<ProfRec> pending;
  std::vector<hipEvent_t> pool;
  int count[P_NSLOTS] = {};
  double ms[P_NSLOTS] = {};
  double bytes[P_NSLOTS] = {};
} g_prof;

hipEvent_t prof_event() {
  if (!g_prof.pool.empty()) {
    hipEvent_t e = g_prof.pool.back();
    g_prof.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void prof_collect() {
  for (ProfRec &r : g_prof.pending) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      g_prof.count[r.slot] += 1;
      g_prof.ms[r.slot] += ms;
      g_prof.bytes[r.slot] += r.bytes;
    }
    g_prof.pool.push_back(r.a);
    g_prof.pool.push_back(r.b);
  }
  g_prof.pending.clear();
}
}  // namespace

// Wraps one pass launch: expr must evaluate to hipError_t.  bytes is the
// pass's algorithmic traffic (each field it touches read or written once).
#define PROF(slot_, nbytes_, stream_, expr_)                              \
  do {                                                                    \
    if (g_prof.on) {                                                      \
      ProfRec r_;                                                         \
      r_.slot = (slot_);                                                  \
      r_.bytes = (double)(nbytes_);                                       \
      r_.a = prof_event();                                                \
      r_.b = prof_event();                                                \
      (void)hipEventRecord(r_.a, (stream_));                              \
      HIPCHK(expr_);                                                      \
      (void)hipEventRecord(r_.b, (stream_));                              \
      g_prof.pending.push_back(r_);                                       \
      if (g_prof.pending.size() > 4096) prof_collect();                   \
    } else {                                                              \
      HIPCHK(expr_);                                                      \
    }                                                                     \
  } while (0)

extern "C" int ghx_prof_enable(int on) {
  prof_collect();
  g_prof.on = on != 0;
  return GHX_OK;
}
extern "C" int ghx_prof_reset(void) {
  prof_collect();
  for (int k = 0; k < P_NSLOTS; ++k) { g_prof.count[k] = 0; g_prof.ms[k] = 0; g_prof.bytes[k] = 0; }
  return GHX_OK;
}
extern "C" int ghx_prof_read(int slot, const char **name, int *count, double *total_ms, double *total_bytes) {
  ARGCHK(slot >= 0, "bad slot");
  if (slot >= P_NSLOTS) return set_err(GHX_ERR_ARG, "slot out of range (%d slots)", (int)P_NSLOTS);
  prof_collect();
  if (name) *name = kProfNames[slot];
  if (count) *count = g_prof.count[slot];
  if (total_ms) *total_ms = g_prof.ms[slot];
  if (total_bytes) *total_bytes = g_prof.bytes[slot];
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct ghx_plan {
  int nx, ny, nz, nxh, dir;
  FFTCtx ctx;
  cplx *tw = nullptr;       // all twiddle tables
  cplx *q = nullptr;        // Q scratch, q_fields fields
  int q_fields = 0;
  bool timing = false;
  double t_fft = 0.0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool ev_pending = false;
};

static void fill_tw(std::vector<cplx> &h, size_t off, int n) {
  const long double twopi = 6.283185307179586476925286766559L;
  for (int k = 0; k < n; ++k) {
    long double a = twopi * (long double)k / (long double)n;
    h[off + k].x = (double)cosl(a);
    h[off + k].y = -(double)sinl(a);
  }
}

static size_t nc_of(int nx, int ny, int nz) { return (size_t)(nx / 2 + 1) * ny * nz; }

static GridK make_gridk(int nx, int ny, int nz, int arbsize, double Lx, double Ly, double Lz) {
  GridK g;
  g.nx = nx; g.ny = ny; g.nz = nz; g.nxh = nx / 2 + 1;
  g.anis = (arbsize || nx != ny || ny != nz) ? 1 : 0;
  g.dkx = arbsize ? 1.0 / Lx : 1.0;
  g.dky = arbsize ? 1.0 / Ly : 1.0;
  g.dkz = arbsize ? 1.0 / Lz : 1.0;
  g.rmp = g.rmq = g.rms = 1.0;
  if (g.anis) {
    g.rmp = 1.0 / ((double)nx * (double)nx);
    g.rmq = 1.0 / ((double)ny * (double)ny);
    g.rms = 1.0 / ((double)nz * (double)nz);
  }
  g.kmax = 1.0 / 9.0;
  if (!arbsize && !g.anis) g.kmax = g.kmax * ((double)nx * (double)nx);
  double t = (double)nx * (double)ny * (double)nz;
  g.inv_n6 = 1.0 / (t * t);
  g.tiny = 0.0;
  return g;
}

static int plan_init(ghx_plan *p, int nx, int ny, int nz, int dir) {
  if (nx < 2 || nx % 2 || ny < 2 || nz < 2)
    return set_err(GHX_ERR_SIZE, "grid %dx%dx%d: nx must be even and all sizes >= 2", nx, ny, nz);
  if (!ghx::fft_size_supported(nx / 2) || !ghx::fft_size_supported(ny) || !ghx::fft_size_supported(nz))
    return set_err(GHX_ERR_SIZE, "grid %dx%dx%d has no compiled FFT schedule (supported line lengths: "
                   "2,3,4,5,6,8,10,12,16,24,32,48,64,96,128,192,256,384,512,1024; x uses nx/2)", nx, ny, nz);
  if (!ghx_device_ok()) return set_err(GHX_ERR_NODEVICE, "no GPU visible to HIP");
  p->nx = nx; p->ny = ny; p->nz = nz; p->nxh = nx / 2 + 1; p->dir = dir;
  std::vector<cplx> h((size_t)nx / 2 + nx + ny + nz);
  fill_tw(h, 0, nx / 2);
  fill_tw(h, nx / 2, nx);
  fill_tw(h, nx / 2 + nx, ny);
  fill_tw(h, nx / 2 + nx + ny, nz);
  HIPCHK(hipMalloc(&p->tw, h.size() * sizeof(cplx)));
  HIPCHK(hipMemcpy(p->tw, h.data(), h.size() * sizeof(cplx), hipMemcpyHostToDevice));
  p->ctx.gk = make_gridk(nx, ny, nz, 0, 1, 1, 1);
  p->ctx.tw_xh = p->tw;
  p->ctx.tw_x = p->tw + nx / 2;
  p->ctx.tw_y = p->tw + nx / 2 + nx;
  p->ctx.tw_z = p->tw + nx / 2 + nx + ny;
  p->ctx.stream = 0;
  return GHX_OK;
}

// nf scratch slots, each large enough for a field in the Q or P layout
static int plan_scratch(ghx_plan *p, int nf) {
  if (p->q_fields >= nf) return GHX_OK;
  if (p->q) HIPCHK(hipFree(p->q));
  p->q = nullptr;
  p->q_fields = 0;
  HIPCHK(hipMalloc(&p->q, (size_t)nf * ghx::slot_points(p->ctx.gk) * sizeof(cplx)));
  p->q_fields = nf;
  return GHX_OK;
}

static void plan_release(ghx_plan *p) {
  if (!p) return;
  if (p->q) (void)hipFree(p->q);
  if (p->tw) (void)hipFree(p->tw);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
}

static void timer_collect(ghx_plan *p) {
  if (p->ev_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p->ev1) == hipSuccess && hipEventElapsedTime(&ms, p->ev0, p->ev1) == hipSuccess)
      p->t_fft += ms * 1e-3;
    p->ev_pending = false;
  }
}
static void timer_begin(ghx_plan *p) {
  if (!p->timing) return;
  timer_collect(p);
  (void)hipEventRecord(p->ev0, p->ctx.stream);
}
static void timer_end(ghx_plan *p) {
  if (!p->timing) return;
  (void)hipEventRecord(p->ev1, p->ctx.stream);
  p->ev_pending = true;
}

extern "C" int ghx_plan_create(ghx_plan **plan, int nx, int ny, int nz, int fftdir, int flags) {
  (void)flags;
  ARGCHK(plan, "null plan pointer");
  ARGCHK(fftdir == GHX_FFTW_REAL_TO_COMPLEX || fftdir == GHX_FFTW_COMPLEX_TO_REAL, "fftdir must be -1 or +1");
  ghx_plan *p = new (std::nothrow) ghx_plan();
  if (!p) return set_err(GHX_ERR_NOMEM, "host allocation failed");
  int rc = plan_init(p, nx, ny, nz, fftdir);
  if (rc) { plan_release(p); delete p; return rc; }
  *plan = p;
  return GHX_OK;
}

extern "C" int ghx_plan_destroy(ghx_plan *plan) {
  if (!plan) return GHX_OK;
  plan_release(plan);
  delete plan;
  return GHX_OK;
}

extern "C" int ghx_plan_set_stream(ghx_plan *plan, void *stream) {
  ARGCHK(plan, "null plan");
  plan->ctx.stream = (hipStream_t)stream;
  return GHX_OK;
}

// prune: the caller reads only the kept modes of `out` (the dealias set), so
// the y and z passes skip the rest (those modes of `out` keep what they held)
static int do_r2c(ghx_plan *p, int nf, const double *const *in, void *const *out, int prune = 0) {
  ARGCHK(p && in && out && nf >= 1 && nf <= ghx::kMaxFields, "bad r2c arguments");
  int rc = plan_scratch(p, 2 * nf);
  if (rc) return rc;
  const size_t nc = nc_of(p->nx, p->ny, p->nz), slot = ghx::slot_points(p->ctx.gk);
  FieldSet pp, q, o;
  for (int f = 0; f < nf; ++f) {
    ARGCHK(in[f] && out[f], "null field pointer");
    pp.p[f] = p->q + f * slot;
    q.p[f] = p->q + (nf + f) * slot;
    o.p[f] = (cplx *)out[f];
  }
  const double sc = (double)nc * sizeof(cplx), sr = (double)p->nx * p->ny * p->nz * sizeof(double);
  timer_begin(p);
  PROF(P_XFWD, nf * (sr + sc), p->ctx.stream, ghx::xfwd(p->ctx, nf, in, pp));
  PROF(P_YFWD, nf * 2 * sc, p->ctx.stream, ghx::yfwd(p->ctx, nf, pp, q, prune));
  PROF(P_ZFWD, nf * 2 * sc, p->ctx.stream, ghx::zfwd(p->ctx, nf, q, o, prune));
  timer_end(p);
  return GHX_OK;
}

static int do_c2r(ghx_plan *p, int nf, const void *const *in, double *const *out) {
  ARGCHK(p && in && out && nf >= 1 && nf <= ghx::kMaxFields, "bad c2r arguments");
  int rc = plan_scratch(p, 2 * nf);
  if (rc) return rc;
  const size_t nc = nc_of(p->nx, p->ny, p->nz), slot = ghx::slot_points(p->ctx.gk);
  FieldSet q, pp, i;
  for (int f = 0; f < nf; ++f) {
    ARGCHK(in[f] && out[f], "null field pointer");
    q.p[f] = p->q + f * slot;
    pp.p[f] = p->q + (nf + f) * slot;
    i.p[f] = (cplx *)in[f];
  }
  const double sc = (double)nc * sizeof(cplx), sr = (double)p->nx * p->ny * p->nz * sizeof(double);
  timer_begin(p);
  PROF(P_ZINV, nf * 2 * sc, p->ctx.stream, ghx::zinv(p->ctx, 0, nf, i, q));
  PROF(P_YINV, nf * 2 * sc, p->ctx.stream, ghx::yinv(p->ctx, nf, q, pp));
  PROF(P_XINV, nf * (sr + sc), p->ctx.stream, ghx::xinv(p->ctx, nf, pp, out));
  timer_end(p);
  return GHX_OK;
}

extern "C" int ghx_r2c(ghx_plan *plan, const double *in, void *out) { return do_r2c(plan, 1, &in, &out); }
extern "C" int ghx_c2r(ghx_plan *plan, const void *in, double *out) { return do_c2r(plan, 1, &in, &out); }
extern "C" int ghx_r2c_many(ghx_plan *plan, int nf, const double *const *in, void *const *out) {
  return do_r2c(plan, nf, in, out);
}
extern "C" int ghx_c2r_many(ghx_plan *plan, int nf, const void *const *in, double *const *out) {
  return do_c2r(plan, nf, in, out);
}

extern "C" int ghx_r2c_host(ghx_plan *p, const double *in, void *out) {
  ARGCHK(p && in && out, "bad arguments");
  const size_t nr = (size_t)p->nx * p->ny * p->nz, nc = nc_of(p->nx, p->ny, p->nz);
  double *din = nullptr;
  void *dout = nullptr;
  HIPCHK(hipMalloc(&din, nr * sizeof(double)));
  hipError_t e = hipMalloc(&dout, nc * sizeof(cplx));
  if (e != hipSuccess) { (void)hipFree(din); HIPCHK(e); }
  int rc = GHX_OK;
  if ((e = hipMemcpy(din, in, nr * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) rc = set_err(GHX_ERR_HIP, "%s", hipGetErrorString(e));
  if (!rc) rc = ghx_r2c(p, din, dout);
  if (!rc && (e = hipMemcpyAsync(out, dout, nc * sizeof(cplx), hipMemcpyDeviceToHost, p->ctx.stream)) != hipSuccess)
    rc = set_err(GHX_ERR_HIP, "%s", hipGetErrorString(e));
  if (!rc && (e = hipStreamSynchronize(p->ctx.stream)) != hipSuccess) rc = set_err(GHX_ERR_HIP, "%s", hipGetErrorString(e));
  (void)hipFree(din);
  (void)hipFree(dout);
  return rc;
}

extern "C" int ghx_c2r_host(ghx_plan *p, const void *in, double *out) {
  ARGCHK(p && in && out, "bad arguments");
  const size_t nr = (size_t)p->nx * p->ny * p->nz, nc = nc_of(p->nx, p->ny, p->nz);
  void *din = nullptr;
  double *dout = nullptr;
  HIPCHK(hipMalloc(&din, nc * sizeof(cplx)));
  hipError_t e = hipMalloc(&dout, nr * sizeof(double));
  if (e != hipSuccess) { (void)hipFree(din); HIPCHK(e); }
  int rc = GHX_OK;
  if ((e = hipMemcpy(din, in, nc * sizeof(cplx), hipMemcpyHostToDevice)) != hipSuccess) rc = set_err(GHX_ERR_HIP, "%s", hipGetErrorString(e));
  if (!rc) rc = ghx_c2r(p, din, dout);
  if (!rc && (e = hipMemcpyAsync(out, dout, nr * sizeof(double), hipMemcpyDeviceToHost, p->ctx.stream)) != hipSuccess)
    rc = set_err(GHX_ERR_HIP, "%s", hipGetErrorString(e));
  if (!rc && (e = hipStreamSynchronize(p->ctx.stream)) != hipSuccess) rc = set_err(GHX_ERR_HIP, "%s", hipGetErrorString(e));
  (void)hipFree(din);
  (void)hipFree(dout);
  return rc;
}

extern "C" int ghx_plan_timers(ghx_plan *p, double *ffttime, double *tratime, double *comtime, double *tottime) {
  ARGCHK(p, "null plan");
  timer_collect(p);
  if (ffttime) *ffttime = p->t_fft;
  if (tratime) *tratime = 0.0;
  if (comtime) *comtime = 0.0;
  if (tottime) *tottime = p->t_fft;
  return GHX_OK;
}

extern "C" int ghx_plan_reset_timers(ghx_plan *p) {
  ARGCHK(p, "null plan");
  if (!p->ev0) {
    HIPCHK(hipEventCreate(&p->ev0));
    HIPCHK(hipEventCreate(&p->ev1));
  }
  timer_collect(p);
  p->timing = true;
  p->t_fft = 0.0;
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// grid
// ---------------------------------------------------------------------------
struct ghx_grid {
  ghx_plan plan;
  GridK gk;
  double Dkk = 1.0;
  double L[3] = {1, 1, 1};  // box lengths in units of 2 pi (1 unless arbsize)
  int nmax = 0;
  hipStream_t stream = 0;
  double *work = nullptr;   // reductions / spectra
  size_t work_cap = 0;
  cplx *ctmp = nullptr;     // complex temporaries
  int ctmp_fields = 0;
  double *rtmp = nullptr;   // real temporaries
  int rtmp_fields = 0;
  size_t nc = 0, nr = 0;
  double ps[3] = {1, 1, 1};  // pruned shares (rows, lines, modes), computed on first use
  bool ps_valid = false;
};

static int grid_ctmp(ghx_grid *g, int nf) {
  if (g->ctmp_fields >= nf) return GHX_OK;
  if (g->ctmp) HIPCHK(hipFree(g->ctmp));
  g->ctmp = nullptr;
  g->ctmp_fields = 0;
  HIPCHK(hipMalloc(&g->ctmp, (size_t)nf * g->nc * sizeof(cplx)));
  g->ctmp_fields = nf;
  return GHX_OK;
}
static int grid_rtmp(ghx_grid *g, int nf) {
  if (g->rtmp_fields >= nf) return GHX_OK;
  if (g->rtmp) HIPCHK(hipFree(g->rtmp));
  g->rtmp = nullptr;
  g->rtmp_fields = 0;
  HIPCHK(hipMalloc(&g->rtmp, (size_t)nf * g->nr * sizeof(double)));
  g->rtmp_fields = nf;
  return GHX_OK;
}
static int grid_work(ghx_grid *g, size_t n) {
  if (g->work_cap >= n) return GHX_OK;
  if (g->work) HIPCHK(hipFree(g->work));
  g->work = nullptr;
  g->work_cap = 0;
  HIPCHK(hipMalloc(&g->work, n * sizeof(double)));
  g->work_cap = n;
  return GHX_OK;
}

extern "C" int ghx_grid_create(ghx_grid **grid, int nx, int ny, int nz, double Lx, double Ly, double Lz, int arbsize,
                               double Dkk) {
  ARGCHK(grid, "null grid pointer");
  if (arbsize) ARGCHK(Lx > 0 && Ly > 0 && Lz > 0, "box lengths must be positive");
  ghx_grid *g = new (std::nothrow) ghx_grid();
  if (!g) return set_err(GHX_ERR_NOMEM, "host allocation failed");
  int rc = plan_init(&g->plan, nx, ny, nz, GHX_FFTW_REAL_TO_COMPLEX);
  if (rc) { plan_release(&g->plan); delete g; return rc; }
  GridK k = make_gridk(nx, ny, nz, arbsize, Lx, Ly, Lz);
  // Dkk, nmax, tiny: main3D.fpp:905-909, 1728-1735
  double dkk = arbsize ? Dkk : 0.0;
  if (dkk < 1e-5) dkk = std::fmin(k.dkx, std::fmin(k.dky, k.dkz));
  double mx = std::fmax(nx * k.dkx, std::fmax(ny * k.dky, nz * k.dkz));
  g->nmax = (int)(mx / dkk);
  k.tiny = std::fmin(1e-5, 0.1 / ((double)g->nmax * (double)g->nmax));
  g->Dkk = dkk;
  if (arbsize) { g->L[0] = Lx; g->L[1] = Ly; g->L[2] = Lz; }
  g->gk = k;
  g->plan.ctx.gk = k;
  g->nc = nc_of(nx, ny, nz);
  g->nr = (size_t)nx * ny * nz;
  *grid = g;
  return GHX_OK;
}

extern "C" int ghx_grid_destroy(ghx_grid *g) {
  if (!g) return GHX_OK;
  plan_release(&g->plan);
  if (g->work) (void)hipFree(g->work);
  if (g->ctmp) (void)hipFree(g->ctmp);
  if (g->rtmp) (void)hipFree(g->rtmp);
  delete g;
  return GHX_OK;
}

extern "C" int ghx_grid_set_stream(ghx_grid *g, void *stream) {
  ARGCHK(g, "null grid");
  g->stream = (hipStream_t)stream;
  g->plan.ctx.stream = (hipStream_t)stream;
  return GHX_OK;
}

extern "C" int ghx_grid_info(const ghx_grid *g, double *kmax, double *tiny, int *nmax, double *Dkk, int *anis) {
  ARGCHK(g, "null grid");
  if (kmax) *kmax = g->gk.kmax;
  if (tiny) *tiny = g->gk.tiny;
  if (nmax) *nmax = g->nmax;
  if (Dkk) *Dkk = g->Dkk;
  if (anis) *anis = g->gk.anis;
  return GHX_OK;
}

extern "C" int ghx_grid_r2c(ghx_grid *g, const double *in, void *out) {
  ARGCHK(g, "null grid");
  return ghx_r2c(&g->plan, in, out);
}
extern "C" int ghx_grid_c2r(ghx_grid *g, const void *in, double *out) {
  ARGCHK(g, "null grid");
  return ghx_c2r(&g->plan, in, out);
}

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
#define C_(p) ((const cplx *)(p))
#define M_(p) ((cplx *)(p))

extern "C" int ghx_derivk3(ghx_grid *g, const void *a, void *b, int dir) {
  ARGCHK(g && a && b && dir >= 1 && dir <= 3, "bad derivk3 arguments");
  HIPCHK(ghx::op_derivk3(g->gk, g->stream, C_(a), M_(b), dir));
  return GHX_OK;
}
extern "C" int ghx_laplak3(ghx_grid *g, const void *a, void *b) {
  ARGCHK(g && a && b, "bad laplak3 arguments");
  HIPCHK(ghx::op_laplak3(g->gk, g->stream, C_(a), M_(b)));
  return GHX_OK;
}
extern "C" int ghx_rotor3(ghx_grid *g, const void *a, const void *b, void *c, int dir) {
  ARGCHK(g && a && b && c && dir >= 1 && dir <= 3, "bad rotor3 arguments");
  HIPCHK(ghx::op_rotor3(g->gk, g->stream, C_(a), C_(b), M_(c), dir));
  return GHX_OK;
}
extern "C" int ghx_nonlhd3(ghx_grid *g, const void *a, const void *b, const void *c, void *out, int dir) {
  ARGCHK(g && a && b && c && out && dir >= 1 && dir <= 3, "bad nonlhd3 arguments");
  HIPCHK(ghx::op_nonlhd3(g->gk, g->stream, C_(a), C_(b), C_(c), M_(out), dir));
  return GHX_OK;
}
extern "C" int ghx_nonlin3(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                           const void *f, void *out, int dir) {
  ARGCHK(g && a && b && c && d && e && f && out && dir >= 1 && dir <= 3, "bad nonlin3 arguments");
  HIPCHK(ghx::op_nonlin3(g->gk, g->stream, C_(a), C_(b), C_(c), C_(d), C_(e), C_(f), M_(out), dir));
  return GHX_OK;
}
extern "C" int ghx_gauge3(ghx_grid *g, const void *a, const void *b, const void *c, void *out, int dir) {
  ARGCHK(g && a && b && c && out && dir >= 1 && dir <= 3, "bad gauge3 arguments");
  HIPCHK(ghx::op_gauge3(g->gk, g->stream, C_(a), C_(b), C_(c), M_(out), dir));
  return GHX_OK;
}

// (P x Q) products through the fused passes.  mode 1: P = curl(a,b,c), Q = (a,b,c)
// (prodre3); mode 0: P = (a,b,c), Q = (d,e,f) (vector3).
// Scratch slots of the fused chain: Q6 (slots 0-5) -> P6 (6-11) -> P3 (reuses
// 0-2) -> Q3 (reuses 6-8).
struct Chain {
  FieldSet q6, p6, p3, q3;
};
static Chain chain_slots(cplx *base, size_t slot) {
  Chain c;
  for (int k = 0; k < 6; ++k) { c.q6.p[k] = base + k * slot; c.p6.p[k] = base + (6 + k) * slot; }
  for (int k = 0; k < 3; ++k) { c.p3.p[k] = base + k * slot; c.q3.p[k] = base + (6 + k) * slot; }
  return c;
}

static int fused_cross(ghx_grid *g, int mode, const void *a, const void *b, const void *c, const void *d,
                       const void *e, const void *f, void *x, void *y, void *z) {
  int rc = plan_scratch(&g->plan, 12);
  if (rc) return rc;
  const size_t nc = g->nc;
  const FFTCtx &ctx = g->plan.ctx;
  Chain ch = chain_slots(g->plan.q, ghx::slot_points(g->gk));
  FieldSet in, out;
  FieldSet &q6 = ch.q6;
  out.p[0] = M_(x); out.p[1] = M_(y); out.p[2] = M_(z);
  const double sc = (double)nc * sizeof(cplx);
  FieldSet pq;  // P fields first, then Q
  if (mode == 1) {
    in.p[0] = M_(a); in.p[1] = M_(b); in.p[2] = M_(c);
    PROF(P_ZINV_CURL, 9 * sc, ctx.stream, ghx::zinv(ctx, 1, 3, in, q6));  // q6 = (u, curl u)
    for (int k = 0; k < 3; ++k) { pq.p[k] = ch.p6.p[3 + k]; pq.p[3 + k] = ch.p6.p[k]; }
  } else {
    in.p[0] = M_(a); in.p[1] = M_(b); in.p[2] = M_(c); in.p[3] = M_(d); in.p[4] = M_(e); in.p[5] = M_(f);
    PROF(P_ZINV, 12 * sc, ctx.stream, ghx::zinv(ctx, 0, 6, in, q6));
    for (int k = 0; k < 6; ++k) pq.p[k] = ch.p6.p[k];
  }
  PROF(P_YINV, 12 * sc, ctx.stream, ghx::yinv(ctx, 6, q6, ch.p6));
  PROF(P_XCROSS, 9 * sc, ctx.stream, ghx::xcross(ctx, 0, pq, ch.p3));
  PROF(P_YFWD, 6 * sc, ctx.stream, ghx::yfwd(ctx, 3, ch.p3, ch.q3));
  PROF(P_ZFWD, 6 * sc, ctx.stream, ghx::zfwd(ctx, 3, ch.q3, out));
  return GHX_OK;
}

extern "C" int ghx_prodre3(ghx_grid *g, const void *a, const void *b, const void *c, void *d, void *e, void *f) {
  ARGCHK(g && a && b && c && d && e && f, "bad prodre3 arguments");
  return fused_cross(g, 1, a, b, c, nullptr, nullptr, nullptr, d, e, f);
}
extern "C" int ghx_vector3(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                           const void *f, void *x, void *y, void *z) {
  ARGCHK(g && a && b && c && d && e && f && x && y && z, "bad vector3 arguments");
  return fused_cross(g, 0, a, b, c, d, e, f, x, y, z);
}

// analysis/trunc3D.fpp:190-320: c (spectrum on g) truncated onto the smaller
// grid t; a c2r on t then gives the truncated field on t's points.
extern "C" int ghx_truncate(ghx_grid *g, const void *c, ghx_grid *t, void *out) {
  ARGCHK(g && c && t && out, "bad truncate arguments");
  const GridK &a = g->gk, &b = t->gk;
  if (b.nx > a.nx || b.ny > a.ny || b.nz > a.nz || b.nx % 2 || b.ny % 2 || b.nz % 2)
    return set_err(GHX_ERR_SIZE, "truncated grid %dx%dx%d must be even and no larger than %dx%dx%d", b.nx, b.ny,
                   b.nz, a.nx, a.ny, a.nz);
  HIPCHK(ghx::op_truncate(a, b, t->stream, C_(c), M_(out)));
  return GHX_OK;
}

// advect3 (pseudospec3D_phd.f90:23-115): Z^-1 of (a, b, c) and grad d in one
// pass, Y^-1, the dot product fused into the x pass, Y, Z.
extern "C" int ghx_advect3(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, void *e) {
  ARGCHK(g && a && b && c && d && e, "bad advect3 arguments");
  int rc = plan_scratch(&g->plan, 12);
  if (rc) return rc;
  const FFTCtx &ctx = g->plan.ctx;
  const size_t slot = ghx::slot_points(g->gk);
  FieldSet in, q6, p6, p1, q1, out;
  in.p[0] = M_(a); in.p[1] = M_(b); in.p[2] = M_(c); in.p[3] = M_(d);
  for (int k = 0; k < 6; ++k) { q6.p[k] = g->plan.q + k * slot; p6.p[k] = g->plan.q + (6 + k) * slot; }
  p1.p[0] = g->plan.q;
  q1.p[0] = g->plan.q + 6 * slot;
  out.p[0] = M_(e);
  const double sc = (double)g->nc * sizeof(cplx);
  PROF(P_ZINV, 10 * sc, ctx.stream, ghx::zinv(ctx, 4, 4, in, q6));
  PROF(P_YINV, 12 * sc, ctx.stream, ghx::yinv(ctx, 6, q6, p6));
  PROF(P_XDOT, 7 * sc, ctx.stream, ghx::xdot(ctx, p6, p1));
  PROF(P_YFWD, 2 * sc, ctx.stream, ghx::yfwd(ctx, 1, p1, q1));
  PROF(P_ZFWD, 2 * sc, ctx.stream, ghx::zfwd(ctx, 1, q1, out));
  return GHX_OK;
}

static int reduce(ghx_grid *g, int kind, const void *a, const void *b, const void *c, const void *d, const void *e,
                  const void *f, double *out) {
  int rc = grid_work(g, ghx::reduce_work_doubles());
  if (rc) return rc;
  HIPCHK(ghx::op_reduce(g->gk, g->stream, kind, C_(a), C_(b), C_(c), C_(d), C_(e), C_(f), g->work, out));
  return GHX_OK;
}

static int curl3(ghx_grid *g, const void *a, const void *b, const void *c, cplx *o) {
  HIPCHK(ghx::op_rotor3(g->gk, g->stream, C_(b), C_(c), o, 1));
  HIPCHK(ghx::op_rotor3(g->gk, g->stream, C_(a), C_(c), o + g->nc, 2));
  HIPCHK(ghx::op_rotor3(g->gk, g->stream, C_(a), C_(b), o + 2 * g->nc, 3));
  return GHX_OK;
}

extern "C" int ghx_energy(ghx_grid *g, const void *a, const void *b, const void *c, int kin, double *out) {
  ARGCHK(g && a && b && c && out && kin >= 0 && kin <= 2, "bad energy arguments");
  if (kin == 1) return reduce(g, 0, a, b, c, nullptr, nullptr, nullptr, out);
  int rc = grid_ctmp(g, 3);
  if (rc) return rc;
  cplx *t = g->ctmp;
  if (kin == 0) {
    if ((rc = curl3(g, a, b, c, t))) return rc;
  } else {
    HIPCHK(ghx::op_laplak3(g->gk, g->stream, C_(a), t));
    HIPCHK(ghx::op_laplak3(g->gk, g->stream, C_(b), t + g->nc));
    HIPCHK(ghx::op_laplak3(g->gk, g->stream, C_(c), t + 2 * g->nc));
  }
  return reduce(g, 0, t, t + g->nc, t + 2 * g->nc, nullptr, nullptr, nullptr, out);
}

extern "C" int ghx_helicity(ghx_grid *g, const void *a, const void *b, const void *c, double *out) {
  ARGCHK(g && a && b && c && out, "bad helicity arguments");
  int rc = grid_ctmp(g, 3);
  if (rc) return rc;
  if ((rc = curl3(g, a, b, c, g->ctmp))) return rc;
  return reduce(g, 1, a, b, c, g->ctmp, g->ctmp + g->nc, g->ctmp + 2 * g->nc, out);
}

extern "C" int ghx_cross(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                         const void *f, int kin, double *out) {
  ARGCHK(g && a && b && c && d && e && f && out && (kin == 0 || kin == 1), "bad cross arguments");
  return reduce(g, kin == 1 ? 1 : 2, a, b, c, d, e, f, out);
}

extern "C" int ghx_divergence(ghx_grid *g, const void *a, const void *b, const void *c, double *out) {
  ARGCHK(g && a && b && c && out, "bad divergence arguments");
  return reduce(g, 3, a, b, c, nullptr, nullptr, nullptr, out);
}

extern "C" int ghx_maxabs(ghx_grid *g, const void *a, const void *b, const void *c, int kin, double *out) {
  ARGCHK(g && a && b && c && out && kin >= 0 && kin <= 2, "bad maxabs arguments");
  int rc = grid_ctmp(g, 3);
  if (!rc) rc = grid_rtmp(g, 3);
  if (!rc) rc = grid_work(g, ghx::reduce_work_doubles());
  if (rc) return rc;
  cplx *t = g->ctmp;
  const void *src[3];
  if (kin == 0) {
    if ((rc = curl3(g, a, b, c, t))) return rc;
    src[0] = t; src[1] = t + g->nc; src[2] = t + 2 * g->nc;
  } else if (kin == 1) {
    HIPCHK(ghx::op_laplak3(g->gk, g->stream, C_(a), t));
    HIPCHK(ghx::op_laplak3(g->gk, g->stream, C_(b), t + g->nc));
    HIPCHK(ghx::op_laplak3(g->gk, g->stream, C_(c), t + 2 * g->nc));
    src[0] = t; src[1] = t + g->nc; src[2] = t + 2 * g->nc;
  } else {
    src[0] = a; src[1] = b; src[2] = c;
  }
  double *r[3] = {g->rtmp, g->rtmp + g->nr, g->rtmp + 2 * g->nr};
  if ((rc = do_c2r(&g->plan, 3, src, r))) return rc;
  double m = 0;
  HIPCHK(ghx::op_maxnorm(g->gk, g->stream, r[0], r[1], r[2], g->work, &m));
  *out = m / ((double)g->gk.nx * (double)g->gk.ny * (double)g->gk.nz);
  return GHX_OK;
}

extern "C" int ghx_spectrumc(ghx_grid *g, const void *a, const void *b, const void *c, int kin, int hel, double *Ek,
                             double *Hk) {
  ARGCHK(g && a && b && c && kin >= 0 && kin <= 2 && (hel == 0 || hel == 1), "bad spectrumc arguments");
  const int nb = g->nmax / 2 + 1;
  int rc = grid_work(g, ghx::spectrum_work_doubles(g->gk, nb));
  if (rc) return rc;
  cplx *t = nullptr;
  if (kin == 0 || hel == 1) {
    if ((rc = grid_ctmp(g, 3))) return rc;
    t = g->ctmp;
    if ((rc = curl3(g, a, b, c, t))) return rc;
  }
  if (kin <= 1 && Ek) {
    if (kin == 1)
      HIPCHK(ghx::op_spectrum(g->gk, g->stream, 0, g->Dkk, nb, C_(a), C_(b), C_(c), nullptr, nullptr, nullptr, g->work, Ek));
    else
      HIPCHK(ghx::op_spectrum(g->gk, g->stream, 0, g->Dkk, nb, t, t + g->nc, t + 2 * g->nc, nullptr, nullptr, nullptr,
                              g->work, Ek));
  }
  if (hel == 1 && Hk)
    HIPCHK(ghx::op_spectrum(g->gk, g->stream, 1, g->Dkk, nb, C_(a), C_(b), C_(c), t, t + g->nc, t + 2 * g->nc,
                            g->work, Hk));
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// The energy spectrum without device allocation, for callers that cannot call
// hipMalloc (code under stream capture, fixed-memory serving loops).  The
// per-block shell partials and their totals live in `scratch`, a file-scope
// device array of kNoAllocDoubles doubles that every GPU the library runs on
// holds until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): each device's
// address is looked up once with hipGetSymbolAddress, with that device
// current, and only that address is used.  Every call on one GPU shares its
// array, so calls that overlap on one GPU (two host threads, two streams
// without ordering) race; each call writes every partial it then sums, so
// nothing needs clearing between calls.
// ---------------------------------------------------------------------------
static constexpr size_t kNoAllocDoubles = (size_t)1 << 19;  // 4 MiB
static __device__ __align__(256) double scratch[kNoAllocDoubles];
static constexpr int kNoAllocDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kNoAllocDevices];

static int scratch_address(double **out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices)
    return set_err(GHX_ERR_ARG, "spectrumc_noalloc: device %d beyond the %d the scratch table holds", dev,
                   kNoAllocDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_err(GHX_ERR_HIP, "spectrumc_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return GHX_OK;
}

extern "C" int ghx_spectrumc_noalloc(ghx_grid *g, const void *a, const void *b, const void *c, double *Ek) {
  ARGCHK(g && a && b && c && Ek, "bad spectrumc_noalloc arguments");
  const int nb = g->nmax / 2 + 1;
  const size_t need = ghx::spectrum_work_doubles(g->gk, nb);
  if (need > kNoAllocDoubles)
    return set_err(GHX_ERR_SIZE, "spectrumc_noalloc: %zu doubles of shell partials exceed the %zu-double scratch",
                   need, kNoAllocDoubles);
  double *w = nullptr;
  int rc = scratch_address(&w);
  if (rc) return rc;
  HIPCHK(ghx::op_spectrum(g->gk, g->stream, 0, g->Dkk, nb, C_(a), C_(b), C_(c), nullptr, nullptr, nullptr, w, Ek));
  return GHX_OK;
}

// entrans (pseudospec3D_hd.f90:1629-1820): shell sums of
// Re(a d* + b e* + c f*) / N^6, times kk2 when kin = 0.  Ek: nmax/2+1 host
// doubles, the reduced totals (GHOST's files hold Ek/Dkk).
extern "C" int ghx_entrans(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                           const void *f, int kin, double *Ek) {
  ARGCHK(g && a && b && c && d && e && f && Ek && kin >= 0, "bad entrans arguments");
  const int nb = g->nmax / 2 + 1;
  int rc = grid_work(g, ghx::spectrum_work_doubles(g->gk, nb));
  if (rc) return rc;
  HIPCHK(ghx::op_spectrum(g->gk, g->stream, kin == 0 ? 2 : 1, g->Dkk, nb, C_(a), C_(b), C_(c), C_(d), C_(e), C_(f),
                          g->work, Ek));
  return GHX_OK;
}

// heltrans (:1823-1949): curl(a,b,c) against (d,e,f)
extern "C" int ghx_heltrans(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                            const void *f, double *Hk) {
  ARGCHK(g && a && b && c && d && e && f && Hk, "bad heltrans arguments");
  const int nb = g->nmax / 2 + 1;
  int rc = grid_work(g, ghx::spectrum_work_doubles(g->gk, nb));
  if (!rc) rc = grid_ctmp(g, 3);
  if (!rc) rc = curl3(g, a, b, c, g->ctmp);
  if (rc) return rc;
  const cplx *t = g->ctmp;
  HIPCHK(ghx::op_spectrum(g->gk, g->stream, 1, g->Dkk, nb, t, t + g->nc, t + 2 * g->nc, C_(d), C_(e), C_(f), g->work,
                          Hk));
  return GHX_OK;
}

// crosspecc (pseudospec3D_mhd.f90:603-712): Re(v . conj(curl A)) per shell,
// shells int(sqrt(kk2)/Dkk + shift + .501) with the offset in working precision
extern "C" int ghx_crosspecc(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                             const void *f, double shift, double *Ck) {
  ARGCHK(g && a && b && c && d && e && f && Ck, "bad crosspecc arguments");
  const int nb = g->nmax / 2 + 1;
  int rc = grid_work(g, ghx::spectrum_work_doubles(g->gk, nb));
  if (!rc) rc = grid_ctmp(g, 3);
  if (!rc) rc = curl3(g, d, e, f, g->ctmp);
  if (rc) return rc;
  const cplx *t = g->ctmp;
  HIPCHK(ghx::op_spectrum(g->gk, g->stream, 1, g->Dkk, nb, C_(a), C_(b), C_(c), t, t + g->nc, t + 2 * g->nc, g->work,
                          Ck, shift + 0.501));
  return GHX_OK;
}

// spectr1d (pseudospec3D_hd.f90:1418-1626): the 1D spectrum of one component
// along k_dir, bins 1..n_dir/2+1 of Ek (nmax/2+1 host doubles; GHOST's
// files hold 0.5 Ek / Dk_dir)
extern "C" int ghx_spectr1d(ghx_grid *g, const void *a, int dir, double *Ek) {
  ARGCHK(g && a && Ek && dir >= 1 && dir <= 3, "bad spectr1d arguments");
  const int nb = g->nmax / 2 + 1;
  int rc = grid_work(g, ghx::spectrum_work_doubles(g->gk, nb));
  if (rc) return rc;
  for (int q = 0; q < nb; ++q) Ek[q] = 0.0;
  HIPCHK(ghx::op_spectr1d(g->gk, g->stream, dir, g->L[dir - 1], nb, C_(a), g->work, Ek));
  return GHX_OK;
}

// ---- scalar diagnostics (pseudospec3D_phd.f90) ----------------------------

// variance (:116-217): <th^2> (kin = 1) or <kk2 |th|^2> (kin = 0)
extern "C" int ghx_variance(ghx_grid *g, const void *a, int kin, double *out) {
  ARGCHK(g && a && out && (kin == 0 || kin == 1), "bad variance arguments");
  return reduce(g, kin == 1 ? 6 : 7, a, a, a, nullptr, nullptr, nullptr, out);
}

// product (:219-284): <a b>
extern "C" int ghx_product(ghx_grid *g, const void *a, const void *b, double *out) {
  ARGCHK(g && a && b && out, "bad product arguments");
  return reduce(g, 8, a, a, a, b, b, b, out);
}

// pscheck (:286-333) without its file output: {<th^2>, <|grad th|^2>, <th s>}
extern "C" int ghx_pscheck(ghx_grid *g, const void *a, const void *b, double *out) {
  ARGCHK(g && a && b && out, "bad pscheck arguments");
  int rc = ghx_variance(g, a, 1, &out[0]);
  if (!rc) rc = ghx_variance(g, a, 0, &out[1]);
  if (!rc) rc = ghx_product(g, a, b, &out[2]);
  return rc;
}

// spectrscc (:516-610): |th|^2 per shell int(sqrt(kk2)/Dkk + shift + .501),
// the offset in working precision
extern "C" int ghx_spectrscc(ghx_grid *g, const void *a, double shift, double *Ek) {
  ARGCHK(g && a && Ek, "bad spectrscc arguments");
  const int nb = g->nmax / 2 + 1;
  int rc = grid_work(g, ghx::spectrum_work_doubles(g->gk, nb));
  if (rc) return rc;
  HIPCHK(ghx::op_spectrum(g->gk, g->stream, 3, g->Dkk, nb, C_(a), C_(a), C_(a), C_(a), C_(a), C_(a), g->work, Ek,
                          shift + 0.501));
  return GHX_OK;
}

// sctrans (:612-727): Re(a b*) per shell (single-precision .501 offset)
extern "C" int ghx_sctrans(ghx_grid *g, const void *a, const void *b, double *Ek) {
  ARGCHK(g && a && b && Ek, "bad sctrans arguments");
  const int nb = g->nmax / 2 + 1;
  int rc = grid_work(g, ghx::spectrum_work_doubles(g->gk, nb));
  if (rc) return rc;
  HIPCHK(ghx::op_spectrum(g->gk, g->stream, 4, g->Dkk, nb, C_(a), C_(a), C_(a), C_(b), C_(b), C_(b), g->work, Ek));
  return GHX_OK;
}

// difucx (dir 1, :729-828) / difucz (dir 3, :830-940): plane means of v_i,
// th and v_i th along x or z; out = 3 n host doubles
static int difuc(ghx_grid *g, const void *a, const void *b, int dir, double *out) {
  ARGCHK(g && a && b && out, "bad difuc arguments");
  const int n = dir == 1 ? g->gk.nx : g->gk.nz;
  int rc = grid_rtmp(g, 2);
  if (!rc) rc = grid_work(g, 3 * (size_t)n);
  if (rc) return rc;
  double *r1 = g->rtmp, *r2 = g->rtmp + g->nr;
  const void *in[2] = {a, b};
  double *ro[2] = {r1, r2};
  if ((rc = do_c2r(&g->plan, 2, in, ro))) return rc;
  const double nx = g->gk.nx, ny = g->gk.ny, nz = g->gk.nz;
  const double tmp = dir == 1 ? 1.0 / (nx * (ny * ny) * (nz * nz)) : 1.0 / ((nx * nx) * (ny * ny) * nz);
  const double tmq = dir == 1 ? 1.0 / ((nx * nx) * (ny * ny * ny) * (nz * nz * nz))
                              : 1.0 / ((nx * nx * nx) * (ny * ny * ny) * (nz * nz));
  HIPCHK(ghx::op_plane_means(g->gk, g->stream, r1, r2, dir, tmp, tmq, g->work, out));
  return GHX_OK;
}
extern "C" int ghx_difucx(ghx_grid *g, const void *a, const void *b, double *out) { return difuc(g, a, b, 1, out); }
extern "C" int ghx_difucz(ghx_grid *g, const void *a, const void *b, double *out) { return difuc(g, a, b, 3, out); }

// ---- axisymmetric spectra (pseudospec3D_rot.f90), preferred axis z --------

// bins of the plane spectra (nz/2+1) and of the cylinder spectra
// (nmaxperp/2+1, nmaxperp = int(max(nx Dkx, ny Dky)/Dkk), main3D.fpp:1729)
static int nb_para(const ghx_grid *g) { return g->gk.nz / 2 + 1; }
static int nb_perp(const ghx_grid *g) {
  const double a = g->gk.nx * g->gk.dkx, b = g->gk.ny * g->gk.dky;
  return (int)((a > b ? a : b) / g->Dkk) / 2 + 1;
}
extern "C" int ghx_axis_bins(const ghx_grid *g, int *nbpara, int *nbperp) {
  ARGCHK(g, "null grid");
  if (nbpara) *nbpara = nb_para(g);
  if (nbperp) *nbperp = nb_perp(g);
  return GHX_OK;
}

static int axispec(ghx_grid *g, int bins, int kind, const void *a, const void *b, const void *c, const void *d,
                   const void *e, const void *f, double *out) {
  const int nb = bins == 0 ? nb_para(g) : nb_perp(g);
  ARGCHK(3 * (size_t)nb * sizeof(double) <= 48 * 1024, "axisymmetric spectrum: too many bins for one block");
  int rc = grid_work(g, ghx::axispec_work_doubles(g->gk, nb));
  if (rc) return rc;
  if (!d) { d = a; e = b; f = c; }
  HIPCHK(ghx::op_axispec(g->gk, g->stream, bins, kind, nb, g->Dkk, g->L[2], C_(a), C_(b), C_(c), C_(d), C_(e), C_(f),
                         g->work, out));
  return GHX_OK;
}

// specpara (:25-335) / specperp (:338-686) without file output: Ek (kin 1:
// of (a,b,c); kin 0: of its curl; kin 2: skipped) and Hk (hel = 1: Re(a .
// conj(curl a))), each 3 nb host doubles {total, xy, z}; NULL skips one.
static int axis_ek(ghx_grid *g, int bins, const void *a, const void *b, const void *c, int kin, int hel, double *Ek,
                   double *Hk) {
  ARGCHK(g && a && b && c && kin >= 0 && kin <= 2, "bad specpara/specperp arguments");
  int rc = GHX_OK;
  const bool need_curl = kin == 0 || hel == 1;
  if (need_curl && ((rc = grid_ctmp(g, 3)) || (rc = curl3(g, a, b, c, g->ctmp)))) return rc;
  const cplx *t0 = g->ctmp, *t1 = g->ctmp + g->nc, *t2 = g->ctmp + 2 * g->nc;
  if (kin == 1 && Ek && (rc = axispec(g, bins, 0, a, b, c, nullptr, nullptr, nullptr, Ek))) return rc;
  if (kin == 0 && Ek && (rc = axispec(g, bins, 0, t0, t1, t2, nullptr, nullptr, nullptr, Ek))) return rc;
  if (hel == 1 && Hk && (rc = axispec(g, bins, 1, a, b, c, t0, t1, t2, Hk))) return rc;
  return GHX_OK;
}
extern "C" int ghx_specpara(ghx_grid *g, const void *a, const void *b, const void *c, int kin, int hel, double *Ek,
                            double *Hk) {
  return axis_ek(g, 0, a, b, c, kin, hel, Ek, Hk);
}
extern "C" int ghx_specperp(ghx_grid *g, const void *a, const void *b, const void *c, int kin, int hel, double *Ek,
                            double *Hk) {
  return axis_ek(g, 1, a, b, c, kin, hel, Ek, Hk);
}

// entpara (:689-875) / entperp (:878-1064): Re(a d* + b e* + c f*) per bin,
// times kk2 when kin = 0; Ek: nb host doubles
static int axis_ent(ghx_grid *g, int bins, const void *a, const void *b, const void *c, const void *d, const void *e,
                    const void *f, int kin, double *Ek) {
  ARGCHK(g && a && b && c && d && e && f && Ek && kin >= 0, "bad entpara/entperp arguments");
  const int nb = bins == 0 ? nb_para(g) : nb_perp(g);
  std::vector<double> o(3 * (size_t)nb);
  int rc = axispec(g, bins, kin == 0 ? 2 : 1, a, b, c, d, e, f, o.data());
  if (rc) return rc;
  for (int q = 0; q < nb; ++q) Ek[q] = o[q];
  return GHX_OK;
}
extern "C" int ghx_entpara(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                           const void *f, int kin, double *Ek) {
  return axis_ent(g, 0, a, b, c, d, e, f, kin, Ek);
}
extern "C" int ghx_entperp(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                           const void *f, int kin, double *Ek) {
  return axis_ent(g, 1, a, b, c, d, e, f, kin, Ek);
}

// heltpara (:1067-1196) / heltperp (:1199-1328): curl(a,b,c) against (d,e,f)
static int axis_helt(ghx_grid *g, int bins, const void *a, const void *b, const void *c, const void *d,
                     const void *e, const void *f, double *Hk) {
  ARGCHK(g && a && b && c && d && e && f && Hk, "bad heltpara/heltperp arguments");
  int rc = grid_ctmp(g, 3);
  if (!rc) rc = curl3(g, a, b, c, g->ctmp);
  if (rc) return rc;
  return axis_ent(g, bins, g->ctmp, g->ctmp + g->nc, g->ctmp + 2 * g->nc, d, e, f, 1, Hk);
}
extern "C" int ghx_heltpara(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                            const void *f, double *Hk) {
  return axis_helt(g, 0, a, b, c, d, e, f, Hk);
}
extern "C" int ghx_heltperp(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                            const void *f, double *Hk) {
  return axis_helt(g, 1, a, b, c, d, e, f, Hk);
}

// spec2D (:1331-1629) without file output: Ek, Hk of nbperp * nbpara host
// doubles, kperp fastest (GHOST's Ek(kperp, kpara)); kin and hel as specpara
extern "C" int ghx_spec2d(ghx_grid *g, const void *a, const void *b, const void *c, int kin, int hel, double *Ek,
                          double *Hk) {
  ARGCHK(g && a && b && c && kin >= 0 && kin <= 2, "bad spec2d arguments");
  const int np = nb_perp(g), nzb = nb_para(g);
  int rc = grid_work(g, (size_t)np * nzb);
  if (rc) return rc;
  const bool need_curl = kin == 0 || hel == 1;
  if (need_curl && ((rc = grid_ctmp(g, 3)) || (rc = curl3(g, a, b, c, g->ctmp)))) return rc;
  const cplx *t0 = g->ctmp, *t1 = g->ctmp + g->nc, *t2 = g->ctmp + 2 * g->nc;
  if (kin <= 1 && Ek) {
    const cplx *x = kin == 1 ? C_(a) : t0, *y = kin == 1 ? C_(b) : t1, *z = kin == 1 ? C_(c) : t2;
    HIPCHK(ghx::op_spec2d(g->gk, g->stream, 0, np, nzb, g->Dkk, g->L[2], x, y, z, x, y, z, g->work, Ek));
  }
  if (hel == 1 && Hk)
    HIPCHK(ghx::op_spec2d(g->gk, g->stream, 1, np, nzb, g->Dkk, g->L[2], C_(a), C_(b), C_(c), t0, t1, t2, g->work, Hk));
  return GHX_OK;
}

// ---- Boussinesq diagnostics (pseudospec3D_bouss.f90) ---------------------

// havgcomp (:19-555): horizontal averages as functions of z, gsh = nz host
// doubles.  itype 0 <(du_perp/dz)^2>, 1 <(dth/dz)^2>, 2 <u_z dth/dz>, 3
// <u^2 + v^2>, 4 <w^2>, 5 <u_perp . (curl u)_perp>, 6 <omega_z th>, 7 <PV^2>
// (PV = fo dth/dz - bv omega_z + omega . grad th - fo bv), 8 <omega . curl
// omega>, 9 the shear of itype 0 (GHOST's first branch serves 0 and 9 and
// returns), 10 <u_z th>.  For itype 2 GHOST reads a work array the call never
// fills; the documented correlation with dth/dz is formed here.
extern "C" int ghx_havgcomp(ghx_grid *g, const void *u, const void *v, const void *w, const void *s, double fo,
                            double bv, int itype, double *gsh) {
  ARGCHK(g && u && v && w && s && gsh && itype >= 0 && itype <= 10, "bad havgcomp arguments");
  int rc = grid_ctmp(g, 4);
  if (!rc) rc = grid_rtmp(g, 4);
  if (!rc) rc = grid_work(g, (size_t)g->gk.nz);
  if (rc) return rc;
  const size_t nc = g->nc, nr = g->nr;
  cplx *c1 = g->ctmp, *c2 = c1 + nc, *c3 = c2 + nc, *c4 = c3 + nc;
  double *r1 = g->rtmp, *r2 = r1 + nr, *r3 = r2 + nr, *r4 = r3 + nr;
  const double nx = g->gk.nx, ny = g->gk.ny, nz = g->gk.nz;
  const double tmp = 1.0 / (nx * nx * nx * ny * ny * ny * nz * nz);
  auto c2r = [&](int n, const void *a0, double *o0, const void *a1 = nullptr, double *o1 = nullptr,
                 const void *a2 = nullptr, double *o2 = nullptr) {
    const void *in[3] = {a0, a1, a2};
    double *out[3] = {o0, o1, o2};
    return do_c2r(&g->plan, n, in, out);
  };
  auto zsum = [&](int kind, double scale, double *out) {
    HIPCHK(ghx::op_zplane(g->gk, g->stream, kind, r1, r2, r3, r4, fo, bv, scale, g->work, out));
    return GHX_OK;
  };
  switch (itype) {
    case 0:
    case 9:
      if ((rc = ghx_derivk3(g, u, c1, 3)) || (rc = ghx_derivk3(g, v, c2, 3)) || (rc = c2r(2, c1, r1, c2, r2)))
        return rc;
      return zsum(0, tmp, gsh);
    case 1:
      if ((rc = ghx_derivk3(g, s, c1, 3)) || (rc = c2r(1, c1, r1))) return rc;
      return zsum(1, tmp, gsh);
    case 2:
      if ((rc = ghx_derivk3(g, s, c1, 3)) || (rc = c2r(2, c1, r1, w, r2))) return rc;
      return zsum(2, tmp, gsh);
    case 3:
      if ((rc = c2r(2, u, r1, v, r2))) return rc;
      return zsum(0, tmp, gsh);
    case 4:
      if ((rc = c2r(1, w, r1))) return rc;
      return zsum(1, tmp, gsh);
    case 5:
      if ((rc = ghx_derivk3(g, v, c1, 3)) || (rc = ghx_derivk3(g, u, c2, 3))) return rc;
      if ((rc = c2r(2, u, r1, c1, r2)) || (rc = c2r(2, v, r3, c2, r4))) return rc;
      return zsum(3, tmp, gsh);
    case 6:
    case 10:
      if (itype == 6) {
        if ((rc = ghx_rotor3(g, u, v, c1, 3)) || (rc = c2r(2, c1, r1, s, r2))) return rc;
      } else if ((rc = c2r(2, w, r1, s, r2))) {
        return rc;
      }
      return zsum(2, tmp, gsh);
    case 7:
      // r1 = dth/dz, r3 = omega_z, r4 = -omega . grad th (advect3)
      if ((rc = ghx_derivk3(g, s, c4, 3)) || (rc = curl3(g, u, v, w, c1))) return rc;
      if ((rc = ghx_advect3(g, c1, c2, c3, s, c1))) return rc;
      if ((rc = ghx_rotor3(g, u, v, c2, 3))) return rc;
      if ((rc = c2r(3, c4, r1, c2, r3, c1, r4))) return rc;
      return zsum(5, 1.0 / (nx * ny), gsh);
    case 8: {
      // omega . curl omega, one component at a time: r3 (r1 - r2)
      const int n = g->gk.nz;
      std::vector<double> part(3 * (size_t)n);
      if ((rc = curl3(g, u, v, w, c1))) return rc;  // c1, c2, c3 = omega_x, omega_y, omega_z
      if ((rc = ghx_derivk3(g, c3, c4, 2)) || (rc = c2r(2, c4, r1, c1, r3))) return rc;
      if ((rc = ghx_derivk3(g, c2, c4, 3)) || (rc = c2r(1, c4, r2))) return rc;
      if ((rc = zsum(4, 1.0, part.data()))) return rc;
      if ((rc = ghx_derivk3(g, c1, c4, 3)) || (rc = c2r(2, c4, r1, c2, r3))) return rc;
      if ((rc = ghx_derivk3(g, c3, c4, 1)) || (rc = c2r(1, c4, r2))) return rc;
      if ((rc = zsum(4, 1.0, part.data() + n))) return rc;
      if ((rc = ghx_derivk3(g, c2, c4, 1)) || (rc = c2r(2, c4, r1, c3, r3))) return rc;
      if ((rc = ghx_derivk3(g, c1, c4, 2)) || (rc = c2r(1, c4, r2))) return rc;
      if ((rc = zsum(4, 1.0, part.data() + 2 * n))) return rc;
      for (int k = 0; k < n; ++k) gsh[k] = (part[k] + part[n + k] + part[2 * n + k]) * tmp;
      return GHX_OK;
    }
  }
  return GHX_OK;
}

// tbouss (:612-709) without its file output: the mean, maximum and minimum
// over z of havgcomp's itype 0..10 (11 host doubles each)
extern "C" int ghx_tbouss(ghx_grid *g, const void *u, const void *v, const void *w, const void *s, double fo,
                          double bv, double *avg, double *mx, double *mn) {
  ARGCHK(g && u && v && w && s && avg && mx && mn, "bad tbouss arguments");
  std::vector<double> gsh((size_t)g->gk.nz);
  for (int i = 0; i < 11; ++i) {
    int rc = ghx_havgcomp(g, u, v, w, s, fo, bv, i, gsh.data());
    if (rc) return rc;
    double sum = 0.0, hi = gsh[0], lo = gsh[0];
    for (double x : gsh) {
      sum += x;
      hi = x > hi ? x : hi;
      lo = x < lo ? x : lo;
    }
    avg[i] = sum / (double)g->gk.nz;
    mx[i] = hi;
    mn[i] = lo;
  }
  return GHX_OK;
}

// spectpv (:712-812): |omega . grad th|^2 over planes int(|kz| Lz + 1),
// nz/2+1 host doubles
extern "C" int ghx_spectpv(ghx_grid *g, const void *u, const void *v, const void *w, const void *s, double *Ek) {
  ARGCHK(g && u && v && w && s && Ek, "bad spectpv arguments");
  int rc = grid_ctmp(g, 4);
  if (!rc) rc = curl3(g, u, v, w, g->ctmp);
  if (rc) return rc;
  cplx *c = g->ctmp, *a = g->ctmp + 3 * g->nc;
  if ((rc = ghx_advect3(g, c, c + g->nc, c + 2 * g->nc, s, a))) return rc;
  const int nb = nb_para(g);
  std::vector<double> o(3 * (size_t)nb);
  if ((rc = axispec(g, 0, 3, a, a, a, nullptr, nullptr, nullptr, o.data()))) return rc;
  for (int q = 0; q < nb; ++q) Ek[q] = o[q];
  return GHX_OK;
}

// hdcheck (pseudospec3D_hd.f90:984-1105) without its file output: out =
// {eng, ens, pot, khe, div}: <v^2>, <omega^2>, the injection rate <v.f>,
// the kinetic helicity (hel = 1) and <(div v)^2> (chk = 1); entries that
// hel/chk do not ask for are 0.
extern "C" int ghx_hdcheck(ghx_grid *g, const void *a, const void *b, const void *c, const void *d, const void *e,
                           const void *f, int hel, int chk, double *out) {
  ARGCHK(g && a && b && c && d && e && f && out, "bad hdcheck arguments");
  for (int q = 0; q < 5; ++q) out[q] = 0.0;
  int rc = GHX_OK;
  if (chk == 1 && (rc = ghx_divergence(g, a, b, c, &out[4]))) return rc;
  if ((rc = ghx_energy(g, a, b, c, 1, &out[0]))) return rc;
  if ((rc = ghx_energy(g, a, b, c, 0, &out[1]))) return rc;
  if (hel == 1 && (rc = ghx_helicity(g, a, b, c, &out[3]))) return rc;
  return ghx_cross(g, a, b, c, d, e, f, 1, &out[2]);
}

// mhdcheck (pseudospec3D_mhd.f90:339-547) without its file output: out =
// {eng, ens, cur, engk, engm, helk, helm, crh, asq, divk, divm}; entries that
// hel/crs/chk do not ask for are 0.  crs = 2 (the Hall generalised helicity)
// is not provided.
extern "C" int ghx_mhdcheck(ghx_grid *g, const void *a, const void *b, const void *c, const void *ma, const void *mb,
                            const void *mc, int hel, int crs, int chk, double *out) {
  ARGCHK(g && a && b && c && ma && mb && mc && out, "bad mhdcheck arguments");
  ARGCHK(crs >= 0 && crs <= 1, "mhdcheck: crs = 2 (Hall generalised helicity) is not provided");
  for (int q = 0; q < 11; ++q) out[q] = 0.0;
  int rc = GHX_OK;
  if (chk == 1) {
    if ((rc = ghx_divergence(g, a, b, c, &out[9]))) return rc;
    if ((rc = ghx_divergence(g, ma, mb, mc, &out[10]))) return rc;
  }
  double engk = 0, ens = 0, engm = 0, cur = 0;
  if ((rc = ghx_energy(g, a, b, c, 1, &engk))) return rc;
  if ((rc = ghx_energy(g, a, b, c, 0, &ens))) return rc;
  // the magnetic field curl A lives past the three temporaries energy uses
  if ((rc = grid_ctmp(g, 6))) return rc;
  cplx *bm = g->ctmp + 3 * g->nc;
  if ((rc = curl3(g, ma, mb, mc, bm))) return rc;
  const cplx *b0 = bm, *b1 = bm + g->nc, *b2 = bm + 2 * g->nc;
  if ((rc = ghx_energy(g, b0, b1, b2, 1, &engm))) return rc;
  if ((rc = ghx_energy(g, b0, b1, b2, 0, &cur))) return rc;
  out[0] = engk + engm;
  out[1] = ens;
  out[2] = cur;
  out[3] = engk;
  out[4] = engm;
  if (hel == 1) {
    if ((rc = ghx_helicity(g, a, b, c, &out[5]))) return rc;
    if ((rc = ghx_helicity(g, ma, mb, mc, &out[6]))) return rc;
  }
  if (crs >= 1) {
    if ((rc = ghx_energy(g, ma, mb, mc, 1, &out[8]))) return rc;
    if ((rc = ghx_cross(g, a, b, c, b0, b1, b2, 1, &out[7]))) return rc;
  }
  return GHX_OK;
}

extern "C" int ghx_phaseshift(ghx_grid *g, void *a, double cre, double cim) {
  ARGCHK(g && a, "bad phaseshift arguments");
  HIPCHK(ghx::op_phaseshift(g->gk, g->stream, M_(a), cre, cim));
  return GHX_OK;
}

extern "C" int ghx_normalize(ghx_grid *g, void *fx, void *fy, void *fz, double f0, int kin) {
  ARGCHK(g && fx && fy && fz, "bad normalize arguments");
  double e = 0;
  int rc = ghx_energy(g, fx, fy, fz, kin, &e);
  if (rc) return rc;
  HIPCHK(ghx::op_normalize_apply(g->gk, g->stream, M_(fx), M_(fy), M_(fz), f0, e));
  return GHX_OK;
}

extern "C" int ghx_sharp(ghx_grid *g, const void *a, void *b, double delta) {
  ARGCHK(g && a && b && delta > 0, "bad sharp arguments");
  const double pi = 3.14159265358979323846;
  HIPCHK(ghx::op_sharp(g->gk, g->stream, C_(a), M_(b), (pi / delta) * (pi / delta)));
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// field output and restart (binary files of binio.hip)
// ---------------------------------------------------------------------------
// main3D.fpp:2675-2710 (VELOC_): C = v * rmp with rmp = 1/(nx ny nz); with
// outs >= 1 the vorticity rotor3(C) is transformed and written as wx, wy, wz
// first; then vx, vy, vz = c2r(C).
extern "C" int ghx_hd_output(ghx_grid *g, ghx_ioplan *io, const char *odir, const char *ext, int outs,
                             const void *vx, const void *vy, const void *vz) {
  ARGCHK(g && io && odir && ext && vx && vy && vz, "bad hd_output arguments");
  int rc = grid_ctmp(g, outs >= 1 ? 6 : 3);
  if (!rc) rc = grid_rtmp(g, 1);
  if (rc) return rc;
  const double rmp = 1.0 / ((double)g->gk.nx * (double)g->gk.ny * (double)g->gk.nz);
  cplx *C[6];
  for (int f = 0; f < 6; ++f) C[f] = g->ctmp + (size_t)f * g->nc;
  const void *v[3] = {vx, vy, vz};
  for (int c = 0; c < 3; ++c) HIPCHK(ghx::op_scale(g->gk, g->stream, C[c], C_(v[c]), rmp));
  double *R = g->rtmp;
  static const char *const wn[3] = {"wx", "wy", "wz"}, *const vn[3] = {"vx", "vy", "vz"};
  if (outs >= 1) {
    HIPCHK(ghx::op_rotor3(g->gk, g->stream, C[1], C[2], C[3], 1));
    HIPCHK(ghx::op_rotor3(g->gk, g->stream, C[0], C[2], C[4], 2));
    HIPCHK(ghx::op_rotor3(g->gk, g->stream, C[0], C[1], C[5], 3));
    for (int c = 0; c < 3; ++c) {
      const void *src = C[3 + c];
      if ((rc = do_c2r(&g->plan, 1, &src, &R))) return rc;
      if ((rc = ghx_io_write(io, odir, wn[c], ext, R))) return rc;
    }
  }
  for (int c = 0; c < 3; ++c) {
    const void *src = C[c];
    if ((rc = do_c2r(&g->plan, 1, &src, &R))) return rc;
    if ((rc = ghx_io_write(io, odir, vn[c], ext, R))) return rc;
  }
  return GHX_OK;
}

// main3D.fpp:2143-2169 (VELOC_, stat > 0): vx, vy, vz read from their files
// and transformed with r2c.
extern "C" int ghx_hd_restart(ghx_grid *g, ghx_ioplan *io, const char *idir, const char *ext, void *vx, void *vy,
                              void *vz) {
  ARGCHK(g && io && idir && ext && vx && vy && vz, "bad hd_restart arguments");
  int rc = grid_rtmp(g, 1);
  if (rc) return rc;
  static const char *const vn[3] = {"vx", "vy", "vz"};
  void *v[3] = {vx, vy, vz};
  double *R = g->rtmp;
  for (int c = 0; c < 3; ++c) {
    HIPCHK(hipStreamSynchronize(g->stream));  // the previous transform has read R
    if ((rc = ghx_io_read(io, idir, vn[c], ext, R))) return rc;
    const double *src = R;
    if ((rc = do_r2c(&g->plan, 1, &src, &v[c]))) return rc;
  }
  HIPCHK(hipStreamSynchronize(g->stream));
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// initial conditions / forcing
// ---------------------------------------------------------------------------
extern "C" double ghx_randu(int32_t *idum) {
  // Park-Miller minimal standard, Schrage factorisation, XOR mask; the scale
  // am = 1./im is a single-precision constant in GHOST.
  const int32_t iq = 127773, ir = 2836, mask = 123459876, ia = 16807, im = 2147483647;
  const double am = (double)(1.0f / (float)im);
  int32_t s = *idum ^ mask;
  const int32_t k = s / iq;
  s = ia * (s - k * iq) - ir * k;
  if (s < 0) s += im;
  double r = am * (double)s;
  r = (r - 0.5) * 2;
  *idum = s ^ mask;
  return r;
}

extern "C" int ghx_initialv_tg(ghx_grid *g, double kdn, double kup, double u0, void *vx, void *vy, void *vz) {
  ARGCHK(g && vx && vy && vz, "bad initialv_tg arguments");
  int rc = grid_rtmp(g, 2);
  if (rc) return rc;
  double *r1 = g->rtmp, *r2 = g->rtmp + g->nr;
  HIPCHK(ghx::op_tg_real(g->gk, g->stream, r1, r2, (int)kdn, (int)kup));
  HIPCHK(ghx::op_zero(g->stream, vz, g->nc * sizeof(cplx)));
  const double *in[2] = {r1, r2};
  void *out[2] = {vx, vy};
  if ((rc = do_r2c(&g->plan, 2, in, out))) return rc;
  return ghx_normalize(g, vx, vy, vz, u0, 1);
}

// Host-side random-phase fill of three fields in the loop order of
// examples/initialfv.f90_random:16-135 (kx=0 plane with its Hermitian mirrors,
// then kx>0 planes), one kx plane at a time, uploaded to c1..c3.
namespace {
struct HostGrid {
  int nx, ny, nz;
  GridK k;
  double kxv(int i) const { return ghxw(i, nx) * k.dkx; }
  double kyv(int j) const { return ghxw(j, ny) * k.dky; }
  double kzv(int q) const { return ghxw(q, nz) * k.dkz; }
  static double ghxw(int q, int n) { return (double)(q < n / 2 ? q : q - n); }
  double kn2(int i, int j, int q) const {
    double a = ghxw(i, nx), b = ghxw(j, ny), c = ghxw(q, nz);
    return k.rmp * (a * a) + k.rmq * (b * b) + k.rms * (c * c);
  }
  double kk2(int i, int j, int q) const {
    if (!k.anis) return kn2(i, j, q);
    double a = kxv(i), b = kyv(j), c = kzv(q);
    return a * a + b * b + c * c;
  }
};

template <typename Band, typename Amp>
int random_fill(ghx_grid *g, Band inband, Amp amp, int32_t *seed, cplx *c1, cplx *c2, cplx *c3) {
  const double pi = 3.14159265358979323846;
  HostGrid hg{g->gk.nx, g->gk.ny, g->gk.nz, g->gk};
  const int ny = hg.ny, nz = hg.nz;
  const size_t plane = (size_t)ny * nz;
  std::vector<cplx> h(3 * plane);
  cplx *C[3] = {h.data(), h.data() + plane, h.data() + 2 * plane};
  cplx *D[3] = {c1, c2, c3};
  auto draw = [&](double fac, double dump) {
    double phase = 2 * pi * ghx_randu(seed);
    cplx r;
    r.x = fac * std::cos(phase) * dump;
    r.y = fac * std::sin(phase) * dump;
    return r;
  };
  auto conjc = [](cplx a) { a.y = -a.y; return a; };
  const cplx zero = {0.0, 0.0};
  // kx = 0 plane, m(k,j) = j*nz + k
  for (int f = 0; f < 3; ++f) C[f][0] = zero;
  auto pair = [&](size_t m, size_t mm, int j, int q) {
    if (inband(0, j, q)) {
      double dump = amp(0, j, q);
      for (int f = 0; f < 3; ++f) {
        C[f][m] = draw(1.0, dump);
        C[f][mm] = conjc(C[f][m]);
      }
    } else {
      for (int f = 0; f < 3; ++f) { C[f][m] = zero; C[f][mm] = zero; }
    }
  };
  for (int j = 2; j <= ny / 2 + 1; ++j) pair((size_t)(j - 1) * nz, (size_t)(ny - j + 1) * nz, j - 1, 0);
  for (int q = 2; q <= nz / 2 + 1; ++q) pair((size_t)(q - 1), (size_t)(nz - q + 1), 0, q - 1);
  for (int j = 2; j <= ny; ++j)
    for (int q = 2; q <= nz / 2 + 1; ++q)
      pair((size_t)(j - 1) * nz + (q - 1), (size_t)(ny - j + 1) * nz + (nz - q + 1), j - 1, q - 1);
  for (int f = 0; f < 3; ++f) HIPCHK(hipMemcpy(D[f], C[f], plane * sizeof(cplx), hipMemcpyHostToDevice));
  for (int i = 1; i < g->gk.nxh; ++i) {
    for (int j = 0; j < ny; ++j)
      for (int q = 0; q < nz; ++q) {
        size_t m = (size_t)j * nz + q;
        if (inband(i, j, q)) {
          double dump = amp(i, j, q);
          for (int f = 0; f < 3; ++f) C[f][m] = draw(2.0, dump);
        } else {
          for (int f = 0; f < 3; ++f) C[f][m] = zero;
        }
      }
    for (int f = 0; f < 3; ++f)
      HIPCHK(hipMemcpy(D[f] + (size_t)i * plane, C[f], plane * sizeof(cplx), hipMemcpyHostToDevice));
  }
  return GHX_OK;
}

}  // namespace

static int finish_random(ghx_grid *g, double f0, void *fx, void *fy, void *fz) {
  cplx *C = g->ctmp;
  // rotor3(C2,C3,fx,1); rotor3(C1,C3,fy,2); rotor3(C1,C2,fz,3); normalize
  HIPCHK(ghx::op_rotor3(g->gk, g->stream, C + g->nc, C + 2 * g->nc, M_(fx), 1));
  HIPCHK(ghx::op_rotor3(g->gk, g->stream, C, C + 2 * g->nc, M_(fy), 2));
  HIPCHK(ghx::op_rotor3(g->gk, g->stream, C, C + g->nc, M_(fz), 3));
  return ghx_normalize(g, fx, fy, fz, f0, 1);
}

extern "C" int ghx_initialfv_random(ghx_grid *g, double kdn, double kup, double f0, int32_t *seed, void *fx, void *fy,
                                    void *fz) {
  ARGCHK(g && seed && fx && fy && fz, "bad initialfv_random arguments");
  int rc = grid_ctmp(g, 3);
  if (rc) return rc;
  HostGrid hg{g->gk.nx, g->gk.ny, g->gk.nz, g->gk};
  auto band = [&](int i, int j, int q) {
    double k2 = hg.kk2(i, j, q);
    return k2 <= kup * kup && k2 >= kdn * kdn;
  };
  auto amp = [&](int i, int j, int q) { return 1. / std::sqrt(hg.kk2(i, j, q)); };
  if ((rc = random_fill(g, band, amp, seed, g->ctmp, g->ctmp + g->nc, g->ctmp + 2 * g->nc))) return rc;
  return finish_random(g, f0, fx, fy, fz);
}

extern "C" int ghx_initialv_k4exp(ghx_grid *g, double k0, double u0, int32_t *seed, void *vx, void *vy, void *vz) {
  ARGCHK(g && seed && vx && vy && vz && k0 > 0, "bad initialv_k4exp arguments");
  int rc = grid_ctmp(g, 3);
  if (rc) return rc;
  HostGrid hg{g->gk.nx, g->gk.ny, g->gk.nz, g->gk};
  auto band = [&](int i, int j, int q) {
    double k2 = hg.kn2(i, j, q);
    return k2 <= g->gk.kmax && k2 >= g->gk.tiny;
  };
  auto amp = [&](int i, int j, int q) { return std::exp(-hg.kk2(i, j, q) / (k0 * k0)); };
  if ((rc = random_fill(g, band, amp, seed, g->ctmp, g->ctmp + g->nc, g->ctmp + 2 * g->nc))) return rc;
  return finish_random(g, u0, vx, vy, vz);
}

// ---------------------------------------------------------------------------
// RK substeps
// ---------------------------------------------------------------------------

// Shares of the spectral volume touched by the pruned passes (dealiased
// fields): fx = kx rows with kept modes, fl = (kx,ky) lines with kept modes,
// fm = kept modes (tiny <= kn2 <= kmax).
struct PruneShares {
  double fx, fl, fm;
};
static PruneShares prune_shares_compute(const GridK &g) {
  auto kn2 = [&](int i, int j, int k) {
    double a = (double)(i < g.nx / 2 ? i : i - g.nx), b = (double)(j < g.ny / 2 ? j : j - g.ny),
           c = (double)(k < g.nz / 2 ? k : k - g.nz);
    return g.rmp * (a * a) + g.rmq * (b * b) + g.rms * (c * c);
  };
  double rows = 0, lines = 0, modes = 0;
  for (int i = 0; i < g.nxh; ++i) {
    if (kn2(i, 0, 0) <= g.kmax) rows += 1;
    for (int j = 0; j < g.ny; ++j) {
      if (kn2(i, j, 0) > g.kmax) continue;
      lines += 1;
      for (int k = 0; k < g.nz; ++k) {
        double q = kn2(i, j, k);
        if (q <= g.kmax && q >= g.tiny) modes += 1;
      }
    }
  }
  const double nl = (double)g.nxh * g.ny;
  return PruneShares{rows / g.nxh, lines / nl, modes / (nl * g.nz)};
}
static PruneShares prune_shares(ghx_grid *g) {
  if (!g->ps_valid) {
    PruneShares p = prune_shares_compute(g->gk);
    g->ps[0] = p.fx; g->ps[1] = p.fl; g->ps[2] = p.fm;
    g->ps_valid = true;
  }
  return PruneShares{g->ps[0], g->ps[1], g->ps[2]};
}

// Fused HD substep: q holds 9 Q-layout fields of scratch.
// omegaz != 0: the rotating-frame step of include/roth_rkstep2.f90 (Coriolis
// force added to the products in the last pass).
static int hd_substep(ghx_grid *g, cplx *q, int o, double dt, double nu, const cplx *const *v, const cplx *const *c,
                      const cplx *const *f, const uint8_t *fline, cplx *const *out, int prune = 0,
                      double omegaz = 0.0) {
  const size_t nc = g->nc;
  const FFTCtx &ctx = g->plan.ctx;
  Chain ch = chain_slots(q, ghx::slot_points(g->gk));
  FieldSet in, pq;
  for (int k = 0; k < 3; ++k) in.p[k] = const_cast<cplx *>(v[k]);
  for (int k = 0; k < 3; ++k) { pq.p[k] = ch.p6.p[3 + k]; pq.p[3 + k] = ch.p6.p[k]; }
  const double sc = (double)nc * sizeof(cplx);
  // Algorithmic traffic of each pass.  Pruned: fx = share of kx rows carrying
  // kept modes, fl = share of (kx,ky) lines, fm = share of kept modes.
  const PruneShares ps = (prune && g_prof.on) ? prune_shares(g) : PruneShares{1.0, 1.0, 1.0};
  PROF(P_ZINV_CURL, (3 * ps.fm + 6 * ps.fl) * sc, ctx.stream, ghx::zinv(ctx, 1, 3, in, ch.q6, prune));
  PROF(P_YINV, (6 * ps.fl + 6 * ps.fx) * sc, ctx.stream, ghx::yinv(ctx, 6, ch.q6, ch.p6, prune));
  PROF(P_XCROSS, 9 * ps.fx * sc, ctx.stream, ghx::xcross(ctx, 0, pq, ch.p3, prune));
  PROF(P_YFWD, (3 * ps.fx + 3 * ps.fl) * sc, ctx.stream, ghx::yfwd(ctx, 3, ch.p3, ch.q3, prune));
  ghx::RKParams rk;
  rk.dt = dt; rk.nu = nu; rk.rmp = 1. / (double)o;
  rk.om2 = 2 * omegaz;
  for (int k = 0; k < 3; ++k) {
    rk.v[k] = v[k]; rk.c[k] = c[k]; rk.out[k] = out[k];
    rk.f[k] = f ? f[k] : nullptr;
  }
  rk.fline = f ? fline : nullptr;
  // reads the three products on the transformed lines, v and C (once if they
  // alias) and the forcing on the kept modes, writes v_out on the kept modes
  // (on every mode without pruning)
  const int nvc = (const void *)c[0] == (const void *)v[0] ? 3 : 6;
  PROF(P_ZFWD_RK_HD, (3 * ps.fl + nvc * ps.fm + 3 * (prune ? ps.fm : 1.0)) * sc, ctx.stream,
       ghx::zfwd_rk_hd(ctx, ch.q3, rk, prune));
  return GHX_OK;
}

// lines (kx,ky) on which any of the fields is non-zero
__global__ void k_line_flags(const cplx *a, const cplx *b, const cplx *c, int nz, size_t nlines, uint8_t *flag) {
  for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < nlines; l += (size_t)gridDim.x * blockDim.x) {
    uint8_t nzf = 0;
    const size_t base = l * nz;
    for (int k = 0; k < nz; ++k) {
      cplx x = a[base + k], y = b[base + k], z = c[base + k];
      if (x.x != 0.0 || x.y != 0.0 || y.x != 0.0 || y.y != 0.0 || z.x != 0.0 || z.y != 0.0) nzf = 1;
    }
    flag[l] = nzf;
  }
}

static int line_flags(ghx_grid *g, const cplx *a, const cplx *b, const cplx *c, uint8_t *flag) {
  size_t nl = (size_t)g->gk.nxh * g->gk.ny;
  unsigned nb = (unsigned)((nl + 255) / 256);
  hipLaunchKernelGGL(k_line_flags, dim3(nb), dim3(256), 0, g->stream, a, b, c, g->gk.nz, nl, flag);
  HIPCHK(hipGetLastError());
  return GHX_OK;
}

static int rkstep2(ghx_grid *g, int o, double dt, double nu, double omegaz, const void *const *v_in,
                   const void *const *C, const void *const *f, void *const *v_out) {
  ARGCHK(g && o >= 1 && v_in && C && v_out, "bad rkstep2 arguments");
  int rc = plan_scratch(&g->plan, 12);
  if (rc) return rc;
  const cplx *v[3], *c[3], *ff[3];
  cplx *out[3];
  for (int k = 0; k < 3; ++k) {
    v[k] = C_(v_in[k]); c[k] = C_(C[k]); out[k] = M_(v_out[k]);
    ff[k] = f ? C_(f[k]) : nullptr;
  }
  uint8_t *fl = nullptr;
  if (f) {
    // dense forcing: every line is read
    size_t nl = (size_t)g->gk.nxh * g->gk.ny;
    if ((rc = grid_work(g, (nl + 7) / 8 + 1))) return rc;
    fl = (uint8_t *)g->work;
    if ((rc = line_flags(g, ff[0], ff[1], ff[2], fl))) return rc;
  }
  return hd_substep(g, g->plan.q, o, dt, nu, v, c, f ? ff : nullptr, fl, out, 0, omegaz);
}

extern "C" int ghx_hd_rkstep2(ghx_grid *g, int o, double dt, double nu, const void *const *v_in, const void *const *C,
                              const void *const *f, void *const *v_out) {
  return rkstep2(g, o, dt, nu, 0.0, v_in, C, f, v_out);
}

// include/roth_rkstep2.f90:1-49: the HD substep in a frame rotating about z
extern "C" int ghx_roth_rkstep2(ghx_grid *g, int o, double dt, double nu, double omegaz, const void *const *v_in,
                                const void *const *C, const void *const *f, void *const *v_out) {
  return rkstep2(g, o, dt, nu, omegaz, v_in, C, f, v_out);
}

struct ghx_hd {
  ghx_grid *g;
  int ord;
  double dt, nu;
  cplx *mem = nullptr;  // v(3), v1(3), f(3), 12 scratch slots
  cplx *v[3], *v1[3], *f[3], *q;
  uint8_t *fline = nullptr;
  bool forced = false;
  bool clean = false;  // v known to be zero outside the kept modes
  ghx_tracers *tr = nullptr;  // Lagrangian tracers advanced with the flow
  // learned subgrid closure (LES): network, filter width and its buffers
  ghx_mlp *sgs = nullptr;
  double sgs_delta = 0.0;
  void *sgs_mem = nullptr;  // grad 9 + tau 6 real, tau 6 + force 3 spectral
  uint8_t *sgs_fline = nullptr;
  double *sgs_nuk = nullptr;  // spectral viscosity per shell of the closure (device), sgs_nb shells
  int sgs_nb = 0;
  double omegaz = 0.0;  // rotation rate about z (include/roth_rkstep2.f90)
};

static int tracers_begin_step(ghx_tracers *t);
static int tracers_stage(ghx_tracers *t, cplx *q, const cplx *const *v, double dt, double xk, int prune);

// Count of non-zero modes of (a,b,c) outside the dealias set.
static int count_unkept(ghx_grid *g, const cplx *a, const cplx *b, const cplx *c, double *n) {
  return reduce(g, 4, a, b, c, nullptr, nullptr, nullptr, n);
}

extern "C" int ghx_hd_create(ghx_hd **s, ghx_grid *g, int ord, double dt, double nu) {
  ARGCHK(s && g && ord >= 1, "bad hd_create arguments");
  ghx_hd *h = new (std::nothrow) ghx_hd();
  if (!h) return set_err(GHX_ERR_NOMEM, "host allocation failed");
  h->g = g; h->ord = ord; h->dt = dt; h->nu = nu;
  (void)prune_shares(g);  // host-side table of pruned shares, computed once per grid
  const size_t nc = g->nc;
  hipError_t e = hipMalloc(&h->mem, (9 * nc + 12 * ghx::slot_points(g->gk)) * sizeof(cplx));
  if (e != hipSuccess) { delete h; HIPCHK(e); }
  for (int k = 0; k < 3; ++k) {
    h->v[k] = h->mem + k * nc;
    h->v1[k] = h->mem + (3 + k) * nc;
    h->f[k] = h->mem + (6 + k) * nc;
  }
  h->q = h->mem + 9 * nc;  // 12 scratch slots (Q/P layouts)
  e = hipMalloc(&h->fline, (size_t)g->gk.nxh * g->gk.ny);
  if (e != hipSuccess) { (void)hipFree(h->mem); delete h; HIPCHK(e); }
  e = ghx::op_zero(g->stream, h->mem, 9 * nc * sizeof(cplx));
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) { (void)hipFree(h->mem); (void)hipFree(h->fline); delete h; HIPCHK(e); }
  *s = h;
  return GHX_OK;
}

extern "C" int ghx_hd_destroy(ghx_hd *s) {
  if (!s) return GHX_OK;
  (void)hipFree(s->mem);
  (void)hipFree(s->fline);
  if (s->sgs_mem) (void)hipFree(s->sgs_mem);
  if (s->sgs_fline) (void)hipFree(s->sgs_fline);
  if (s->sgs_nuk) (void)hipFree(s->sgs_nuk);
  delete s;
  return GHX_OK;
}

extern "C" int ghx_hd_fields(ghx_hd *s, void **v, void **f) {
  ARGCHK(s, "null solver");
  s->clean = false;  // the caller may write through these pointers
  for (int k = 0; k < 3; ++k) {
    if (v) v[k] = s->v[k];
    if (f) f[k] = s->f[k];
  }
  return GHX_OK;
}

extern "C" int ghx_hd_set_rotation(ghx_hd *s, double omegaz) {
  ARGCHK(s, "null solver");
  s->omegaz = omegaz;
  return GHX_OK;
}

extern "C" int ghx_hd_set_forcing(ghx_hd *s, int enabled) {
  ARGCHK(s, "null solver");
  s->forced = enabled != 0;
  if (s->forced) {
    int rc = line_flags(s->g, s->f[0], s->f[1], s->f[2], s->fline);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s->g->stream));
  }
  return GHX_OK;
}

// main3D.fpp:3202-3231.  The state lives in v; C (step start) is v itself
// for the first substep, so the intermediate substeps write v1 and the last
// one writes back into v element-wise over C.
extern "C" int ghx_hd_touch(ghx_hd *s) {
  ARGCHK(s, "null solver");
  s->clean = false;
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// learned LES (no counterpart in GHOST)
// ---------------------------------------------------------------------------
// Physical velocity gradients d u_i / d x_j (i major) of a spectral state:
// derivk3 scaled by 1/(nx ny nz), then c2r of the nine fields.  One fused
// Z^-1 takes v to the nine scaled derivatives (the same operations as derivk3
// then the scale), then Y^-1 and X^-1 of nine fields.
// prune: v is dealiased (zero outside the kept modes), so the passes skip
// what is zero (the result is the same)
static int velocity_gradients(ghx_grid *g, const void *const *v, double *const *grad9, int prune = 0) {
  int rc = plan_scratch(&g->plan, 18);
  if (rc) return rc;
  ghx_plan *p = &g->plan;
  const size_t slot = ghx::slot_points(p->ctx.gk);
  FieldSet in, q, pp;
  for (int k = 0; k < 3; ++k) in.p[k] = const_cast<cplx *>(C_(v[k]));
  for (int f = 0; f < 9; ++f) {
    q.p[f] = p->q + f * slot;
    pp.p[f] = p->q + (9 + f) * slot;
  }
  HIPCHK(ghx::zinv(p->ctx, 6, 9, in, q, prune, nullptr));
  HIPCHK(ghx::yinv(p->ctx, 9, q, pp, prune));
  HIPCHK(ghx::xinv(p->ctx, 9, pp, grad9, prune));
  return GHX_OK;
}

extern "C" int ghx_velocity_gradients(ghx_grid *g, const void *const *v, double *const *grad9) {
  ARGCHK(g && v && grad9 && v[0] && v[1] && v[2], "bad velocity_gradients arguments");
  for (int k = 0; k < 9; ++k) ARGCHK(grad9[k], "null gradient field");
  return velocity_gradients(g, v, grad9);
}

// Training pairs of the closure from a fine-grid state v (spectral, GHOST's
// normalisation): ub = sharp(v, kc); on the small grid t, vsmall = the
// truncation of ub in t's normalisation (the LES state) and tau6 = the
// physical subgrid stress F(u_i u_j) - F(ub_i ub_j) sampled on t's points,
// F = sharp(kc) then trunc3D's truncation onto t.
extern "C" int ghx_sgs_targets(ghx_grid *g, ghx_grid *t, const void *const *v, double kc, double *const *tau6,
                               void *const *vsmall) {
  ARGCHK(g && t && v && tau6 && vsmall && kc > 0, "bad sgs_targets arguments");
  const GridK &a = g->gk, &b = t->gk;
  if (b.nx > a.nx || b.ny > a.ny || b.nz > a.nz)
    return set_err(GHX_ERR_SIZE, "the LES grid must not be larger than the fine grid");
  int rc = grid_ctmp(g, 6);
  if (!rc) rc = grid_rtmp(g, 12);
  if (!rc) rc = grid_ctmp(t, 6);
  if (rc) return rc;
  const size_t nc = g->nc, nr = g->nr;
  cplx *ub[3];
  double *r[12];
  for (int k = 0; k < 3; ++k) ub[k] = g->ctmp + k * nc;
  for (int k = 0; k < 12; ++k) r[k] = g->rtmp + k * nr;
  for (int k = 0; k < 3; ++k) HIPCHK(ghx::op_sharp(a, g->stream, C_(v[k]), ub[k], kc * kc));
  const void *srcv[3] = {v[0], v[1], v[2]}, *srcb[3] = {ub[0], ub[1], ub[2]};
  if ((rc = do_c2r(&g->plan, 3, srcv, r))) return rc;
  if ((rc = do_c2r(&g->plan, 3, srcb, r + 3))) return rc;
  // the LES state: truncation (which scales by 1/(nx ny nz)) times t's size
  const double st = (double)b.nx * (double)b.ny * (double)b.nz;
  for (int k = 0; k < 3; ++k) {
    HIPCHK(ghx::op_truncate(a, b, g->stream, ub[k], M_(vsmall[k])));
    HIPCHK(ghx::op_scale(b, g->stream, M_(vsmall[k]), C_(vsmall[k]), st));
  }
  const double s = 1.0 / ((double)a.nx * (double)a.ny * (double)a.nz);
  const double *u[3] = {r[0], r[1], r[2]}, *uf[3] = {r[3], r[4], r[5]};
  HIPCHK(ghx::op_sgs_products(a, g->stream, u, uf, s * s, r + 6));
  const double *srcp[6] = {r[6], r[7], r[8], r[9], r[10], r[11]};
  void *pc[6];
  for (int k = 0; k < 6; ++k) pc[k] = g->ctmp + k * nc;
  if ((rc = do_r2c(&g->plan, 6, srcp, pc))) return rc;
  const void *tc[6];
  for (int k = 0; k < 6; ++k) {
    cplx *c = M_(pc[k]);
    HIPCHK(ghx::op_sharp(a, g->stream, c, c, kc * kc));
    cplx *d = t->ctmp + k * t->nc;
    HIPCHK(ghx::op_truncate(a, b, g->stream, c, d));
    tc[k] = d;
  }
  HIPCHK(hipStreamSynchronize(g->stream));  // t's transforms may run on another stream
  if ((rc = do_c2r(&t->plan, 6, tc, tau6))) return rc;
  HIPCHK(hipStreamSynchronize(t->stream));
  return GHX_OK;
}

// The subgrid forcing of one substep: gradients of vin, tau from the network,
// its transform, and f + P(-div tau) into the solver's buffers.
static int sgs_force(ghx_hd *s, const cplx *const *vin, const cplx **fs, int prune) {
  ghx_grid *g = s->g;
  const size_t nc = g->nc, nr = g->nr;
  double *gr = static_cast<double *>(s->sgs_mem);
  double *tr = gr + 9 * nr;
  cplx *tc = reinterpret_cast<cplx *>(tr + 6 * nr), *fc = tc + 6 * nc;
  const cplx *f0[3] = {s->f[0], s->f[1], s->f[2]};
  cplx *out[3] = {fc, fc + nc, fc + 2 * nc};
  if (!s->sgs) {  // the spectral viscosity alone
    HIPCHK(ghx::op_sgs_force(g->gk, g->stream, nullptr, s->forced ? f0 : nullptr, out, vin, s->sgs_nuk, s->sgs_nb,
                             g->Dkk));
    for (int k = 0; k < 3; ++k) fs[k] = out[k];
    return GHX_OK;
  }
  double *grad9[9], *tau6[6];
  for (int k = 0; k < 9; ++k) grad9[k] = gr + k * nr;
  for (int k = 0; k < 6; ++k) tau6[k] = tr + k * nr;
  int rc;
  const void *v[3] = {vin[0], vin[1], vin[2]};
  if ((rc = velocity_gradients(g, v, grad9, prune))) return rc;
  if ((rc = ghx_mlp_set_stream(s->sgs, g->stream))) return rc;
  const double *cg[9];
  for (int k = 0; k < 9; ++k) cg[k] = grad9[k];
  if ((rc = ghx_mlp_stress(s->sgs, (long long)nr, cg, s->sgs_delta, tau6))) return rc;
  const double *ct[6];
  void *tcv[6];
  for (int k = 0; k < 6; ++k) { ct[k] = tau6[k]; tcv[k] = tc + k * nc; }
  // k_sgs_force reads the stress spectra on the kept modes only
  if ((rc = do_r2c(&g->plan, 6, ct, tcv, 1))) return rc;
  const cplx *tau[6];
  for (int k = 0; k < 6; ++k) tau[k] = tc + k * nc;
  HIPCHK(ghx::op_sgs_force(g->gk, g->stream, tau, s->forced ? f0 : nullptr, out, s->sgs_nuk ? vin : nullptr,
                           s->sgs_nuk, s->sgs_nb, g->Dkk));
  for (int k = 0; k < 3; ++k) fs[k] = out[k];
  return GHX_OK;
}

static int sgs_buffers(ghx_hd *s);

// Attach the closure (m = NULL detaches): before every substep the forcing
// becomes f + P(-div tau) with tau = m(grad u, delta).
extern "C" int ghx_hd_set_sgs(ghx_hd *s, ghx_mlp *m, double delta) {
  ARGCHK(s, "null solver");
  if (!m) {
    s->sgs = nullptr;
    return GHX_OK;
  }
  ARGCHK(delta > 0, "filter width must be positive");
  int rc = sgs_buffers(s);
  if (rc) return rc;
  s->sgs = m;
  s->sgs_delta = delta;
  return GHX_OK;
}

// Spectral part of the closure: an extra viscosity nuk[q] on the shells
// q + 1 = int(sqrt(kk2)/Dkk + .501) = 1..nb, -nuk kk2 v added to the force of
// every substep (with or without a network attached); nb = 0 removes it.
extern "C" int ghx_hd_set_sgs_spectral(ghx_hd *s, int nb, const double *nuk) {
  ARGCHK(s && nb >= 0 && (nb == 0 || nuk), "bad hd_set_sgs_spectral arguments");
  if (s->sgs_nuk) HIPCHK(hipFree(s->sgs_nuk));
  s->sgs_nuk = nullptr;
  s->sgs_nb = 0;
  if (nb == 0) return GHX_OK;
  int rc = sgs_buffers(s);
  if (rc) return rc;
  HIPCHK(hipMalloc(&s->sgs_nuk, nb * sizeof(double)));
  HIPCHK(hipMemcpy(s->sgs_nuk, nuk, nb * sizeof(double), hipMemcpyHostToDevice));
  s->sgs_nb = nb;
  return GHX_OK;
}

static int sgs_buffers(ghx_hd *s) {
  ghx_grid *g = s->g;
  if (!s->sgs_mem) {
    HIPCHK(hipMalloc(&s->sgs_mem, 15 * g->nr * sizeof(double) + 9 * g->nc * sizeof(cplx)));
    const size_t nl = (size_t)g->gk.nxh * g->gk.ny;
    HIPCHK(hipMalloc(&s->sgs_fline, nl));
    HIPCHK(hipMemset(s->sgs_fline, 1, nl));  // the subgrid force reaches every line of the dealias set
  }
  return GHX_OK;
}

static int hd_step_once(ghx_hd *s) {
  if (!s->clean) {
    double n = 0;
    int rc = count_unkept(s->g, s->v[0], s->v[1], s->v[2], &n);
    if (rc) return rc;
    s->clean = (n == 0.0);
  }
  const int prune = s->clean && !getenv("GHX_NO_PRUNE") ? 1 : 0;
  const cplx *fc[3] = {s->f[0], s->f[1], s->f[2]};
  const cplx *const *f = s->forced ? fc : nullptr;
  const cplx *v0[3] = {s->v[0], s->v[1], s->v[2]};
  const cplx *v1[3] = {s->v1[0], s->v1[1], s->v1[2]};
  if (s->tr) {
    int rc = tracers_begin_step(s->tr);
    if (rc) return rc;
  }
  for (int o = s->ord; o >= 1; --o) {
    const bool first = (o == s->ord), last = (o == 1);
    const cplx *const *vin = first ? v0 : v1;
    cplx *const *out = last ? s->v : s->v1;
    const cplx *const *fo = f;
    const uint8_t *fl = s->fline;
    const cplx *fs[3];
    if (s->sgs || s->sgs_nuk) {  // the forcing of this substep is f + P(-div tau(vin)) - nu(k) kk2 vin
      int rc = sgs_force(s, vin, fs, prune);
      if (rc) return rc;
      fo = fs;
      fl = s->sgs_fline;
    }
    int rc = hd_substep(s->g, s->q, o, s->dt, s->nu, vin, v0, fo, fl, out, prune, s->omegaz);
    if (rc) return rc;
    // main3D.fpp:3233-3267: the tracers take a stage on the velocity this
    // substep produced, which is dealiased whatever the input was
    if (s->tr) {
      const cplx *vo[3] = {out[0], out[1], out[2]};
      rc = tracers_stage(s->tr, s->q, vo, s->dt, 1.0 / (double)o, getenv("GHX_NO_PRUNE") ? 0 : 1);
      if (rc) return rc;
    }
  }
  s->clean = true;  // every substep zeroes the modes outside the dealias set
  return GHX_OK;
}

extern "C" int ghx_hd_step(ghx_hd *s, int nsteps) {
  ARGCHK(s && nsteps >= 0, "bad hd_step arguments");
  for (int n = 0; n < nsteps; ++n) {
    int rc = hd_step_once(s);
    if (rc) return rc;
  }
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// MHD
// ---------------------------------------------------------------------------
struct ghx_mhd {
  ghx_grid *g;
  int ord;
  double dt, nu, mu;
  cplx *mem = nullptr;  // s(6) s1(6) f(3) m(3), 24 scratch slots
  cplx *s[6], *s1[6], *f[3], *m[3], *q;
  uint8_t *fline = nullptr;
  bool forced = false;
  bool clean = false;
  ghx_tracers *tr = nullptr;
};

extern "C" int ghx_mhd_create(ghx_mhd **out, ghx_grid *g, int ord, double dt, double nu, double mu) {
  ARGCHK(out && g && ord >= 1, "bad mhd_create arguments");
  ghx_mhd *h = new (std::nothrow) ghx_mhd();
  if (!h) return set_err(GHX_ERR_NOMEM, "host allocation failed");
  h->g = g; h->ord = ord; h->dt = dt; h->nu = nu; h->mu = mu;
  (void)prune_shares(g);
  const size_t nc = g->nc;
  hipError_t e = hipMalloc(&h->mem, (18 * nc + 24 * ghx::slot_points(g->gk)) * sizeof(cplx));
  if (e != hipSuccess) { delete h; HIPCHK(e); }
  for (int k = 0; k < 6; ++k) { h->s[k] = h->mem + k * nc; h->s1[k] = h->mem + (6 + k) * nc; }
  for (int k = 0; k < 3; ++k) { h->f[k] = h->mem + (12 + k) * nc; h->m[k] = h->mem + (15 + k) * nc; }
  h->q = h->mem + 18 * nc;  // 24 scratch slots
  e = hipMalloc(&h->fline, (size_t)g->gk.nxh * g->gk.ny);
  if (e != hipSuccess) { (void)hipFree(h->mem); delete h; HIPCHK(e); }
  e = ghx::op_zero(g->stream, h->mem, 18 * nc * sizeof(cplx));
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) { (void)hipFree(h->mem); (void)hipFree(h->fline); delete h; HIPCHK(e); }
  *out = h;
  return GHX_OK;
}

extern "C" int ghx_mhd_destroy(ghx_mhd *s) {
  if (!s) return GHX_OK;
  (void)hipFree(s->mem);
  (void)hipFree(s->fline);
  delete s;
  return GHX_OK;
}

extern "C" int ghx_mhd_fields(ghx_mhd *s, void **v, void **a, void **f, void **m) {
  ARGCHK(s, "null solver");
  s->clean = false;
  for (int k = 0; k < 3; ++k) {
    if (v) v[k] = s->s[k];
    if (a) a[k] = s->s[3 + k];
    if (f) f[k] = s->f[k];
    if (m) m[k] = s->m[k];
  }
  return GHX_OK;
}

__global__ void k_or_flags(uint8_t *a, const uint8_t *b, size_t n) {
  for (size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (size_t)gridDim.x * blockDim.x)
    a[l] = a[l] | b[l];
}

extern "C" int ghx_mhd_set_forcing(ghx_mhd *s, int enabled) {
  ARGCHK(s, "null solver");
  s->forced = enabled != 0;
  if (s->forced) {
    ghx_grid *g = s->g;
    size_t nl = (size_t)g->gk.nxh * g->gk.ny;
    int rc = grid_work(g, nl / 8 + 2);
    if (rc) return rc;
    uint8_t *tmp = (uint8_t *)g->work;
    if ((rc = line_flags(g, s->f[0], s->f[1], s->f[2], s->fline))) return rc;
    if ((rc = line_flags(g, s->m[0], s->m[1], s->m[2], tmp))) return rc;
    hipLaunchKernelGGL(k_or_flags, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, g->stream, s->fline, tmp, nl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g->stream));
  }
  return GHX_OK;
}

// include/mhd_rkstep2.f90:4-65 fused: one inverse chain for (v, A) producing
// (u, w, b, j); three cross products (w x u, j x b, u x b); one forward chain
// ending in nonlin3 / gauge3 and the update.
extern "C" int ghx_mhd_touch(ghx_mhd *s) {
  ARGCHK(s, "null solver");
  s->clean = false;
  return GHX_OK;
}

static int mhd_substep(ghx_mhd *s, int o, const cplx *const *vin, cplx *const *out, int prune) {
  ghx_grid *g = s->g;
  const size_t nc = g->nc, slot = ghx::slot_points(g->gk);
  const FFTCtx &ctx = g->plan.ctx;
  // Q12 (slots 0-11) -> P12 (12-23) -> P9 (reuses 0-8) -> Q9 (reuses 12-20)
  FieldSet in, q12, p12, p9, q9, pq;
  for (int k = 0; k < 6; ++k) in.p[k] = const_cast<cplx *>(vin[k]);
  for (int k = 0; k < 12; ++k) { q12.p[k] = s->q + k * slot; p12.p[k] = s->q + (12 + k) * slot; }
  for (int k = 0; k < 9; ++k) { p9.p[k] = s->q + k * slot; q9.p[k] = s->q + (12 + k) * slot; }
  const double sc = (double)nc * sizeof(cplx);
  const PruneShares ps = (prune && g_prof.on) ? prune_shares(g) : PruneShares{1.0, 1.0, 1.0};
  PROF(P_ZINV_MHD, (6 * ps.fm + 12 * ps.fl) * sc, ctx.stream, ghx::zinv(ctx, 2, 6, in, q12, prune));  // u w b j
  // the Y^-1 pass takes at most kMaxFields fields per launch
  PROF(P_YINV, (6 * ps.fl + 6 * ps.fx) * sc, ctx.stream, ghx::yinv(ctx, 6, q12, p12, prune));
  {
    FieldSet qa, pa;
    for (int k = 0; k < 6; ++k) { qa.p[k] = q12.p[6 + k]; pa.p[k] = p12.p[6 + k]; }
    PROF(P_YINV, (6 * ps.fl + 6 * ps.fx) * sc, ctx.stream, ghx::yinv(ctx, 6, qa, pa, prune));
  }
  FieldSet o3;
  // w x u
  for (int k = 0; k < 3; ++k) { pq.p[k] = p12.p[3 + k]; pq.p[3 + k] = p12.p[k]; o3.p[k] = p9.p[k]; }
  PROF(P_XCROSS, 9 * ps.fx * sc, ctx.stream, ghx::xcross(ctx, 0, pq, o3, prune));
  // j x b
  for (int k = 0; k < 3; ++k) { pq.p[k] = p12.p[9 + k]; pq.p[3 + k] = p12.p[6 + k]; o3.p[k] = p9.p[3 + k]; }
  PROF(P_XCROSS, 9 * ps.fx * sc, ctx.stream, ghx::xcross(ctx, 0, pq, o3, prune));
  // u x b
  for (int k = 0; k < 3; ++k) { pq.p[k] = p12.p[k]; pq.p[3 + k] = p12.p[6 + k]; o3.p[k] = p9.p[6 + k]; }
  PROF(P_XCROSS, 9 * ps.fx * sc, ctx.stream, ghx::xcross(ctx, 0, pq, o3, prune));
  PROF(P_YFWD, (9 * ps.fx + 9 * ps.fl) * sc, ctx.stream, ghx::yfwd(ctx, 9, p9, q9, prune));
  ghx::MHDRKParams rk;
  rk.dt = s->dt; rk.nu = s->nu; rk.mu = s->mu; rk.rmp = 1. / (double)o;
  for (int k = 0; k < 6; ++k) {
    rk.v[k] = vin[k];
    rk.c[k] = s->s[k];
    rk.out[k] = out[k];
  }
  for (int k = 0; k < 3; ++k) {
    rk.f[k] = s->forced ? s->f[k] : nullptr;
    rk.f[3 + k] = s->forced ? s->m[k] : nullptr;
  }
  rk.fline = s->forced ? s->fline : nullptr;
  const int nvc = vin[0] == s->s[0] ? 6 : 12;
  PROF(P_ZFWD_RK_MHD, (9 * ps.fl + nvc * ps.fm + 6 * (prune ? ps.fm : 1.0)) * sc, ctx.stream,
       ghx::zfwd_rk_mhd(ctx, q9, rk, prune));
  return GHX_OK;
}

extern "C" int ghx_mhd_step(ghx_mhd *s, int nsteps) {
  ARGCHK(s && nsteps >= 0, "bad mhd_step arguments");
  const cplx *v0[6], *v1[6];
  for (int k = 0; k < 6; ++k) { v0[k] = s->s[k]; v1[k] = s->s1[k]; }
  for (int n = 0; n < nsteps; ++n) {
    if (!s->clean) {
      double a = 0, b = 0;
      int rc = count_unkept(s->g, s->s[0], s->s[1], s->s[2], &a);
      if (!rc) rc = count_unkept(s->g, s->s[3], s->s[4], s->s[5], &b);
      if (rc) return rc;
      s->clean = (a == 0.0 && b == 0.0);
    }
    const int prune = s->clean && !getenv("GHX_NO_PRUNE") ? 1 : 0;
    if (s->tr) {
      int rc = tracers_begin_step(s->tr);
      if (rc) return rc;
    }
    for (int o = s->ord; o >= 1; --o) {
      const bool first = (o == s->ord), last = (o == 1);
      cplx *const *out = last ? s->s : s->s1;
      int rc = mhd_substep(s, o, first ? v0 : v1, out, prune);
      if (rc) return rc;
      if (s->tr) {  // main3D.fpp:3233-3267, velocity after the substep
        const cplx *vo[3] = {out[0], out[1], out[2]};
        rc = tracers_stage(s->tr, s->q, vo, s->dt, 1.0 / (double)o, getenv("GHX_NO_PRUNE") ? 0 : 1);
        if (rc) return rc;
      }
    }
    s->clean = true;
  }
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// Lagrangian tracers (lagpart/gparts_mod.f90, gpsplineint_mod.fpp), one rank
// ---------------------------------------------------------------------------
struct ghx_tracers {
  ghx_grid *g;
  long long n;
  int zcompat;
  double *mem = nullptr;  // pos(3n) x0(3n) lv(3n)
  double *pos[3], *x0[3], *lv[3];
  double *spl = nullptr;  // spline prefilter factors: nx/2+1 + ny + nz
};

// Spectral form of GPSplineInt_CompSpline3D (gpsplineint_mod.fpp:1158-1579):
// the periodic spline system (1/6, 2/3, 1/6) along each axis is the circulant
// whose symbol is b(k) = 2/3 + cos(2 pi k / n) / 3, so the control points of
// c2r(v / (nx ny nz)) are c2r(v sx[kx] sy[ky] sz[kz]) with s = 1 / b and the
// 1 / (nx ny nz) folded into sx.
static int tracer_prefilter(ghx_tracers *t) {
  const GridK &k = t->g->gk;
  std::vector<double> h((size_t)k.nxh + k.ny + k.nz);
  const long double twopi = 6.283185307179586476925286766559L;
  auto inv_b = [&](int q, int n) {
    return 1.0L / (2.0L / 3.0L + cosl(twopi * (long double)q / (long double)n) / 3.0L);
  };
  const long double rmp = 1.0L / ((long double)k.nx * k.ny * k.nz);
  for (int i = 0; i < k.nxh; ++i) h[i] = (double)(rmp * inv_b(i, k.nx));
  for (int j = 0; j < k.ny; ++j) h[k.nxh + j] = (double)inv_b(j, k.ny);
  for (int q = 0; q < k.nz; ++q) h[k.nxh + k.ny + q] = (double)inv_b(q, k.nz);
  HIPCHK(hipMalloc(&t->spl, h.size() * sizeof(double)));
  HIPCHK(hipMemcpy(t->spl, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  return GHX_OK;
}

extern "C" int ghx_tracers_create(ghx_tracers **out, ghx_grid *g, long long n, int zcompat) {
  ARGCHK(out && g && n >= 0, "bad tracers_create arguments");
  if (g->gk.nx < 4 || g->gk.ny < 4 || g->gk.nz < 4)
    return set_err(GHX_ERR_SIZE, "cubic-spline tracers need at least 4 points along each axis");
  ghx_tracers *t = new (std::nothrow) ghx_tracers();
  if (!t) return set_err(GHX_ERR_NOMEM, "host allocation failed");
  t->g = g; t->n = n; t->zcompat = zcompat ? 1 : 0;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  hipError_t e = hipMalloc(&t->mem, 9 * nn * sizeof(double));
  if (e != hipSuccess) { delete t; HIPCHK(e); }
  for (int c = 0; c < 3; ++c) {
    t->pos[c] = t->mem + c * nn;
    t->x0[c] = t->mem + (3 + c) * nn;
    t->lv[c] = t->mem + (6 + c) * nn;
  }
  int rc = tracer_prefilter(t);
  if (!rc) {
    e = ghx::op_zero(g->stream, t->mem, 9 * nn * sizeof(double));
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = set_err(GHX_ERR_HIP, "zeroing tracer arrays: %s", hipGetErrorString(e));
  }
  if (rc) {
    (void)hipFree(t->mem);
    if (t->spl) (void)hipFree(t->spl);
    delete t;
    return rc;
  }
  *out = t;
  return GHX_OK;
}

extern "C" int ghx_tracers_destroy(ghx_tracers *t) {
  if (!t) return GHX_OK;
  (void)hipFree(t->mem);
  (void)hipFree(t->spl);
  delete t;
  return GHX_OK;
}

extern "C" int ghx_tracers_positions(ghx_tracers *t, double **px, double **py, double **pz) {
  ARGCHK(t, "null tracers");
  if (px) *px = t->pos[0];
  if (py) *py = t->pos[1];
  if (pz) *pz = t->pos[2];
  return GHX_OK;
}

extern "C" int ghx_tracers_velocities(ghx_tracers *t, double **ux, double **uy, double **uz) {
  ARGCHK(t, "null tracers");
  if (ux) *ux = t->lv[0];
  if (uy) *uy = t->lv[1];
  if (uz) *uz = t->lv[2];
  return GHX_OK;
}

// GPart_SetStepRKK (gparts_mod.f90:1839-1908): x0 = x.
static int tracers_begin_step(ghx_tracers *t) {
  const double *src[3] = {t->pos[0], t->pos[1], t->pos[2]};
  HIPCHK(ghx::tracer_copy(t->g->stream, t->x0, src, t->n));
  return GHX_OK;
}

extern "C" int ghx_tracers_set_step(ghx_tracers *t) {
  ARGCHK(t, "null tracers");
  return tracers_begin_step(t);
}

// One stage on 9 scratch slots of q: Q (0-2) -> P (3-5) -> real control
// points of the three components interleaved (slots 6-8, contiguous), then
// the interpolation and position update.
static int tracers_stage(ghx_tracers *t, cplx *q, const cplx *const *v, double dt, double xk, int prune) {
  ghx_grid *g = t->g;
  const FFTCtx &ctx = g->plan.ctx;
  const size_t slot = ghx::slot_points(g->gk);
  FieldSet in, qq, pp;
  double *r3 = reinterpret_cast<double *>(q + 6 * slot);
  for (int c = 0; c < 3; ++c) {
    in.p[c] = const_cast<cplx *>(v[c]);
    qq.p[c] = q + c * slot;
    pp.p[c] = q + (3 + c) * slot;
  }
  const double sc = (double)g->nc * sizeof(cplx), sr = (double)g->nr * sizeof(double);
  const PruneShares ps = (prune && g_prof.on) ? prune_shares(g) : PruneShares{1.0, 1.0, 1.0};
  PROF(P_ZINV_SPLINE, (3 * ps.fm + 3 * ps.fl) * sc, ctx.stream, ghx::zinv(ctx, 3, 3, in, qq, prune, t->spl));
  PROF(P_YINV_TRACER, (3 * ps.fl + 3 * ps.fx) * sc, ctx.stream, ghx::yinv(ctx, 3, qq, pp, prune));
  PROF(P_XINV_TRACER, 3 * (ps.fx * sc + sr), ctx.stream, ghx::xinv3i(ctx, pp, r3, prune));
  ghx::TracerStage st;
  const int n[3] = {g->gk.nx, g->gk.ny, g->gk.nz};
  st.coef3 = r3;
  for (int c = 0; c < 3; ++c) {
    st.pos[c] = t->pos[c];
    st.x0[c] = t->x0[c];
    st.lv[c] = t->lv[c];
    // GPart_StepRKK: dtfact = dt * xk * invdel, invdel = n / (2 pi L) (gparts_mod.f90:207-213)
    const double pi = 3.14159265358979323846;
    st.dtf[c] = dt * xk * ((double)n[c] / (2 * pi * g->L[c]));
  }
  st.nx = n[0]; st.ny = n[1]; st.nz = n[2];
  st.np = t->n;
  st.zcompat = t->zcompat;
  // per particle: position and x0 read, 3 x 64 control points, position and velocity written
  PROF(P_TRACER_STAGE, (double)t->n * (24 + 24 + 3 * 64 * 8 + 24 + 24), ctx.stream, ghx::tracer_stage(ctx.stream, st));
  return GHX_OK;
}

extern "C" int ghx_tracers_step(ghx_tracers *t, const void *const *v, double dt, double xk) {
  ARGCHK(t && v && v[0] && v[1] && v[2], "bad tracers_step arguments");
  int rc = plan_scratch(&t->g->plan, 9);
  if (rc) return rc;
  const cplx *vv[3] = {C_(v[0]), C_(v[1]), C_(v[2])};
  return tracers_stage(t, t->g->plan.q, vv, dt, xk, 0);
}

// GPart_EulerToLag (gparts_mod.f90:2216-2269) of one real field r[z][y][x].
extern "C" int ghx_tracers_euler_to_lag(ghx_tracers *t, const double *field, double *out) {
  ARGCHK(t && field && out, "bad tracers_euler_to_lag arguments");
  ghx_grid *g = t->g;
  int rc = grid_ctmp(g, 1);
  if (!rc) rc = plan_scratch(&g->plan, 3);
  if (rc) return rc;
  void *spec = g->ctmp;
  if ((rc = do_r2c(&g->plan, 1, &field, &spec))) return rc;
  const FFTCtx &ctx = g->plan.ctx;
  const size_t slot = ghx::slot_points(g->gk);
  FieldSet in, qq, pp;
  in.p[0] = g->ctmp;
  qq.p[0] = g->plan.q;
  pp.p[0] = g->plan.q + slot;
  double *r = reinterpret_cast<double *>(g->plan.q + 2 * slot);
  HIPCHK(ghx::zinv(ctx, 3, 1, in, qq, 0, t->spl));
  HIPCHK(ghx::yinv(ctx, 1, qq, pp, 0));
  HIPCHK(ghx::xinv(ctx, 1, pp, &r, 0));
  const double *pos[3] = {t->pos[0], t->pos[1], t->pos[2]};
  HIPCHK(ghx::spline_interp(ctx.stream, r, pos, t->n, g->gk.nx, g->gk.ny, g->gk.nz, t->zcompat, out));
  return GHX_OK;
}

extern "C" int ghx_hd_set_tracers(ghx_hd *s, ghx_tracers *t) {
  ARGCHK(s, "null solver");
  ARGCHK(!t || t->g == s->g, "tracers and solver must share the grid");
  s->tr = t;
  return GHX_OK;
}

extern "C" int ghx_mhd_set_tracers(ghx_mhd *s, ghx_tracers *t) {
  ARGCHK(s, "null solver");
  ARGCHK(!t || t->g == s->g, "tracers and solver must share the grid");
  s->tr = t;
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// Boussinesq (include/bouss_rkstep{1,2}.f90)
// ---------------------------------------------------------------------------
struct ghx_bouss {
  ghx_grid *g;
  int ord;
  double dt, nu, kappa, xmom, xtemp;
  cplx *mem = nullptr;  // s(4) s1(4) f(4), 18 scratch slots
  cplx *s[4], *s1[4], *f[4], *q;
  uint8_t *fline = nullptr;
  bool forced = false;
  bool clean = false;
  ghx_tracers *tr = nullptr;
  // the rotating solver (include/rotbouss_rkstep2.f90): Coriolis force of
  // rotation omegaz about z, and no ord = 3 correction
  bool rot = false;
  double omegaz = 0.0;
  cplx *corr = nullptr;  // the ord = 3 correction term, 3 fields
};

extern "C" int ghx_bouss_create(ghx_bouss **out, ghx_grid *g, int ord, double dt, double nu, double kappa,
                                double xmom, double xtemp) {
  ARGCHK(out && g, "bad bouss_create arguments");
  ARGCHK(ord >= 1 && ord <= 3, "Boussinesq solver: ord 1 to 3");
  ghx_bouss *h = new (std::nothrow) ghx_bouss();
  if (!h) return set_err(GHX_ERR_NOMEM, "host allocation failed");
  h->g = g; h->ord = ord; h->dt = dt; h->nu = nu; h->kappa = kappa; h->xmom = xmom; h->xtemp = xtemp;
  (void)prune_shares(g);
  const size_t nc = g->nc;
  hipError_t e = hipMalloc(&h->mem, (12 * nc + 18 * ghx::slot_points(g->gk)) * sizeof(cplx));
  if (e != hipSuccess) { delete h; HIPCHK(e); }
  for (int k = 0; k < 4; ++k) {
    h->s[k] = h->mem + k * nc;
    h->s1[k] = h->mem + (4 + k) * nc;
    h->f[k] = h->mem + (8 + k) * nc;
  }
  h->q = h->mem + 12 * nc;
  e = hipMalloc(&h->fline, (size_t)g->gk.nxh * g->gk.ny);
  if (e == hipSuccess) e = ghx::op_zero(g->stream, h->mem, 12 * nc * sizeof(cplx));
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) { (void)hipFree(h->mem); (void)hipFree(h->fline); delete h; HIPCHK(e); }
  *out = h;
  return GHX_OK;
}

extern "C" int ghx_bouss_destroy(ghx_bouss *s) {
  if (!s) return GHX_OK;
  (void)hipFree(s->corr);
  (void)hipFree(s->mem);
  (void)hipFree(s->fline);
  delete s;
  return GHX_OK;
}

extern "C" int ghx_bouss_fields(ghx_bouss *s, void **v, void **th, void **f, void **fs) {
  ARGCHK(s, "null solver");
  s->clean = false;
  for (int k = 0; k < 3; ++k) {
    if (v) v[k] = s->s[k];
    if (f) f[k] = s->f[k];
  }
  if (th) *th = s->s[3];
  if (fs) *fs = s->f[3];
  return GHX_OK;
}

extern "C" int ghx_bouss_touch(ghx_bouss *s) {
  ARGCHK(s, "null solver");
  s->clean = false;
  return GHX_OK;
}

extern "C" int ghx_bouss_set_rotation(ghx_bouss *s, double omegaz) {
  ARGCHK(s, "null solver");
  s->rot = true;
  s->omegaz = omegaz;
  return GHX_OK;
}

extern "C" int ghx_bouss_set_forcing(ghx_bouss *s, int enabled) {
  ARGCHK(s, "null solver");
  s->forced = enabled != 0;
  if (s->forced) {
    ghx_grid *g = s->g;
    size_t nl = (size_t)g->gk.nxh * g->gk.ny;
    int rc = grid_work(g, nl / 8 + 2);
    if (rc) return rc;
    uint8_t *tmp = (uint8_t *)g->work;
    if ((rc = line_flags(g, s->f[0], s->f[1], s->f[2], s->fline))) return rc;
    if ((rc = line_flags(g, s->f[3], s->f[3], s->f[3], tmp))) return rc;
    hipLaunchKernelGGL(k_or_flags, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, g->stream, s->fline, tmp, nl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g->stream));
  }
  return GHX_OK;
}

extern "C" int ghx_bouss_set_tracers(ghx_bouss *s, ghx_tracers *t) {
  ARGCHK(s, "null solver");
  ARGCHK(!t || t->g == s->g, "tracers and solver must share the grid");
  s->tr = t;
  return GHX_OK;
}

// include/bouss_rkstep2.f90:1-71 fused: one inverse chain for (v, th)
// producing (u, w, grad th); w x u and -(u . grad th) in the x pass; one
// forward chain ending in buoyancy, projection, heat current and the updates.
static int bouss_substep(ghx_bouss *s, int o, const cplx *const *vin, cplx *const *out, int prune) {
  ghx_grid *g = s->g;
  const size_t nc = g->nc, slot = ghx::slot_points(g->gk);
  const FFTCtx &ctx = g->plan.ctx;
  // Q9 (slots 0-8) -> P9 (9-17) -> P4 (reuses 0-3) -> Q4 (reuses 9-12)
  FieldSet in, q9, p9, pq, o3, pd, o1, p4, q4;
  for (int k = 0; k < 4; ++k) in.p[k] = const_cast<cplx *>(vin[k]);
  for (int k = 0; k < 9; ++k) { q9.p[k] = s->q + k * slot; p9.p[k] = s->q + (9 + k) * slot; }
  for (int k = 0; k < 4; ++k) { p4.p[k] = s->q + k * slot; q4.p[k] = s->q + (9 + k) * slot; }
  const double sc = (double)nc * sizeof(cplx);
  const PruneShares ps = (prune && g_prof.on) ? prune_shares(g) : PruneShares{1.0, 1.0, 1.0};
  PROF(P_ZINV_BOUSS, (4 * ps.fm + 9 * ps.fl) * sc, ctx.stream, ghx::zinv(ctx, 5, 4, in, q9, prune));
  PROF(P_YINV, (9 * ps.fl + 9 * ps.fx) * sc, ctx.stream, ghx::yinv(ctx, 9, q9, p9, prune));
  // w x u in GHOST's operand order (P = w, Q = u)
  for (int k = 0; k < 3; ++k) { pq.p[k] = p9.p[3 + k]; pq.p[3 + k] = p9.p[k]; o3.p[k] = p4.p[k]; }
  PROF(P_XCROSS, 9 * ps.fx * sc, ctx.stream, ghx::xcross(ctx, 0, pq, o3, prune));
  for (int k = 0; k < 3; ++k) { pd.p[k] = p9.p[k]; pd.p[3 + k] = p9.p[6 + k]; }
  o1.p[0] = p4.p[3];
  PROF(P_XDOT, 7 * ps.fx * sc, ctx.stream, ghx::xdot(ctx, pd, o1, prune));
  PROF(P_YFWD, (4 * ps.fx + 4 * ps.fl) * sc, ctx.stream, ghx::yfwd(ctx, 4, p4, q4, prune));
  ghx::BoussRKParams rk;
  rk.dt = s->dt; rk.nu = s->nu; rk.kappa = s->kappa; rk.xmom = s->xmom; rk.xtemp = s->xtemp;
  rk.rmp = 1. / (double)o;
  rk.om2 = s->rot ? 2 * s->omegaz : 0.0;
  for (int k = 0; k < 4; ++k) {
    rk.v[k] = vin[k];
    rk.c[k] = s->s[k];
    rk.out[k] = out[k];
    rk.f[k] = s->forced ? s->f[k] : nullptr;
  }
  rk.fline = s->forced ? s->fline : nullptr;
  const int nvc = vin[0] == s->s[0] ? 4 : 8;
  PROF(P_ZFWD_RK_BOUSS, (4 * ps.fl + nvc * ps.fm + 4 * (prune ? ps.fm : 1.0)) * sc, ctx.stream,
       ghx::zfwd_rk_bouss(ctx, q4, rk, prune));
  return GHX_OK;
}

// The correction term of include/bouss_rkstep2.f90:73-145 from the step-start
// state c = (u, th): x = N(L(u) + N(u)) with N(w) = nonlhd3 of prodre3(w) plus
// the buoyancy xmom th, L the Laplacian (no viscosity, as written there) and
// the sum masked to the kept modes.  The caller adds 2 x dt^3/24 to the
// velocity after the last substep.  x may not alias c.
static int bouss_ord3_term(ghx_grid *g, double xmom, const cplx *const *c, cplx *const *x) {
  int rc = fused_cross(g, 1, c[0], c[1], c[2], nullptr, nullptr, nullptr, x[0], x[1], x[2]);
  if (rc) return rc;
  const cplx *w[3] = {x[0], x[1], x[2]};
  HIPCHK(ghx::op_bouss3_nonl(g->gk, g->stream, 0, w, c[3], xmom, c, x));
  if ((rc = fused_cross(g, 1, x[0], x[1], x[2], nullptr, nullptr, nullptr, x[0], x[1], x[2]))) return rc;
  HIPCHK(ghx::op_bouss3_nonl(g->gk, g->stream, 1, w, c[3], xmom, nullptr, x));
  return GHX_OK;
}

// include/bouss_rkstep2.f90:73-145 on caller fields, for the last substep
// (o = 1) of an ord = 3 step: c holds (v_x, v_y, v_z, th) at the step start,
// v (3 fields) the velocity after the substep, corrected in place.  c is not
// changed (GHOST leaves L(u) + N(u) in C1-C3, which nothing reads before the
// next step's copy).
extern "C" int ghx_bouss_ord3(ghx_grid *g, double dt, double xmom, const void *const *c, void *const *v) {
  ARGCHK(g && c && v, "bad bouss_ord3 arguments");
  const size_t nc = g->nc;
  int rc = grid_work(g, 6 * nc);
  if (rc) return rc;
  cplx *w = (cplx *)g->work;
  cplx *x[3] = {w, w + nc, w + 2 * nc};
  const cplx *ci[4] = {C_(c[0]), C_(c[1]), C_(c[2]), C_(c[3])};
  for (int k = 0; k < 3; ++k)
    for (int m = 0; m < 4; ++m) ARGCHK(v[k] != c[m], "bouss_ord3: v may not alias c");
  if ((rc = bouss_ord3_term(g, xmom, ci, x))) return rc;
  cplx *vo[3] = {M_(v[0]), M_(v[1]), M_(v[2])};
  const cplx *xi[3] = {x[0], x[1], x[2]};
  HIPCHK(ghx::op_bouss3_add(g->gk, g->stream, vo, xi, dt * dt * dt / 24.0));
  return GHX_OK;
}

// Forcing-line flags for the stateless substeps below: the lines (kx, ky)
// where any of the nf fields is non-zero, OR-ed, in g->work (2 nl bytes).
static int stateless_flags(ghx_grid *g, const cplx *const *fa, int nf, uint8_t **flags) {
  const size_t nl = (size_t)g->gk.nxh * g->gk.ny;
  int rc = grid_work(g, (2 * nl + 7) / 8 + 1);
  if (rc) return rc;
  uint8_t *fl = (uint8_t *)g->work, *tmp = fl + nl;
  for (int b = 0; b < nf; b += 3) {
    const cplx *x = fa[b], *y = fa[b + 1 < nf ? b + 1 : b], *z = fa[b + 2 < nf ? b + 2 : b];
    if ((rc = line_flags(g, x, y, z, b == 0 ? fl : tmp))) return rc;
    if (b > 0) {
      hipLaunchKernelGGL(k_or_flags, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, g->stream, fl, tmp, nl);
      HIPCHK(hipGetLastError());
    }
  }
  *flags = fl;
  return GHX_OK;
}

// include/mhd_rkstep2.f90:4-65 for order index o on caller fields: vin, C
// and vout hold (v_x, v_y, v_z, a_x, a_y, a_z); f (mechanical) and m
// (electromotive) forcing, 3 fields each, both given or both NULL.  vout may
// equal vin or C.
extern "C" int ghx_mhd_rkstep2(ghx_grid *g, int o, double dt, double nu, double mu, const void *const *vin,
                               const void *const *C, const void *const *f, const void *const *m, void *const *vout) {
  ARGCHK(g && o >= 1 && vin && C && vout && (f == nullptr) == (m == nullptr), "bad mhd_rkstep2 arguments");
  int rc = plan_scratch(&g->plan, 24);
  if (rc) return rc;
  ghx_mhd h;
  h.g = g; h.ord = o; h.dt = dt; h.nu = nu; h.mu = mu;
  h.q = g->plan.q;
  const cplx *vi[6];
  cplx *vo[6];
  for (int k = 0; k < 6; ++k) { h.s[k] = M_(C[k]); vi[k] = C_(vin[k]); vo[k] = M_(vout[k]); }
  if (f) {
    for (int k = 0; k < 3; ++k) { h.f[k] = M_(f[k]); h.m[k] = M_(m[k]); }
    const cplx *fa[6] = {h.f[0], h.f[1], h.f[2], h.m[0], h.m[1], h.m[2]};
    if ((rc = stateless_flags(g, fa, 6, &h.fline))) return rc;
    h.forced = true;
  }
  rc = mhd_substep(&h, o, vi, vo, 0);
  h.fline = nullptr;  // g->work, not owned
  return rc;
}

// include/bouss_rkstep2.f90:1-71 (rotbouss_rkstep2.f90 with rot = 1: the
// Coriolis force of rotation omegaz about z) for order index o on caller
// fields: vin, C and vout hold (v_x, v_y, v_z, th); f holds (f_x, f_y, f_z,
// f_s) or is NULL.  The ord = 3 correction of bouss_rkstep2.f90:73-145 is not
// applied here.
static int bouss_rkstep2_impl(ghx_grid *g, int o, double dt, double nu, double kappa, double xmom, double xtemp,
                              bool rot, double omegaz, const void *const *vin, const void *const *C,
                              const void *const *f, void *const *vout) {
  ARGCHK(g && o >= 1 && vin && C && vout, "bad bouss_rkstep2 arguments");
  int rc = plan_scratch(&g->plan, 18);
  if (rc) return rc;
  ghx_bouss h;
  h.g = g; h.ord = o; h.dt = dt; h.nu = nu; h.kappa = kappa; h.xmom = xmom; h.xtemp = xtemp;
  h.rot = rot; h.omegaz = omegaz;
  h.q = g->plan.q;
  const cplx *vi[4];
  cplx *vo[4];
  for (int k = 0; k < 4; ++k) { h.s[k] = M_(C[k]); vi[k] = C_(vin[k]); vo[k] = M_(vout[k]); }
  if (f) {
    for (int k = 0; k < 4; ++k) h.f[k] = M_(f[k]);
    const cplx *fa[4] = {h.f[0], h.f[1], h.f[2], h.f[3]};
    if ((rc = stateless_flags(g, fa, 4, &h.fline))) return rc;
    h.forced = true;
  }
  rc = bouss_substep(&h, o, vi, vo, 0);
  h.fline = nullptr;
  return rc;
}
extern "C" int ghx_bouss_rkstep2(ghx_grid *g, int o, double dt, double nu, double kappa, double xmom, double xtemp,
                                 const void *const *vin, const void *const *C, const void *const *f,
                                 void *const *vout) {
  return bouss_rkstep2_impl(g, o, dt, nu, kappa, xmom, xtemp, false, 0.0, vin, C, f, vout);
}
extern "C" int ghx_rotbouss_rkstep2(ghx_grid *g, int o, double dt, double nu, double kappa, double xmom,
                                    double xtemp, double omegaz, const void *const *vin, const void *const *C,
                                    const void *const *f, void *const *vout) {
  return bouss_rkstep2_impl(g, o, dt, nu, kappa, xmom, xtemp, true, omegaz, vin, C, f, vout);
}

extern "C" int ghx_bouss_step(ghx_bouss *s, int nsteps) {
  ARGCHK(s && nsteps >= 0, "bad bouss_step arguments");
  // include/bouss_rkstep2.f90:73-145: an ord = 3 step without rotation adds
  // 2 N(L(u) + N(u)) dt^3/24 after its last substep.  The term depends on the
  // step-start state only, so it is formed before the substeps overwrite it.
  // It is not masked, so the state is never clean afterwards.
  const bool ord3 = s->ord == 3 && !s->rot;
  if (ord3 && !s->corr) HIPCHK(hipMalloc(&s->corr, 3 * s->g->nc * sizeof(cplx)));
  cplx *corr[3] = {s->corr, s->corr + (ord3 ? s->g->nc : 0), s->corr + (ord3 ? 2 * s->g->nc : 0)};
  const cplx *v0[4], *v1[4];
  for (int k = 0; k < 4; ++k) { v0[k] = s->s[k]; v1[k] = s->s1[k]; }
  for (int n = 0; n < nsteps; ++n) {
    if (!s->clean) {
      double a = 0, b = 0;
      int rc = count_unkept(s->g, s->s[0], s->s[1], s->s[2], &a);
      if (!rc) rc = reduce(s->g, 5, s->s[3], s->s[3], s->s[3], nullptr, nullptr, nullptr, &b);
      if (rc) return rc;
      s->clean = (a == 0.0 && b == 0.0);
    }
    const int prune = s->clean && !getenv("GHX_NO_PRUNE") ? 1 : 0;
    if (s->tr) {
      int rc = tracers_begin_step(s->tr);
      if (rc) return rc;
    }
    for (int o = s->ord; o >= 1; --o) {
      const bool first = (o == s->ord), last = (o == 1);
      cplx *const *out = last ? s->s : s->s1;
      if (ord3 && first) {
        int rc = bouss_ord3_term(s->g, s->xmom, v0, corr);
        if (rc) return rc;
      }
      int rc = bouss_substep(s, o, first ? v0 : v1, out, prune);
      if (rc) return rc;
      if (ord3 && last) {
        const cplx *ci[3] = {corr[0], corr[1], corr[2]};
        HIPCHK(ghx::op_bouss3_add(s->g->gk, s->g->stream, out, ci, s->dt * s->dt * s->dt / 24.0));
      }
      if (s->tr) {  // main3D.fpp:3233-3267, velocity after the substep
        const cplx *vo[3] = {out[0], out[1], out[2]};
        // the corrected velocity is not confined to the kept modes
        const int tprune = (ord3 && last) || getenv("GHX_NO_PRUNE") ? 0 : 1;
        rc = tracers_stage(s->tr, s->q, vo, s->dt, 1.0 / (double)o, tprune);
        if (rc) return rc;
      }
    }
    s->clean = !ord3;
  }
  return GHX_OK;
}

// ---------------------------------------------------------------------------
// Ensembles: many independent HD runs advanced together.  Members of equal
// grid size form a group whose fields are laid out member after member, so
// each pass of the fused substep is one launch for the whole group (the
// member index is a grid dimension of the launch).  After the first step the
// whole ensemble step is captured in a HIP graph and replayed.
// ---------------------------------------------------------------------------
struct EnsGroup {
  int n = 0;
  ghx_grid *g = nullptr;
  std::vector<int> members;
  cplx *mem = nullptr;    // v [member][3][nc], v1 [member][3][nc], scratch [member][12][slot]
  double *nu = nullptr;   // per-member viscosity (device)
  size_t nc = 0, slot = 0;
  int tz = 0, ty = 0;     // tile widths of the z and y passes (0: built-in)
  int sched = 0;          // radix schedules of the z and y passes (FFTCtx::sched)
  bool clean = false;
  cplx *v() const { return mem; }
  cplx *v1() const { return mem + members.size() * 3 * nc; }
  cplx *q() const { return mem + members.size() * 6 * nc; }
};

struct ghx_ensemble {
  std::vector<EnsGroup> groups;
  std::vector<std::pair<int, int>> where;  // member -> (group, index in group)
  int ord = 2;
  double dt = 0.0;
  hipStream_t stream = nullptr;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the captured step forks one branch per size group: a stream and a join
  // event per group (default priority), and the fork event
  std::vector<hipStream_t> gstream;
  std::vector<hipEvent_t> gjoin;
  hipEvent_t fork = nullptr;
};

static void ens_drop_graph(ghx_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
}

extern "C" int ghx_ensemble_destroy(ghx_ensemble *e) {
  if (!e) return GHX_OK;
  ens_drop_graph(e);
  for (hipStream_t st : e->gstream) (void)hipStreamDestroy(st);
  for (hipEvent_t ev : e->gjoin) (void)hipEventDestroy(ev);
  if (e->fork) (void)hipEventDestroy(e->fork);
  for (EnsGroup &G : e->groups) {
    if (G.mem) (void)hipFree(G.mem);
    if (G.nu) (void)hipFree(G.nu);
    ghx_grid_destroy(G.g);
  }
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return GHX_OK;
}

extern "C" int ghx_ensemble_create(ghx_ensemble **out, int nmembers, const int *n, const double *nu, double dt,
                                   int ord) {
  ARGCHK(out && nmembers > 0 && n && nu && ord >= 1, "bad ensemble_create arguments");
  ghx_ensemble *e = new (std::nothrow) ghx_ensemble();
  if (!e) return set_err(GHX_ERR_NOMEM, "host allocation failed");
  e->ord = ord;
  e->dt = dt;
  e->where.resize(nmembers);
  int rc = GHX_OK;
  hipError_t he = hipStreamCreate(&e->stream);  // default priority
  if (he != hipSuccess) { delete e; HIPCHK(he); }
  for (int m = 0; m < nmembers && !rc; ++m) {
    int gi = -1;
    for (size_t k = 0; k < e->groups.size(); ++k)
      if (e->groups[k].n == n[m]) gi = (int)k;
    if (gi < 0) {
      EnsGroup G;
      G.n = n[m];
      rc = ghx_grid_create(&G.g, n[m], n[m], n[m], 1.0, 1.0, 1.0, 0, 0.0);
      if (rc) break;
      ghx_grid_set_stream(G.g, e->stream);
      e->groups.push_back(G);
      gi = (int)e->groups.size() - 1;
    }
    e->where[m] = {gi, (int)e->groups[gi].members.size()};
    e->groups[gi].members.push_back(m);
  }
  for (EnsGroup &G : e->groups) {
    if (rc) break;
    const size_t cnt = G.members.size();
    G.nc = G.g->nc;
    G.slot = ghx::slot_points(G.g->gk);
    (void)prune_shares(G.g);
    he = hipMalloc(&G.mem, (cnt * 6 * G.nc + cnt * 12 * G.slot) * sizeof(cplx));
    if (he == hipSuccess) he = hipMalloc(&G.nu, cnt * sizeof(double));
    std::vector<double> h(cnt);
    for (size_t b = 0; b < cnt; ++b) h[b] = nu[G.members[b]];
    if (he == hipSuccess) he = hipMemcpy(G.nu, h.data(), cnt * sizeof(double), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = ghx::op_zero(e->stream, G.mem, cnt * 6 * G.nc * sizeof(cplx));
    if (he != hipSuccess) rc = set_err(he == hipErrorOutOfMemory ? GHX_ERR_NOMEM : GHX_ERR_HIP, "ensemble: %s",
                                       hipGetErrorString(he));
  }
  if (!rc && (he = hipStreamSynchronize(e->stream)) != hipSuccess)
    rc = set_err(GHX_ERR_HIP, "ensemble: %s", hipGetErrorString(he));
  if (rc) { ghx_ensemble_destroy(e); return rc; }
  *out = e;
  return GHX_OK;
}

extern "C" int ghx_ensemble_fields(ghx_ensemble *e, int member, void **v) {
  ARGCHK(e && v && member >= 0 && member < (int)e->where.size(), "bad ensemble_fields arguments");
  EnsGroup &G = e->groups[e->where[member].first];
  const size_t b = e->where[member].second;
  for (int k = 0; k < 3; ++k) v[k] = G.v() + (b * 3 + k) * G.nc;
  G.clean = false;
  return GHX_OK;
}

extern "C" int ghx_ensemble_grid(ghx_ensemble *e, int member, ghx_grid **g) {
  ARGCHK(e && g && member >= 0 && member < (int)e->where.size(), "bad ensemble_grid arguments");
  *g = e->groups[e->where[member].first].g;
  return GHX_OK;
}

extern "C" int ghx_ensemble_touch(ghx_ensemble *e) {
  ARGCHK(e, "null ensemble");
  for (EnsGroup &G : e->groups) G.clean = false;
  return GHX_OK;
}

extern "C" int ghx_ensemble_set_graph(ghx_ensemble *e, int enabled) {
  ARGCHK(e, "null ensemble");
  e->use_graph = enabled != 0;
  ens_drop_graph(e);
  return GHX_OK;
}

// One substep o of every member of G (include/hd_rkstep2.f90, unforced), one
// launch per pass for the whole group.
static int ens_substep(ghx_ensemble *e, EnsGroup &G, int o, const cplx *vin, cplx *out, int prune,
                       hipStream_t st) {
  FFTCtx ctx = G.g->plan.ctx;
  ctx.stream = st;
  ctx.batch = (int)G.members.size();
  ctx.tile_z = G.tz;
  ctx.tile_y = G.ty;
  ctx.sched = G.sched;
  const size_t nc = G.nc, slot = G.slot, sstride = 12 * slot;
  Chain ch = chain_slots(G.q(), slot);
  ch.q6.bstride = ch.p6.bstride = ch.p3.bstride = ch.q3.bstride = sstride;
  FieldSet in, pq;
  for (int k = 0; k < 3; ++k) in.p[k] = const_cast<cplx *>(vin) + k * nc;
  in.bstride = 3 * nc;
  for (int k = 0; k < 3; ++k) { pq.p[k] = ch.p6.p[3 + k]; pq.p[3 + k] = ch.p6.p[k]; }
  pq.bstride = sstride;
  const double sc = (double)nc * sizeof(cplx) * ctx.batch;
  const PruneShares ps = (prune && g_prof.on) ? prune_shares(G.g) : PruneShares{1.0, 1.0, 1.0};
  PROF(P_ZINV_CURL, (3 * ps.fm + 6 * ps.fl) * sc, ctx.stream, ghx::zinv(ctx, 1, 3, in, ch.q6, prune));
  PROF(P_YINV, (6 * ps.fl + 6 * ps.fx) * sc, ctx.stream, ghx::yinv(ctx, 6, ch.q6, ch.p6, prune));
  PROF(P_XCROSS, 9 * ps.fx * sc, ctx.stream, ghx::xcross(ctx, 0, pq, ch.p3, prune));
  PROF(P_YFWD, (3 * ps.fx + 3 * ps.fl) * sc, ctx.stream, ghx::yfwd(ctx, 3, ch.p3, ch.q3, prune));
  ghx::RKParams rk;
  rk.dt = e->dt; rk.nu = 0.0; rk.rmp = 1. / (double)o;
  for (int k = 0; k < 3; ++k) {
    rk.v[k] = vin + k * nc;
    rk.c[k] = G.v() + k * nc;
    rk.out[k] = out + k * nc;
    rk.f[k] = nullptr;
  }
  rk.fline = nullptr;
  rk.bstride = 3 * nc;
  rk.nu_b = G.nu;
  const int nvc = vin == G.v() ? 3 : 6;
  PROF(P_ZFWD_RK_HD, (3 * ps.fl + nvc * ps.fm + 3 * (prune ? ps.fm : 1.0)) * sc, ctx.stream,
       ghx::zfwd_rk_hd(ctx, ch.q3, rk, prune));
  return GHX_OK;
}

static int ens_step_group(ghx_ensemble *e, EnsGroup &G, int prune, hipStream_t st = nullptr) {
  for (int o = e->ord; o >= 1; --o) {
    const bool first = (o == e->ord), last = (o == 1);
    int rc = ens_substep(e, G, o, first ? G.v() : G.v1(), last ? G.v() : G.v1(), prune, st ? st : e->stream);
    if (rc) return rc;
  }
  return GHX_OK;
}

// The captured step: the size groups are independent, so each is a branch of
// the graph (forked from and joined back to the ensemble's stream), and the
// small sizes' passes overlap the large ones' instead of queueing behind
// them.  GHX_ENS_BRANCH=0 captures them one after another.
static int ens_capture(ghx_ensemble *e) {
  const size_t ng = e->groups.size();
  const char *ev = getenv("GHX_ENS_BRANCH");
  const bool branch = ng > 1 && !(ev && atoi(ev) == 0);
  if (branch && e->gstream.size() < ng) {
    if (!e->fork) HIPCHK(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
    while (e->gstream.size() < ng) {
      hipStream_t st = nullptr;
      hipEvent_t j = nullptr;
      HIPCHK(hipStreamCreate(&st));  // default priority
      e->gstream.push_back(st);
      HIPCHK(hipEventCreateWithFlags(&j, hipEventDisableTiming));
      e->gjoin.push_back(j);
    }
  }
  HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc = GHX_OK;
  hipError_t he = hipSuccess;
  if (branch) {
    he = hipEventRecord(e->fork, e->stream);
    for (size_t k = 0; k < ng && he == hipSuccess; ++k) {
      he = hipStreamWaitEvent(e->gstream[k], e->fork, 0);
      if (he == hipSuccess && !rc) rc = ens_step_group(e, e->groups[k], 1, e->gstream[k]);
      // every forked branch is joined, also after a failure, so that the
      // capture can end
      if (he == hipSuccess) he = hipEventRecord(e->gjoin[k], e->gstream[k]);
      if (he == hipSuccess) he = hipStreamWaitEvent(e->stream, e->gjoin[k], 0);
    }
  } else {
    for (EnsGroup &G : e->groups)
      if (!rc) rc = ens_step_group(e, G, 1);
  }
  hipGraph_t graph = nullptr;
  const hipError_t hend = hipStreamEndCapture(e->stream, &graph);
  if (rc || he != hipSuccess || hend != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    HIPCHK(he != hipSuccess ? he : hend);
  }
  e->graph = graph;
  HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  return GHX_OK;
}

// The members' states are dealiased once checked; every step keeps them so.
static int ens_check_clean(ghx_ensemble *e, EnsGroup &G) {
  bool clean = true;
  for (size_t b = 0; b < G.members.size() && clean; ++b) {
    const cplx *v = G.v() + b * 3 * G.nc;
    double n = 0;
    int rc = count_unkept(G.g, v, v + G.nc, v + 2 * G.nc, &n);
    if (rc) return rc;
    clean = (n == 0.0);
  }
  G.clean = clean;
  return GHX_OK;
}

extern "C" int ghx_ensemble_step(ghx_ensemble *e, int nsteps) {
  ARGCHK(e && nsteps >= 0, "bad ensemble_step arguments");
  const bool allow_prune = !getenv("GHX_NO_PRUNE");
  for (int s = 0; s < nsteps; ++s) {
    bool all_clean = true;
    for (EnsGroup &G : e->groups) {
      if (!G.clean) {
        ens_drop_graph(e);
        int rc = ens_check_clean(e, G);
        if (rc) return rc;
      }
      all_clean = all_clean && G.clean;
    }
    if (e->use_graph && all_clean && allow_prune) {
      if (!e->exec) {
        const bool prof_was = g_prof.on;
        g_prof.on = false;  // no profiler events inside the graph
        const int rc = ens_capture(e);
        g_prof.on = prof_was;
        if (rc) return rc;
      }
      HIPCHK(hipGraphLaunch(e->exec, e->stream));
    } else {
      for (EnsGroup &G : e->groups) {
        int rc = ens_step_group(e, G, G.clean && allow_prune ? 1 : 0);
        if (rc) return rc;
      }
    }
    for (EnsGroup &G : e->groups) G.clean = true;
  }
  return GHX_OK;
}

extern "C" int ghx_ensemble_sync(ghx_ensemble *e) {
  ARGCHK(e, "null ensemble");
  HIPCHK(hipStreamSynchronize(e->stream));
  return GHX_OK;
}

// Tile widths per grid size chosen by timing: for every group, each candidate
// (z tile, y tile) runs `steps` steps of that group alone, from a saved copy
// of its state that is put back afterwards; the fastest is kept.
// The radix schedules a size can choose between: the z and y passes each on
// the default schedule or on four points per lane where that differs.
static int ens_schedules(int n) { return (n == 32 || n == 64 || n == 128 || n == 256) ? 4 : 1; }

extern "C" int ghx_ensemble_autotune(ghx_ensemble *e, int steps, int *tz_out, int *ty_out) {
  ARGCHK(e && steps > 0, "bad ensemble_autotune arguments");
  // (z, y) tile widths; the launchers clamp each to what the line length,
  // the lines there are, the block bound and the LDS allow
  static const int cand[][2] = {{4, 8},  {2, 4},  {4, 4},   {8, 8},   {8, 4},  {2, 8},
                                {4, 16}, {8, 16}, {16, 8}, {16, 16}, {32, 16}, {32, 32}};
  ens_drop_graph(e);
  hipEvent_t a = nullptr, b = nullptr;
  HIPCHK(hipEventCreate(&a));
  HIPCHK(hipEventCreate(&b));
  int rc = GHX_OK;
  for (size_t gi = 0; gi < e->groups.size() && !rc; ++gi) {
    EnsGroup &G = e->groups[gi];
    if (!G.clean && (rc = ens_check_clean(e, G))) break;
    const size_t bytes = G.members.size() * 3 * G.nc * sizeof(cplx);
    cplx *save = nullptr;
    hipError_t he = hipMalloc(&save, bytes);
    if (he == hipSuccess) he = hipMemcpyAsync(save, G.v(), bytes, hipMemcpyDeviceToDevice, e->stream);
    if (he != hipSuccess) { rc = set_err(GHX_ERR_HIP, "autotune: %s", hipGetErrorString(he)); break; }
    double best = 1e30;
    int bz = 0, by = 0, bs = 0;
    for (int sc = 0; sc < ens_schedules(G.n) && !rc; ++sc)
      for (const auto &c : cand) {
        G.tz = c[0];
        G.ty = c[1];
        G.sched = sc;
        (void)ens_step_group(e, G, G.clean ? 1 : 0);  // warm up this shape
        (void)hipEventRecord(a, e->stream);
        for (int s = 0; s < steps && !rc; ++s) rc = ens_step_group(e, G, G.clean ? 1 : 0);
        (void)hipEventRecord(b, e->stream);
        float ms = 0.f;
        if ((he = hipEventSynchronize(b)) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
          rc = set_err(GHX_ERR_HIP, "autotune: %s", hipGetErrorString(he));
          break;
        }
        if (ms < best) { best = ms; bz = c[0]; by = c[1]; bs = sc; }
        (void)hipMemcpyAsync(G.v(), save, bytes, hipMemcpyDeviceToDevice, e->stream);
      }
    G.tz = bz;
    G.ty = by;
    G.sched = bs;
    if (tz_out) tz_out[gi] = bz;
    if (ty_out) ty_out[gi] = by;
    (void)hipStreamSynchronize(e->stream);
    (void)hipFree(save);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

extern "C" int ghx_ensemble_schedules(ghx_ensemble *e, int *sched, int *tz, int *ty) {
  ARGCHK(e, "null ensemble");
  for (size_t k = 0; k < e->groups.size(); ++k) {
    if (sched) sched[k] = e->groups[k].sched;
    if (tz) tz[k] = e->groups[k].tz;
    if (ty) ty[k] = e->groups[k].ty;
  }
  return GHX_OK;
}

extern "C" int ghx_ensemble_set_schedule(ghx_ensemble *e, int group, int sched, int tz, int ty) {
  ARGCHK(e && group >= 0 && group < (int)e->groups.size() && sched >= 0 && sched <= 3 && tz >= 0 && ty >= 0,
         "bad ensemble_set_schedule arguments");
  ens_drop_graph(e);
  EnsGroup &G = e->groups[group];
  G.sched = sched < ens_schedules(G.n) ? sched : 0;
  G.tz = tz;
  G.ty = ty;
  return GHX_OK;
}

extern "C" int ghx_ensemble_groups(ghx_ensemble *e, int *ngroups, int *sizes, int *counts) {
  ARGCHK(e && ngroups, "bad ensemble_groups arguments");
  *ngroups = (int)e->groups.size();
  for (size_t k = 0; k < e->groups.size(); ++k) {
    if (sizes) sizes[k] = e->groups[k].n;
    if (counts) counts[k] = (int)e->groups[k].members.size();
  }
  return GHX_OK;
}
