// binio.hip -- GHOST's binary field files for fields resident in HBM.
//
// File format (3D/src/mpiio/binary_io.f90): the global array in Fortran
// order, raw native-endian FP64 reals (io_write/io_read) or FP64 complex pairs
// (io_writec/io_readc), no header or record markers, starting at byte offset
// disp.  A rank's slab (n1, n2, ksta:kend) sits at disp + (ksta-1)*n1*n2*esize.
// A real field r[nz][ny][nx] of this library is exactly Fortran R(nx,ny,nz),
// a spectral field c[kx][ky][kz] exactly C(nz,ny,nx/2+1), so files are
// written and read without any reordering.
//
// Device fields move through two pinned staging buffers: the PCIe copy of one
// chunk runs on the plan's stream while the host writes (or reads) the other
// chunk to (from) the file.  Host arrays are written directly.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "../../include/ghx.h"

int ghx_internal_set_err(int code, const char *msg);

namespace {

int fail(int code, const std::string &msg) { return ghx_internal_set_err(code, msg.c_str()); }

int hip_fail(hipError_t e, const char *what) {
  return fail(e == hipErrorOutOfMemory ? GHX_ERR_NOMEM : GHX_ERR_HIP,
              std::string(what) + ": " + hipGetErrorString(e));
}

#define IOHIP(expr)                                 \
  do {                                              \
    hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return hip_fail(e_, #expr); \
  } while (0)

// Device memory is staged; host, pinned-host and managed memory are accessed
// directly.  Without a usable device every pointer is a host pointer.
bool is_device_ptr(const void *p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

// rarray_byte_swap (utils/gutils_mod.fpp:114-151): byte reversal of every real.
void swap_reals(void *buf, size_t bytes) {
  uint64_t *w = static_cast<uint64_t *>(buf);
  const size_t n = bytes / 8;
  for (size_t i = 0; i < n; ++i) w[i] = __builtin_bswap64(w[i]);
}

}  // namespace

struct ghx_ioplan {
  int n[3] = {0, 0, 0};
  int ksta = 1, kend = 0;
  int esize = 8;     // 8: REAL(GP), 16: COMPLEX(GP)
  int iswap = 0;     // MODULE iovar: byte-swap reals after reading
  int bmangle = 1;   // MODULE iompi: 1 -> dir/fname.nmb.out
  long long disp = 0;  // MODULE iompi: file offset of the array
  size_t chunk = (size_t)64 << 20;
  hipStream_t stream = nullptr;
  void *pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};

  size_t slab_bytes() const { return (size_t)n[0] * n[1] * (size_t)(kend - ksta + 1) * esize; }
  long long slab_offset() const { return disp + (long long)(ksta - 1) * n[0] * n[1] * esize; }
};

namespace {

int make_plan(ghx_ioplan **plan, const int n[3], int ksta, int kend, int esize) {
  if (!plan || !n) return fail(GHX_ERR_ARG, "io_init: null argument");
  if (n[0] < 1 || n[1] < 1 || n[2] < 1 || ksta < 1 || kend < ksta || kend > n[2])
    return fail(GHX_ERR_ARG, "io_init: bad sizes or slab range");
  ghx_ioplan *p = new (std::nothrow) ghx_ioplan;
  if (!p) return fail(GHX_ERR_NOMEM, "io_init: out of host memory");
  for (int d = 0; d < 3; ++d) p->n[d] = n[d];
  p->ksta = ksta;
  p->kend = kend;
  p->esize = esize;
  *plan = p;
  return GHX_OK;
}

int staging(ghx_ioplan *p) {
  if (p->pin[0]) return GHX_OK;
  IOHIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  for (int b = 0; b < 2; ++b) {
    IOHIP(hipHostMalloc(&p->pin[b], p->chunk, hipHostMallocDefault));
    IOHIP(hipEventCreateWithFlags(&p->ev[b], hipEventDisableTiming));
  }
  return GHX_OK;
}

// io_read / io_write (mpiio/binary_io.f90:62-186) and the complex variants
// (:236-331) name the file dir/fname.nmb.out when bmangle = 1; otherwise
// dir/fname, except io_write, which opens fname alone.
std::string file_name(const ghx_ioplan *p, const char *dir, const char *fname, const char *nmb, bool real_write) {
  if (p->bmangle == 1) return std::string(dir) + "/" + fname + "." + nmb + ".out";
  if (real_write) return std::string(fname);
  return std::string(dir) + "/" + fname;
}

int full_pwrite(int fd, const char *buf, size_t len, long long off, const std::string &path) {
  while (len) {
    ssize_t w = pwrite(fd, buf, len, (off_t)off);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) return fail(GHX_ERR_ARG, "io_write: cannot write " + path + ": " + strerror(errno));
    buf += w;
    len -= (size_t)w;
    off += w;
  }
  return GHX_OK;
}

int full_pread(int fd, char *buf, size_t len, long long off, const std::string &path) {
  while (len) {
    ssize_t r = pread(fd, buf, len, (off_t)off);
    if (r < 0 && errno == EINTR) continue;
    if (r < 0) return fail(GHX_ERR_ARG, "io_read: cannot read " + path + ": " + strerror(errno));
    if (r == 0) return fail(GHX_ERR_ARG, "io_read: " + path + " is shorter than the array");
    buf += r;
    len -= (size_t)r;
    off += r;
  }
  return GHX_OK;
}

struct Fd {
  int fd = -1;
  ~Fd() {
    if (fd >= 0) close(fd);
  }
};

int do_write(ghx_ioplan *p, const std::string &path, const void *var) {
  Fd f;
  // MPI_MODE_CREATE + MPI_MODE_WRONLY: create if absent, never truncate
  f.fd = open(path.c_str(), O_WRONLY | O_CREAT, 0644);
  if (f.fd < 0) return fail(GHX_ERR_ARG, "io_write: cannot open " + path + ": " + strerror(errno));
  const size_t bytes = p->slab_bytes();
  const long long off0 = p->slab_offset();
  if (!is_device_ptr(var)) return full_pwrite(f.fd, static_cast<const char *>(var), bytes, off0, path);
  int rc = staging(p);
  if (rc) return rc;
  // fields are produced on any stream: wait for all of them once
  IOHIP(hipDeviceSynchronize());
  const char *src = static_cast<const char *>(var);
  const size_t nch = (bytes + p->chunk - 1) / p->chunk;
  auto issue = [&](size_t i) -> int {
    const size_t o = i * p->chunk, len = bytes - o < p->chunk ? bytes - o : p->chunk;
    IOHIP(hipMemcpyAsync(p->pin[i & 1], src + o, len, hipMemcpyDeviceToHost, p->stream));
    IOHIP(hipEventRecord(p->ev[i & 1], p->stream));
    return GHX_OK;
  };
  if ((rc = issue(0))) return rc;
  for (size_t i = 0; i < nch; ++i) {
    // buffer (i+1)&1 last held chunk i-1, whose file write has returned
    if (i + 1 < nch && (rc = issue(i + 1))) return rc;
    IOHIP(hipEventSynchronize(p->ev[i & 1]));
    const size_t o = i * p->chunk, len = bytes - o < p->chunk ? bytes - o : p->chunk;
    if ((rc = full_pwrite(f.fd, static_cast<const char *>(p->pin[i & 1]), len, off0 + (long long)o, path)))
      return rc;
  }
  return GHX_OK;
}

int do_read(ghx_ioplan *p, const std::string &path, void *var, bool swap) {
  Fd f;
  f.fd = open(path.c_str(), O_RDONLY);
  if (f.fd < 0) return fail(GHX_ERR_ARG, "io_read: cannot open file for reading: " + path);
  const size_t bytes = p->slab_bytes();
  const long long off0 = p->slab_offset();
  if (!is_device_ptr(var)) {
    int rc = full_pread(f.fd, static_cast<char *>(var), bytes, off0, path);
    if (!rc && swap) swap_reals(var, bytes);
    return rc;
  }
  int rc = staging(p);
  if (rc) return rc;
  IOHIP(hipDeviceSynchronize());
  char *dst = static_cast<char *>(var);
  const size_t nch = (bytes + p->chunk - 1) / p->chunk;
  for (size_t i = 0; i < nch; ++i) {
    const size_t o = i * p->chunk, len = bytes - o < p->chunk ? bytes - o : p->chunk;
    if (i >= 2) IOHIP(hipEventSynchronize(p->ev[i & 1]));  // chunk i-2 has left this buffer
    if ((rc = full_pread(f.fd, static_cast<char *>(p->pin[i & 1]), len, off0 + (long long)o, path))) {
      (void)hipStreamSynchronize(p->stream);
      return rc;
    }
    if (swap) swap_reals(p->pin[i & 1], len);
    IOHIP(hipMemcpyAsync(dst + o, p->pin[i & 1], len, hipMemcpyHostToDevice, p->stream));
    IOHIP(hipEventRecord(p->ev[i & 1], p->stream));
  }
  IOHIP(hipStreamSynchronize(p->stream));
  return GHX_OK;
}

int check_call(const ghx_ioplan *p, const char *dir, const char *fname, const char *nmb, const void *var,
               int esize) {
  if (!p || !dir || !fname || !nmb || !var) return fail(GHX_ERR_ARG, "io: null argument");
  if (p->esize != esize)
    return fail(GHX_ERR_ARG, esize == 8 ? "io_read/io_write need a plan from io_init"
                                        : "io_readc/io_writec need a plan from io_initc");
  return GHX_OK;
}

}  // namespace

extern "C" int ghx_io_init(ghx_ioplan **plan, int myrank, const int n[3], int ksta, int kend) {
  (void)myrank;
  return make_plan(plan, n, ksta, kend, 8);
}

extern "C" int ghx_io_initc(ghx_ioplan **plan, int myrank, const int n[3], int ksta, int kend) {
  (void)myrank;
  return make_plan(plan, n, ksta, kend, 16);
}

extern "C" int ghx_io_destroy(ghx_ioplan *plan) {
  if (!plan) return GHX_OK;
  if (plan->stream) (void)hipStreamSynchronize(plan->stream);
  for (int b = 0; b < 2; ++b) {
    if (plan->ev[b]) (void)hipEventDestroy(plan->ev[b]);
    if (plan->pin[b]) (void)hipHostFree(plan->pin[b]);
  }
  if (plan->stream) (void)hipStreamDestroy(plan->stream);
  delete plan;
  return GHX_OK;
}

extern "C" int ghx_io_set_options(ghx_ioplan *plan, int iswap, int bmangle, long long disp) {
  if (!plan || disp < 0) return fail(GHX_ERR_ARG, "io_set_options: bad argument");
  plan->iswap = iswap;
  plan->bmangle = bmangle;
  plan->disp = disp;
  return GHX_OK;
}

extern "C" int ghx_io_write(ghx_ioplan *plan, const char *dir, const char *fname, const char *nmb,
                            const double *var) {
  int rc = check_call(plan, dir, fname, nmb, var, 8);
  return rc ? rc : do_write(plan, file_name(plan, dir, fname, nmb, true), var);
}

extern "C" int ghx_io_read(ghx_ioplan *plan, const char *dir, const char *fname, const char *nmb, double *var) {
  int rc = check_call(plan, dir, fname, nmb, var, 8);
  return rc ? rc : do_read(plan, file_name(plan, dir, fname, nmb, false), var, plan->iswap > 0);
}

extern "C" int ghx_io_writec(ghx_ioplan *plan, const char *dir, const char *fname, const char *nmb,
                             const void *var) {
  int rc = check_call(plan, dir, fname, nmb, var, 16);
  return rc ? rc : do_write(plan, file_name(plan, dir, fname, nmb, false), var);
}

extern "C" int ghx_io_readc(ghx_ioplan *plan, const char *dir, const char *fname, const char *nmb, void *var) {
  int rc = check_call(plan, dir, fname, nmb, var, 16);
  // io_readc does not byte-swap (mpiio/binary_io.f90:236-282)
  return rc ? rc : do_read(plan, file_name(plan, dir, fname, nmb, false), var, false);
}
