// pseudo.hip -- pointwise spectral operators, RK update, diagnostics and
// reductions on device-resident GHOST fields (gfx950, FP64).
//
// Spectral index m = (i*ny + j)*nz + k for kx index i, ky index j, kz index k
// (GHOST's C(nz,ny,nx/2+1) in column-major order).
#include "pseudo.h"

namespace ghx {

namespace {

constexpr int kBlock = 256;

struct Idx {
  int i, j, k;
};
GHX_DEV Idx decompose(size_t m, const GridK &g) {
  Idx r;
  r.k = (int)(m % (size_t)g.nz);
  size_t line = m / (size_t)g.nz;
  r.j = (int)(line % (size_t)g.ny);
  r.i = (int)(line / (size_t)g.ny);
  return r;
}

__host__ __device__ inline size_t ncplx(const GridK &g) { return (size_t)g.nxh * g.ny * g.nz; }
__host__ __device__ inline size_t nreal(const GridK &g) { return (size_t)g.nx * g.ny * g.nz; }
inline unsigned nblocks(size_t n) {
  size_t b = (n + kBlock - 1) / kBlock;
  if (b > 65536u * 4u) b = 65536u * 4u;
  return (unsigned)(b ? b : 1);
}

#define GRID_STRIDE(m, n) for (size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x; m < (n); m += (size_t)gridDim.x * blockDim.x)

// derivk3 (pseudospec3D_hd.f90:22-88): b = i k_dir a
__global__ void k_derivk3(GridK g, const cplx *a, cplx *b, int dir) {
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    double kk = dir == 1 ? wavenum(q.i, g.nx, g.dkx) : dir == 2 ? wavenum(q.j, g.ny, g.dky) : wavenum(q.k, g.nz, g.dkz);
    b[m] = c_ik(kk, a[m]);
  }
}

GHX_DEV double kk2_at(const GridK &g, const Idx &q) {
  if (!g.anis) return kn2_of(g, q.i, q.j, q.k);
  return kk2_of(g, wavenum(q.i, g.nx, g.dkx), wavenum(q.j, g.ny, g.dky), wavenum(q.k, g.nz, g.dkz));
}

// laplak3 (:91-121): b = -kk2 a
__global__ void k_laplak3(GridK g, const cplx *a, cplx *b) {
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    double kk2 = kk2_at(g, q);
    cplx v = a[m];
    b[m] = c_make(-kk2 * v.x, -kk2 * v.y);
  }
}

// rotor3 (:124-200)
__global__ void k_rotor3(GridK g, const cplx *a, const cplx *b, cplx *c, int dir) {
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    double kx = wavenum(q.i, g.nx, g.dkx), ky = wavenum(q.j, g.ny, g.dky), kz = wavenum(q.k, g.nz, g.dkz);
    cplx av = a[m], bv = b[m];
    if (dir == 1) c[m] = c_sub(c_ik(ky, bv), c_ik(kz, av));
    else if (dir == 2) c[m] = c_sub(c_ik(kz, av), c_ik(kx, bv));
    else c[m] = c_sub(c_ik(kx, bv), c_ik(ky, av));
  }
}

// nonlhd3 (:399-509), nonlin3 (mhd :108-227), gauge3 (mhd :230-336)
// mode 0: g = -a_dir + k_dir (k.a)/k^2;  mode 1: g = -a_dir + d_dir + k_dir (k.(a-d))/k^2;
// mode 2: g = a_dir - k_dir (k.a)/k^2.  The k_dir = 0 plane keeps only the first terms.
__global__ void k_project(GridK g, int mode, const cplx *a, const cplx *b, const cplx *c, const cplx *d,
                          const cplx *e, const cplx *f, cplx *o, int dir) {
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    double kx = wavenum(q.i, g.nx, g.dkx), ky = wavenum(q.j, g.ny, g.dky), kz = wavenum(q.k, g.nz, g.dkz);
    cplx A = a[m], B = b[m], Cc = c[m];
    cplx s = dir == 1 ? A : dir == 2 ? B : Cc;
    int plane = dir == 1 ? q.i : dir == 2 ? q.j : q.k;
    double kd = dir == 1 ? kx : dir == 2 ? ky : kz;
    cplx r;
    if (mode == 1) {
      cplx D = d[m], Ee = e[m], F = f[m];
      cplx sd = dir == 1 ? D : dir == 2 ? Ee : F;
      r = c_add(c_make(-s.x, -s.y), sd);
      if (plane != 0) {
        double kk2 = kk2_at(g, q);
        cplx dot = c_add(c_add(c_scale(c_sub(A, D), kx), c_scale(c_sub(B, Ee), ky)), c_scale(c_sub(Cc, F), kz));
        r = c_add(r, c_make(kd * dot.x / kk2, kd * dot.y / kk2));
      }
    } else {
      r = mode == 0 ? c_make(-s.x, -s.y) : s;
      if (plane != 0) {
        double kk2 = kk2_at(g, q);
        cplx dot = c_add(c_add(c_scale(A, kx), c_scale(B, ky)), c_scale(Cc, kz));
        cplx t = c_make(kd * dot.x / kk2, kd * dot.y / kk2);
        r = mode == 0 ? c_add(r, t) : c_sub(r, t);
      }
    }
    o[m] = r;
  }
}

// phaseshift (:2084-2150).  On the kx=0 plane GHOST's loops multiply some
// entries by cdump and some by its conjugate; the kz and ky Nyquist entries
// get both, in the loop order reproduced here.
__global__ void k_phaseshift(GridK g, cplx *a, double cre, double cim) {
  const cplx cd = c_make(cre, cim), jd = c_make(cre, -cim);
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    cplx v = a[m];
    if (q.i > 0) {
      v = c_mul(v, cd);
    } else {
      const int ny = g.ny, nz = g.nz, j = q.j, k = q.k;
      if (k == 0) {
        if (j >= 1 && j <= ny / 2) v = c_mul(v, cd);
        if (j >= ny / 2 && j <= ny - 1 && j >= 1) v = c_mul(v, jd);
      } else if (j == 0) {
        if (k <= nz / 2) v = c_mul(v, cd);
        if (k >= nz / 2) v = c_mul(v, jd);
      } else {
        // loop over j (outer, 1..ny-1) then k (inner, 1..nz/2): entry (k,j) with
        // k <= nz/2 gets cd at iteration (j,k); entry (k,j) with k >= nz/2 gets
        // jd at iteration (ny-j, nz-k).
        const bool gets_cd = k <= nz / 2;
        const bool gets_jd = k >= nz / 2;
        if (gets_cd && gets_jd) {
          if (j <= ny - j) { v = c_mul(v, cd); v = c_mul(v, jd); }
          else { v = c_mul(v, jd); v = c_mul(v, cd); }
        } else if (gets_cd) {
          v = c_mul(v, cd);
        } else {
          v = c_mul(v, jd);
        }
      }
    }
    a[m] = v;
  }
}

__global__ void k_scale(size_t n, cplx *a, const cplx *src, double s) {
  GRID_STRIDE(m, n) {
    cplx v = src[m];
    a[m] = c_make(v.x * s, v.y * s);
  }
}

// module_dns.f90:33-38: f = f*f0/sqrt(tmp)
__global__ void k_normalize(size_t n, cplx *x, cplx *y, cplx *z, double f0, double sq) {
  GRID_STRIDE(m, n) {
    cplx a = x[m], b = y[m], c = z[m];
    x[m] = c_make(a.x * f0 / sq, a.y * f0 / sq);
    y[m] = c_make(b.x * f0 / sq, b.y * f0 / sq);
    z[m] = c_make(c.x * f0 / sq, c.y * f0 / sq);
  }
}

// sharp (pseudospec3D_filt.f90:200-241): keep kn2 <= kc2
__global__ void k_sharp(GridK g, const cplx *a, cplx *b, double kc2) {
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    b[m] = kn2_of(g, q.i, q.j, q.k) <= kc2 ? a[m] : c_make(0.0, 0.0);
  }
}

// Subgrid stress on the fine grid (learned LES, no counterpart in GHOST):
// p_ij = (u_i u_j - ub_i ub_j) s^2 from the unnormalised c2r outputs of the
// velocity u and of its filtered part ub (s = 1/(nx ny nz)); components in
// the order xx, yy, zz, xy, xz, yz.
struct R6 {
  double *p[6];
};
struct CR3 {
  const double *p[3];
};
__global__ void k_sgs_products(size_t n, CR3 u, CR3 ub, double s2, R6 out) {
  constexpr int I[6] = {0, 1, 2, 0, 0, 1}, J[6] = {0, 1, 2, 1, 2, 2};
  GRID_STRIDE(m, n) {
    const double a[3] = {u.p[0][m], u.p[1][m], u.p[2][m]};
    const double b[3] = {ub.p[0][m], ub.p[1][m], ub.p[2][m]};
#pragma unroll
    for (int c = 0; c < 6; ++c) out.p[c][m] = (a[I[c]] * a[J[c]] - b[I[c]] * b[J[c]]) * s2;
  }
}

// Subgrid force of the LES: out = f0 + P(-div tau) - nu(k) kk2 v on the kept
// modes (tiny <= kn2 <= kmax), f0 elsewhere (zero without f0).  tau: six
// spectral fields xx, yy, zz, xy, xz, yz (null: no stress term); P removes
// the component along k, as the projections of nonlhd3 do (the pressure of
// the subgrid force).  nuk (may be null): a viscosity per shell
// int(sqrt(kk2)/Dkk + .501) = 1..nb acting on the velocity v.
struct C6 {
  const cplx *p[6];
};
struct SpecVisc {
  const cplx *v[3];
  const double *nuk;
  int nb;
  double dkk;
};
__global__ void k_sgs_force(GridK g, C6 t, SpecVisc sv, const cplx *f0x, const cplx *f0y, const cplx *f0z, cplx *ox,
                            cplx *oy, cplx *oz) {
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    const cplx z = c_make(0.0, 0.0);
    cplx fx = f0x ? f0x[m] : z, fy = f0y ? f0y[m] : z, fz = f0z ? f0z[m] : z;
    const double kn2 = kn2_of(g, q.i, q.j, q.k);
    if (kn2 <= g.kmax && kn2 >= g.tiny) {
      const double kx = wavenum(q.i, g.nx, g.dkx), ky = wavenum(q.j, g.ny, g.dky), kz = wavenum(q.k, g.nz, g.dkz);
      const double kk2 = g.anis ? kk2_of(g, kx, ky, kz) : kn2;
      if (t.p[0]) {
        const cplx txx = t.p[0][m], tyy = t.p[1][m], tzz = t.p[2][m], txy = t.p[3][m], txz = t.p[4][m],
                   tyz = t.p[5][m];
        // -i k_j tau_ij
        const cplx sx = c_make(-(-kx * txx.y - ky * txy.y - kz * txz.y), -(kx * txx.x + ky * txy.x + kz * txz.x));
        const cplx sy = c_make(-(-kx * txy.y - ky * tyy.y - kz * tyz.y), -(kx * txy.x + ky * tyy.x + kz * tyz.x));
        const cplx sz = c_make(-(-kx * txz.y - ky * tyz.y - kz * tzz.y), -(kx * txz.x + ky * tyz.x + kz * tzz.x));
        const cplx dot = c_add(c_add(c_scale(sx, kx), c_scale(sy, ky)), c_scale(sz, kz));
        fx = c_add(fx, c_sub(sx, c_scale(dot, kx / kk2)));
        fy = c_add(fy, c_sub(sy, c_scale(dot, ky / kk2)));
        fz = c_add(fz, c_sub(sz, c_scale(dot, kz / kk2)));
      }
      if (sv.nuk) {
        const int kmn = (int)(sqrt(kk2) / sv.dkk + 0.501);
        if (kmn >= 1 && kmn <= sv.nb) {
          const double a = -sv.nuk[kmn - 1] * kk2;
          fx = c_add(fx, c_scale(sv.v[0][m], a));
          fy = c_add(fy, c_scale(sv.v[1][m], a));
          fz = c_add(fz, c_scale(sv.v[2][m], a));
        }
      }
    }
    ox[m] = fx;
    oy[m] = fy;
    oz[m] = fz;
  }
}

// analysis/trunc3D.fpp:190-320: out (grid t, smaller) = the modes of c
// (grid g) with kx^2/nxt^2 + ky^2/nyt^2 + kz^2/nzt^2 <= 1/9, scaled by
// 1/(nx ny nz), ky and kz re-indexed onto t's axes.
__global__ void k_truncate(GridK g, GridK t, const cplx *c, cplx *out) {
  const double rmp = 1.0 / ((double)t.nx * (double)t.nx), rmq = 1.0 / ((double)t.ny * (double)t.ny),
               rms = 1.0 / ((double)t.nz * (double)t.nz);
  const double fact = 1.0 / ((double)g.nx * (double)g.ny * (double)g.nz);
  GRID_STRIDE(m, ncplx(t)) {
    Idx q = decompose(m, t);
    const int j = q.j < t.ny / 2 ? q.j : q.j - t.ny + g.ny;
    const int k = q.k < t.nz / 2 ? q.k : q.k - t.nz + g.nz;
    const double kx = (double)(q.i < g.nx / 2 ? q.i : q.i - g.nx), ky = (double)(j < g.ny / 2 ? j : j - g.ny),
                 kz = (double)(k < g.nz / 2 ? k : k - g.nz);
    const double kn2 = __dadd_rn(__dadd_rn(__dmul_rn(rmp, __dmul_rn(kx, kx)), __dmul_rn(rmq, __dmul_rn(ky, ky))),
                                 __dmul_rn(rms, __dmul_rn(kz, kz)));
    const cplx v = c[((size_t)q.i * g.ny + j) * g.nz + k];
    out[m] = kn2 > 1.0 / 9.0 ? c_make(0.0, 0.0) : c_make(v.x * fact, v.y * fact);
  }
}

__global__ void k_zero(size_t n, double *p) {
  GRID_STRIDE(m, n) p[m] = 0.0;
}

__global__ void k_cross_real(size_t n, double tmp, double *r1, double *r2, double *r3, const double *r4,
                             const double *r5, const double *r6, double *r7) {
  GRID_STRIDE(m, n) {
    double a = (r2[m] * r6[m] - r5[m] * r3[m]) * tmp;
    double b = (r3[m] * r4[m] - r6[m] * r1[m]) * tmp;
    double c = (r1[m] * r5[m] - r4[m] * r2[m]) * tmp;
    r7[m] = a;
    r3[m] = b;
    r1[m] = c;
  }
}

__global__ void k_rk_update(GridK g, cplx *v, const cplx *c, const cplx *lap, const cplx *nl, const cplx *f,
                            double nu, double dt, double rmp) {
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    double kn2 = kn2_of(g, q.i, q.j, q.k);
    cplx out = c_make(0.0, 0.0);
    if (kn2 <= g.kmax && kn2 >= g.tiny) {
      cplx L = lap[m], N = nl[m], C = c[m];
      cplx F = f ? f[m] : c_make(0.0, 0.0);
      double sx = nu * L.x + N.x + F.x, sy = nu * L.y + N.y + F.y;
      out = c_make(C.x + dt * sx * rmp, C.y + dt * sy * rmp);
    }
    v[m] = out;
  }
}

// The pointwise steps of the ord = 3 correction, include/bouss_rkstep2.f90:
// 73-145.  n = nonlhd3 of (w4, w5, w6 + xmom th) in the three directions (the
// gravity loops :78-85 and :115-122 followed by nonlhd3 :87-89, :124-126).
// mode 0 (:90-111): out_d = -kk2 c_d + n_d on kept modes, else 0 (laplak3
// then the masked sum); mode 1: out_d = n_d.
__global__ void k_bouss3_nonl(GridK g, int mode, const cplx *w4, const cplx *w5, const cplx *w6, const cplx *th,
                              double xmom, const cplx *c1, const cplx *c2, const cplx *c3, cplx *o1, cplx *o2,
                              cplx *o3) {
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    double kx = wavenum(q.i, g.nx, g.dkx), ky = wavenum(q.j, g.ny, g.dky), kz = wavenum(q.k, g.nz, g.dkz);
    const cplx t = th[m], c6 = w6[m];
    const cplx a = w4[m], b = w5[m], c = c_make(c6.x + xmom * t.x, c6.y + xmom * t.y);
    cplx n1 = c_make(-a.x, -a.y), n2 = c_make(-b.x, -b.y), n3 = c_make(-c.x, -c.y);
    if (q.i != 0 || q.j != 0 || q.k != 0) {
      const double kk2 = kk2_at(g, q);
      const cplx dot = c_add(c_add(c_scale(a, kx), c_scale(b, ky)), c_scale(c, kz));
      if (q.i != 0) n1 = c_add(n1, c_make(kx * dot.x / kk2, kx * dot.y / kk2));
      if (q.j != 0) n2 = c_add(n2, c_make(ky * dot.x / kk2, ky * dot.y / kk2));
      if (q.k != 0) n3 = c_add(n3, c_make(kz * dot.x / kk2, kz * dot.y / kk2));
    }
    if (mode == 0) {
      const double kn2 = kn2_of(g, q.i, q.j, q.k);
      if (kn2 <= g.kmax && kn2 >= g.tiny) {
        const double kk2 = kk2_at(g, q);
        const cplx u1 = c1[m], u2 = c2[m], u3 = c3[m];
        n1 = c_add(c_make(-kk2 * u1.x, -kk2 * u1.y), n1);
        n2 = c_add(c_make(-kk2 * u2.x, -kk2 * u2.y), n2);
        n3 = c_add(c_make(-kk2 * u3.x, -kk2 * u3.y), n3);
      } else {
        n1 = n2 = n3 = c_make(0.0, 0.0);
      }
    }
    o1[m] = n1;
    o2[m] = n2;
    o3[m] = n3;
  }
}

// include/bouss_rkstep2.f90:130-143: v_d = v_d + 2 x_d rmp on every mode
__global__ void k_bouss3_add(size_t n, cplx *v1, cplx *v2, cplx *v3, const cplx *x1, const cplx *x2,
                             const cplx *x3, double rmp) {
  GRID_STRIDE(m, n) {
    cplx a = x1[m], b = x2[m], c = x3[m], p = v1[m], r = v2[m], s = v3[m];
    v1[m] = c_make(p.x + 2.0 * a.x * rmp, p.y + 2.0 * a.y * rmp);
    v2[m] = c_make(r.x + 2.0 * b.x * rmp, r.y + 2.0 * b.y * rmp);
    v3[m] = c_make(s.x + 2.0 * c.x * rmp, s.y + 2.0 * c.y * rmp);
  }
}

// examples/initialv.f90_tg:37-56 (real space, x fastest)
__global__ void k_tg_real(GridK g, double *r1, double *r2, int kdn, int kup) {
  const double pi = 3.14159265358979323846;
  GRID_STRIDE(m, nreal(g)) {
    int i = (int)(m % (size_t)g.nx) + 1;
    size_t rest = m / (size_t)g.nx;
    int j = (int)(rest % (size_t)g.ny) + 1;
    int k = (int)(rest / (size_t)g.ny) + 1;
    double a = 0.0, b = 0.0;
    for (int ki = kdn; ki <= kup; ++ki) {
      double ax = 2 * pi * ki * ((double)i - 1) / (double)g.nx;
      double ay = 2 * pi * ki * ((double)j - 1) / (double)g.ny;
      double az = 2 * pi * ki * ((double)k - 1) / (double)g.nz;
      a = a + sin(ax) * cos(ay) * cos(az);
      b = b - cos(ax) * sin(ay) * cos(az);
    }
    r1[m] = a;
    r2[m] = b;
  }
}

// ---------------------------------------------------------------------------
// Reductions: per-block partial sums in a fixed order, then one block sums
// the partials in index order, so results are reproducible run to run.
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 2048;

GHX_DEV double block_sum(double v) {
  __shared__ double sh[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x / 64); ++q) r += sh[q];
  __syncthreads();
  return r;
}

GHX_DEV double cabs2(cplx a) { return a.x * a.x + a.y * a.y; }
GHX_DEV double cre_dot(cplx a, cplx b) { return a.x * b.x + a.y * b.y; }  // Re(a conj(b))

__global__ void k_reduce(GridK g, int kind, const cplx *a, const cplx *b, const cplx *c, const cplx *d,
                         const cplx *e, const cplx *f, double *part) {
  double acc = 0.0;
  const double tmp = g.inv_n6;
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    const double w = q.i == 0 ? 1.0 : 2.0;
    double s;
    if (kind == 0) {
      s = cabs2(a[m]) + cabs2(b[m]) + cabs2(c[m]);
    } else if (kind == 1 || kind == 2) {
      s = cre_dot(a[m], d[m]) + cre_dot(b[m], e[m]) + cre_dot(c[m], f[m]);
      if (kind == 2) s *= kk2_at(g, q);
    } else if (kind == 3) {
      double kx = wavenum(q.i, g.nx, g.dkx), ky = wavenum(q.j, g.ny, g.dky), kz = wavenum(q.k, g.nz, g.dkz);
      cplx dv = c_add(c_add(c_ik(kx, a[m]), c_ik(ky, b[m])), c_ik(kz, c[m]));
      s = cabs2(dv);
    } else if (kind >= 6) {
      // scalars (pseudospec3D_phd.f90): 6 variance |a|^2, 7 kk2 |a|^2, 8 product Re(a d*)
      s = kind == 8 ? cre_dot(a[m], d[m]) : cabs2(a[m]);
      if (kind == 7) s *= kk2_at(g, q);
    } else {
      // kind 4: count of modes outside tiny <= kn2 <= kmax holding a non-zero
      // value; kind 5: the same outside kn2 <= kmax only (a scalar whose mean,
      // kn2 < tiny, is carried)
      const double kn2 = kn2_of(g, q.i, q.j, q.k);
      const cplx A = a[m], B = b[m], Cc = c[m];
      const bool nz = A.x != 0.0 || A.y != 0.0 || B.x != 0.0 || B.y != 0.0 || Cc.x != 0.0 || Cc.y != 0.0;
      const bool kept = kind == 4 ? (kn2 <= g.kmax && kn2 >= g.tiny) : kn2 <= g.kmax;
      acc += (!kept && nz) ? 1.0 : 0.0;
      continue;
    }
    acc += w * s * tmp;
  }
  double r = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

__global__ void k_sum_partials(const double *part, int n, double *out) {
  double acc = 0.0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) acc += part[q];
  double r = block_sum(acc);
  if (threadIdx.x == 0) *out = r;
}

// spectrumc, entrans, heltrans, crosspecc: shell kmn = int(sqrt(kk2)/Dkk +
// off), bins 1..nb; off is .501 as a single-precision literal, except in
// crosspecc (shift + .501 in working precision).  kind 0: |a|^2+|b|^2+|c|^2;
// 1: Re(a d*) + Re(b e*) + Re(c f*); 2: kind 1 times kk2 (entrans, kin = 0).
__global__ void k_spectrum(GridK g, int kind, double dkk, int nb, double off, const cplx *a, const cplx *b,
                           const cplx *c, const cplx *d, const cplx *e, const cplx *f, double *part) {
  extern __shared__ double bins[];
  for (int q = threadIdx.x; q < nb; q += blockDim.x) bins[q] = 0.0;
  __syncthreads();
  const double tmp = g.inv_n6;
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    double kk2 = kk2_at(g, q);
    int kmn = (int)(sqrt(kk2) / dkk + off);
    if (kmn > 0 && kmn <= nb) {
      const double w = q.i == 0 ? 1.0 : 2.0;
      double s;
      if (kind == 0) s = cabs2(a[m]) + cabs2(b[m]) + cabs2(c[m]);
      else if (kind == 3) s = cabs2(a[m]);             // one scalar (spectrscc)
      else if (kind == 4) s = cre_dot(a[m], d[m]);     // scalar transfer (sctrans)
      else s = cre_dot(a[m], d[m]) + cre_dot(b[m], e[m]) + cre_dot(c[m], f[m]);
      if (kind == 2) s = kk2 * s;
      atomicAdd(&bins[kmn - 1], w * s * tmp);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nb; q += blockDim.x) part[(size_t)blockIdx.x * nb + q] = bins[q];
}

__global__ void k_spectrum_sum(const double *part, int nblk, int nb, double *out) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nb; q += gridDim.x * blockDim.x) {
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += part[(size_t)b * nb + q];
    out[q] = acc;
  }
}

// spectr1d (pseudospec3D_hd.f90:1418-1626): planes of k_dir, bin
// int(|k_dir| L_dir + 1) in 1..n_dir/2+1 (the box-scaled wavenumber times the
// box length, GHOST's arithmetic), |a|^2 / N^6 with the Hermitian weight.
__global__ void k_spectr1d(GridK g, int dir, double ldir, int nb, const cplx *a, double *part) {
  extern __shared__ double bins[];
  for (int q = threadIdx.x; q < nb; q += blockDim.x) bins[q] = 0.0;
  __syncthreads();
  const double tmp = g.inv_n6;
  const int lim = dir == 1 ? g.nx / 2 + 1 : dir == 2 ? g.ny / 2 + 1 : g.nz / 2 + 1;
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    const double kd = dir == 1 ? fabs(wavenum(q.i, g.nx, g.dkx)) * ldir
                      : dir == 2 ? fabs(wavenum(q.j, g.ny, g.dky)) * ldir
                                 : fabs(wavenum(q.k, g.nz, g.dkz)) * ldir;
    const int kmn = (int)(kd + 1);
    if (kmn <= lim && kmn <= nb) atomicAdd(&bins[kmn - 1], (q.i == 0 ? 1.0 : 2.0) * cabs2(a[m]) * tmp);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < nb; q += blockDim.x) part[(size_t)blockIdx.x * nb + q] = bins[q];
}

// Axisymmetric spectra of pseudospec3D_rot.f90 about z.  bins 0: planes
// kmn = int(|kz| Lz + 1) (specpara, entpara, heltpara); 1: cylinders
// kmn = int(sqrt(kx^2 + ky^2)/Dkk + 1) (specperp, entperp, heltperp).
// kind 0: |a|^2 + |b|^2 (xy) and |c|^2 (z); 1: Re(a d*) + Re(b e*) and
// Re(c f*); 2: kind 1 times kk2.  Per block, three rows of nb bins: the total,
// and the xy and z parts (every mode for planes, the kz = 0 plane only for
// cylinders).
__global__ void k_axispec(GridK g, int bins, int kind, int nb, double dkk, double lz, const cplx *a, const cplx *b,
                          const cplx *c, const cplx *d, const cplx *e, const cplx *f, double *part) {
  extern __shared__ double bins3[];
  for (int q = threadIdx.x; q < 3 * nb; q += blockDim.x) bins3[q] = 0.0;
  __syncthreads();
  const double tmp = g.inv_n6;
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    int kmn;
    if (bins == 0) {
      kmn = (int)(fabs(wavenum(q.k, g.nz, g.dkz)) * lz + 1);
    } else {
      const double kx = wavenum(q.i, g.nx, g.dkx), ky = wavenum(q.j, g.ny, g.dky);
      kmn = (int)(sqrt(kx * kx + ky * ky) / dkk + 1);
    }
    if (kmn <= 0 || kmn > nb) continue;
    const double w = (q.i == 0 ? 1.0 : 2.0) * tmp;
    double h, v;
    if (kind == 0) {
      h = cabs2(a[m]) + cabs2(b[m]);
      v = cabs2(c[m]);
    } else if (kind == 3) {  // one field (spectpv)
      h = cabs2(a[m]);
      v = 0.0;
    } else {
      h = cre_dot(a[m], d[m]) + cre_dot(b[m], e[m]);
      v = cre_dot(c[m], f[m]);
      if (kind == 2) {
        const double k2 = kk2_at(g, q);
        h *= k2;
        v *= k2;
      }
    }
    atomicAdd(&bins3[kmn - 1], w * (h + v));
    if (bins == 0 || q.k == 0) {
      atomicAdd(&bins3[nb + kmn - 1], w * h);
      atomicAdd(&bins3[2 * nb + kmn - 1], w * v);
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * nb; q += blockDim.x) part[(size_t)blockIdx.x * 3 * nb + q] = bins3[q];
}

// spec2D (pseudospec3D_rot.f90:1331-1629): E(kperp, kpara), kperp fastest,
// np by nzb bins, accumulated with device-wide atomics into out (zeroed by
// the caller).  kind 0: |a|^2 + |b|^2 + |c|^2; 1: Re(a d* + b e* + c f*).
__global__ void k_spec2d(GridK g, int kind, int np, int nzb, double dkk, double lz, const cplx *a, const cplx *b,
                         const cplx *c, const cplx *d, const cplx *e, const cplx *f, double *out) {
  const double tmp = g.inv_n6;
  GRID_STRIDE(m, ncplx(g)) {
    Idx q = decompose(m, g);
    const double kx = wavenum(q.i, g.nx, g.dkx), ky = wavenum(q.j, g.ny, g.dky);
    const int kmn = (int)(sqrt(kx * kx + ky * ky) / dkk + 1);
    const int kmz = (int)(fabs(wavenum(q.k, g.nz, g.dkz)) * lz + 1);
    if (kmn <= 0 || kmn > np || kmz <= 0 || kmz > nzb) continue;
    const double s = kind == 0 ? cabs2(a[m]) + cabs2(b[m]) + cabs2(c[m])
                               : cre_dot(a[m], d[m]) + cre_dot(b[m], e[m]) + cre_dot(c[m], f[m]);
    atomicAdd(&out[(size_t)(kmz - 1) * np + (kmn - 1)], (q.i == 0 ? 1.0 : 2.0) * s * tmp);
  }
}

// difucx / difucz (pseudospec3D_phd.f90:729-940): for each x (dir 1) or z
// (dir 3) plane s, the sums of r1, r2 and r1 r2 over the plane, scaled by
// tmp, tmp and tmq; out = {<v>(s), <th>(s), <v th>(s)}, n entries each.
__global__ void k_plane_means(const double *r1, const double *r2, int nx, int ny, int nz, int dir, double tmp,
                              double tmq, double *out) {
  const int s = blockIdx.x;
  const int n = dir == 1 ? nx : nz;
  double a1 = 0.0, a2 = 0.0, a3 = 0.0;
  if (dir == 1) {
    for (size_t t = threadIdx.x; t < (size_t)ny * nz; t += blockDim.x) {
      const size_t m = t * nx + s;  // (k, j) = (t / ny, t % ny)
      a1 += r1[m];
      a2 += r2[m];
      a3 += r1[m] * r2[m];
    }
  } else {
    const size_t base = (size_t)s * nx * ny;
    for (size_t t = threadIdx.x; t < (size_t)nx * ny; t += blockDim.x) {
      a1 += r1[base + t];
      a2 += r2[base + t];
      a3 += r1[base + t] * r2[base + t];
    }
  }
  a1 = block_sum(a1);
  a2 = block_sum(a2);
  a3 = block_sum(a3);
  if (threadIdx.x == 0) {
    out[s] = a1 * tmp;
    out[n + s] = a2 * tmp;
    out[2 * n + s] = a3 * tmq;
  }
}

// havgcomp (pseudospec3D_bouss.f90:19-555): per z plane (one block each),
// the sum over x, y of kind 0: r1^2 + r2^2, 1: r1^2, 2: r1 r2, 3: -r1 r2 +
// r3 r4, 4: r3 (r1 - r2), 5: PV^2 with PV = fo r1 t - bv r3 t - r4 t - fo bv
// accumulated in GHOST's order (r1 = dth/dz, r3 = omega_z, r4 = the advect3
// term, t = 1/(nx ny nz)); out[k] = sum times scale.
__global__ void k_zplane(int kind, size_t plane, const double *r1, const double *r2, const double *r3,
                         const double *r4, double fo, double bv, double t, double scale, double *out) {
  const size_t base = (size_t)blockIdx.x * plane;
  double acc = 0.0;
  for (size_t q = threadIdx.x; q < plane; q += blockDim.x) {
    const size_t m = base + q;
    switch (kind) {
      case 0: acc += r1[m] * r1[m] + r2[m] * r2[m]; break;
      case 1: acc += r1[m] * r1[m]; break;
      case 2: acc += r1[m] * r2[m]; break;
      case 3: acc += -(r1[m] * r2[m]) + r3[m] * r4[m]; break;
      case 4: acc += r3[m] * (r1[m] - r2[m]); break;
      default: {
        double pv = 0.0 + fo * r1[m] * t;
        pv = pv - bv * r3[m] * t;
        pv = pv + r4[m] * (-t);
        pv = pv - fo * bv;
        acc += pv * pv;
      }
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = acc * scale;
}

__global__ void k_maxnorm(size_t n, const double *r1, const double *r2, const double *r3, double *part) {
  double v = 0.0;
  GRID_STRIDE(m, n) {
    double s = sqrt(r1[m] * r1[m] + r2[m] * r2[m] + r3[m] * r3[m]);
    v = fmax(v, s);
  }
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_max_partials(const double *part, int n, double *out) {
  __shared__ double sh[kBlock];
  double v = 0.0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) v = fmax(v, part[q]);
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

}  // namespace

#define LAUNCH(kern, n, st, ...)                                                     \
  do {                                                                               \
    hipLaunchKernelGGL(kern, dim3(nblocks(n)), dim3(kBlock), 0, st, __VA_ARGS__);    \
    return hipGetLastError();                                                        \
  } while (0)

hipError_t op_derivk3(const GridK &g, hipStream_t st, const cplx *a, cplx *b, int dir) {
  LAUNCH(k_derivk3, ncplx(g), st, g, a, b, dir);
}
hipError_t op_laplak3(const GridK &g, hipStream_t st, const cplx *a, cplx *b) { LAUNCH(k_laplak3, ncplx(g), st, g, a, b); }
hipError_t op_rotor3(const GridK &g, hipStream_t st, const cplx *a, const cplx *b, cplx *c, int dir) {
  LAUNCH(k_rotor3, ncplx(g), st, g, a, b, c, dir);
}
hipError_t op_nonlhd3(const GridK &g, hipStream_t st, const cplx *a, const cplx *b, const cplx *c, cplx *o, int dir) {
  LAUNCH(k_project, ncplx(g), st, g, 0, a, b, c, (const cplx *)nullptr, (const cplx *)nullptr, (const cplx *)nullptr, o, dir);
}
hipError_t op_nonlin3(const GridK &g, hipStream_t st, const cplx *a, const cplx *b, const cplx *c, const cplx *d,
                      const cplx *e, const cplx *f, cplx *o, int dir) {
  LAUNCH(k_project, ncplx(g), st, g, 1, a, b, c, d, e, f, o, dir);
}
hipError_t op_gauge3(const GridK &g, hipStream_t st, const cplx *a, const cplx *b, const cplx *c, cplx *o, int dir) {
  LAUNCH(k_project, ncplx(g), st, g, 2, a, b, c, (const cplx *)nullptr, (const cplx *)nullptr, (const cplx *)nullptr, o, dir);
}
hipError_t op_phaseshift(const GridK &g, hipStream_t st, cplx *a, double cre, double cim) {
  LAUNCH(k_phaseshift, ncplx(g), st, g, a, cre, cim);
}
hipError_t op_scale(const GridK &g, hipStream_t st, cplx *a, const cplx *src, double s) {
  LAUNCH(k_scale, ncplx(g), st, ncplx(g), a, src, s);
}
hipError_t op_normalize_apply(const GridK &g, hipStream_t st, cplx *x, cplx *y, cplx *z, double f0, double e) {
  LAUNCH(k_normalize, ncplx(g), st, ncplx(g), x, y, z, f0, sqrt(e));
}
hipError_t op_truncate(const GridK &g, const GridK &t, hipStream_t st, const cplx *c, cplx *out) {
  LAUNCH(k_truncate, ncplx(t), st, g, t, c, out);
}

hipError_t op_sgs_products(const GridK &g, hipStream_t st, const double *const *u, const double *const *ub, double s2,
                           double *const *out) {
  CR3 a, b;
  R6 o;
  for (int c = 0; c < 3; ++c) { a.p[c] = u[c]; b.p[c] = ub[c]; }
  for (int c = 0; c < 6; ++c) o.p[c] = out[c];
  LAUNCH(k_sgs_products, nreal(g), st, nreal(g), a, b, s2, o);
}
hipError_t op_sgs_force(const GridK &g, hipStream_t st, const cplx *const *tau, const cplx *const *f0,
                        cplx *const *out, const cplx *const *v, const double *nuk, int nb, double dkk) {
  C6 t;
  for (int c = 0; c < 6; ++c) t.p[c] = tau ? tau[c] : nullptr;
  SpecVisc sv;
  for (int c = 0; c < 3; ++c) sv.v[c] = v ? v[c] : nullptr;
  sv.nuk = v ? nuk : nullptr;
  sv.nb = nb;
  sv.dkk = dkk;
  LAUNCH(k_sgs_force, ncplx(g), st, g, t, sv, f0 ? f0[0] : nullptr, f0 ? f0[1] : nullptr, f0 ? f0[2] : nullptr,
         out[0], out[1], out[2]);
}
hipError_t op_sharp(const GridK &g, hipStream_t st, const cplx *a, cplx *b, double kc2) {
  LAUNCH(k_sharp, ncplx(g), st, g, a, b, kc2);
}
hipError_t op_zero(hipStream_t st, void *p, size_t bytes) {
  size_t n = bytes / sizeof(double);
  LAUNCH(k_zero, n, st, n, (double *)p);
}
hipError_t op_cross_real(const GridK &g, hipStream_t st, double *r1, double *r2, double *r3, const double *r4,
                         const double *r5, const double *r6, double *r7) {
  LAUNCH(k_cross_real, nreal(g), st, nreal(g), g.inv_n6, r1, r2, r3, r4, r5, r6, r7);
}
hipError_t op_rk_update(const GridK &g, hipStream_t st, cplx *v, const cplx *c, const cplx *lap, const cplx *nl,
                        const cplx *f, double nu, double dt, double rmp) {
  LAUNCH(k_rk_update, ncplx(g), st, g, v, c, lap, nl, f, nu, dt, rmp);
}
hipError_t op_bouss3_nonl(const GridK &g, hipStream_t st, int mode, const cplx *const *w, const cplx *th,
                          double xmom, const cplx *const *c, cplx *const *out) {
  const cplx *c1 = c ? c[0] : nullptr, *c2 = c ? c[1] : nullptr, *c3 = c ? c[2] : nullptr;
  LAUNCH(k_bouss3_nonl, ncplx(g), st, g, mode, w[0], w[1], w[2], th, xmom, c1, c2, c3, out[0], out[1], out[2]);
}
hipError_t op_bouss3_add(const GridK &g, hipStream_t st, cplx *const *v, const cplx *const *x, double rmp) {
  LAUNCH(k_bouss3_add, ncplx(g), st, ncplx(g), v[0], v[1], v[2], x[0], x[1], x[2], rmp);
}
hipError_t op_tg_real(const GridK &g, hipStream_t st, double *r1, double *r2, int kdn, int kup) {
  LAUNCH(k_tg_real, nreal(g), st, g, r1, r2, kdn, kup);
}

size_t reduce_work_doubles() { return kRedBlocks + 8; }

hipError_t op_reduce(const GridK &g, hipStream_t st, int kind, const cplx *a, const cplx *b, const cplx *c,
                     const cplx *d, const cplx *e, const cplx *f, double *work, double *out) {
  unsigned nb = nblocks(ncplx(g));
  if (nb > (unsigned)kRedBlocks) nb = kRedBlocks;
  hipLaunchKernelGGL(k_reduce, dim3(nb), dim3(kBlock), 0, st, g, kind, a, b, c, d, e, f, work);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, st, work, (int)nb, work + kRedBlocks);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = hipMemcpyAsync(out, work + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(st);
}

static unsigned spectrum_blocks(const GridK &g) {
  unsigned nb = nblocks(ncplx(g));
  return nb > 1024u ? 1024u : nb;
}
size_t spectrum_work_doubles(const GridK &g, int nb) { return (size_t)spectrum_blocks(g) * nb + nb; }

hipError_t op_spectrum(const GridK &g, hipStream_t st, int kind, double dkk, int nb, const cplx *a, const cplx *b,
                       const cplx *c, const cplx *d, const cplx *e, const cplx *f, double *work, double *out,
                       double off) {
  unsigned nblk = spectrum_blocks(g);
  hipLaunchKernelGGL(k_spectrum, dim3(nblk), dim3(kBlock), nb * sizeof(double), st, g, kind, dkk, nb, off, a, b, c,
                     d, e, f, work);
  double *res = work + (size_t)nblk * nb;
  hipLaunchKernelGGL(k_spectrum_sum, dim3((nb + kBlock - 1) / kBlock), dim3(kBlock), 0, st, work, (int)nblk, nb, res);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = hipMemcpyAsync(out, res, sizeof(double) * nb, hipMemcpyDeviceToHost, st);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(st);
}

size_t axispec_work_doubles(const GridK &g, int nb) { return (size_t)spectrum_blocks(g) * 3 * nb + 3 * nb; }

hipError_t op_axispec(const GridK &g, hipStream_t st, int bins, int kind, int nb, double dkk, double lz,
                      const cplx *a, const cplx *b, const cplx *c, const cplx *d, const cplx *e, const cplx *f,
                      double *work, double *out) {
  unsigned nblk = spectrum_blocks(g);
  hipLaunchKernelGGL(k_axispec, dim3(nblk), dim3(kBlock), 3 * nb * sizeof(double), st, g, bins, kind, nb, dkk, lz, a,
                     b, c, d, e, f, work);
  double *res = work + (size_t)nblk * 3 * nb;
  hipLaunchKernelGGL(k_spectrum_sum, dim3((3 * nb + kBlock - 1) / kBlock), dim3(kBlock), 0, st, work, (int)nblk,
                     3 * nb, res);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = hipMemcpyAsync(out, res, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, st);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(st);
}

hipError_t op_spec2d(const GridK &g, hipStream_t st, int kind, int np, int nzb, double dkk, double lz, const cplx *a,
                     const cplx *b, const cplx *c, const cplx *d, const cplx *e, const cplx *f, double *work,
                     double *out) {
  const size_t n = (size_t)np * nzb;
  hipError_t err = op_zero(st, work, n * sizeof(double));
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(k_spec2d, dim3(spectrum_blocks(g)), dim3(kBlock), 0, st, g, kind, np, nzb, dkk, lz, a, b, c, d,
                     e, f, work);
  err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = hipMemcpyAsync(out, work, sizeof(double) * n, hipMemcpyDeviceToHost, st);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(st);
}

hipError_t op_plane_means(const GridK &g, hipStream_t st, const double *r1, const double *r2, int dir, double tmp,
                          double tmq, double *work, double *out) {
  const int n = dir == 1 ? g.nx : g.nz;
  hipLaunchKernelGGL(k_plane_means, dim3(n), dim3(kBlock), 0, st, r1, r2, g.nx, g.ny, g.nz, dir, tmp, tmq, work);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = hipMemcpyAsync(out, work, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, st);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(st);
}

hipError_t op_zplane(const GridK &g, hipStream_t st, int kind, const double *r1, const double *r2, const double *r3,
                     const double *r4, double fo, double bv, double scale, double *work, double *out) {
  const double t = 1.0 / ((double)g.nx * (double)g.ny * (double)g.nz);
  hipLaunchKernelGGL(k_zplane, dim3(g.nz), dim3(kBlock), 0, st, kind, (size_t)g.nx * g.ny, r1, r2, r3, r4, fo, bv, t,
                     scale, work);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = hipMemcpyAsync(out, work, sizeof(double) * g.nz, hipMemcpyDeviceToHost, st);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(st);
}

hipError_t op_spectr1d(const GridK &g, hipStream_t st, int dir, double ldir, int nb, const cplx *a, double *work,
                       double *out) {
  unsigned nblk = spectrum_blocks(g);
  hipLaunchKernelGGL(k_spectr1d, dim3(nblk), dim3(kBlock), nb * sizeof(double), st, g, dir, ldir, nb, a, work);
  double *res = work + (size_t)nblk * nb;
  hipLaunchKernelGGL(k_spectrum_sum, dim3((nb + kBlock - 1) / kBlock), dim3(kBlock), 0, st, work, (int)nblk, nb, res);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = hipMemcpyAsync(out, res, sizeof(double) * nb, hipMemcpyDeviceToHost, st);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(st);
}

hipError_t op_maxnorm(const GridK &g, hipStream_t st, const double *r1, const double *r2, const double *r3,
                      double *work, double *out) {
  unsigned nb = nblocks(nreal(g));
  if (nb > (unsigned)kRedBlocks) nb = kRedBlocks;
  hipLaunchKernelGGL(k_maxnorm, dim3(nb), dim3(kBlock), 0, st, nreal(g), r1, r2, r3, work);
  hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(kBlock), 0, st, work, (int)nb, work + kRedBlocks);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  err = hipMemcpyAsync(out, work + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st);
  if (err != hipSuccess) return err;
  return hipStreamSynchronize(st);
}

}  // namespace ghx
