// tracers.h -- Lagrangian tracer kernels: periodic cubic-spline interpolation
// of real fields at particle positions and the RK position update of GHOST's
// lagpart module (single rank, positions in grid units).
#pragma once
#include "ghx_common.h"

namespace ghx {

struct TracerStage {
  const double *coef3;    // spline control points of u_x, u_y, u_z interleaved: c[((z ny + y) nx + x) 3 + comp]
  double *pos[3];         // positions, updated in place
  const double *x0[3];    // positions at the start of the RK step
  double *lv[3];          // interpolated velocities (may be null)
  double dtf[3];          // dt * xk * n_c / (2 pi L_c)
  int nx, ny, nz;
  long long np;
  int zcompat;  // 1: GHOST's z stencil (one cell up, clamped top cell)
};

// Interpolate the three fields at pos, then pos = wrap(x0 + dtf * u).
hipError_t tracer_stage(hipStream_t s, const TracerStage &p);
// out[q] = spline interpolant of coef at pos[q] (no position update).
hipError_t spline_interp(hipStream_t s, const double *coef, const double *const *pos, long long np, int nx,
                         int ny, int nz, int zcompat, double *out);
// x0 = pos
hipError_t tracer_copy(hipStream_t s, double *const *dst, const double *const *src, long long np);

}  // namespace ghx
