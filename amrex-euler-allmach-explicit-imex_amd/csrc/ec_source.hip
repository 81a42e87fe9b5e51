// ec_source.hip -- the radial-symmetry geometric source of the cylindrical
// shock-density bubble (euler.problem=shock_density_bubble_cylindrical_clawpack),
// applied in place after each explicit update.  Citations are relative to
// amrex/apps/euler_compare/src/euler_compare/.
//
// apply_shock_density_bubble_cylindrical_source (problems/boundary_conditions.hpp:
// 75-131) treats y as the radius r of an axisymmetric flow and integrates
// dq/dt = -(rho v, rho u v, rho v^2, v (E + p)) / r over dt with the midpoint
// rule (a half step from q0 to q*, then the full step from q0 with the source
// at q*).  It is pointwise: one read and one write of the state, 64 B per
// cell, HBM-bound; one thread per cell in row order keeps every access
// coalesced.  The arithmetic follows the reference's operation order (the
// build contracts no multiply-add), so the result is bit-identical to it.
#include <cmath>

#include "ec_context.hpp"

namespace ec {

namespace {

constexpr int SB = 256;

struct SourceCons {
  double rho, mx, my, E;
};

__device__ __forceinline__ SourceCons source_rate_step(const SourceCons& q0, const SourceCons& at, double h,
                                                       double r, double gm1)
{
  // The source evaluated at state 'at', applied to q0 over h.
  const double inv = at.rho > 0.0 ? 1.0 / at.rho : 0.0;
  const double u = at.mx * inv;
  const double v = at.my * inv;
  const double p = gm1 * (at.E - 0.5 * (at.mx * at.mx + at.my * at.my) * inv);
  SourceCons q;
  q.rho = q0.rho - h * at.my / r;
  q.mx = q0.mx - h * at.rho * u * v / r;
  q.my = q0.my - h * at.rho * v * v / r;
  q.E = q0.E - h * v * (at.E + p) / r;
  return q;
}

__global__ void __launch_bounds__(SB) cylindrical_source_kernel(double* __restrict__ U, long long rs,
                                                                  long long cs, int nx, int ny, double y_lo,
                                                                  double dy, double dt, double gm1)
{
  const long long n = (long long)nx * ny;
  const double half_dt = 0.5 * dt;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(t / nx), i = (int)(t - (long long)j * nx);
    const double r = y_lo + ((double)j + 0.5) * dy;
    if (!(r > 0.0)) continue;
    const long long k = (long long)j * rs + i;
    const SourceCons q0{U[k], U[cs + k], U[2 * cs + k], U[3 * cs + k]};
    const SourceCons qs = source_rate_step(q0, q0, half_dt, r, gm1);
    const SourceCons q1 = source_rate_step(q0, qs, dt, r, gm1);
    U[k] = q1.rho;
    U[cs + k] = q1.mx;
    U[2 * cs + k] = q1.my;
    U[3 * cs + k] = q1.E;
  }
}

}  // namespace

}  // namespace ec

using namespace ec;

extern "C" ec_status ec_cylindrical_source(ec_ctx* ctx, double* U, const ec_layout* lay, double dt, double y_lo)
{
  if (!ctx || !U || !lay) return EC_ERR_INVALID_ARGUMENT;
  const int nx = ctx->grid.nx, ny = ctx->grid.ny;
  if (lay->row_stride < nx || lay->comp_stride < lay->row_stride * ny) return EC_ERR_INVALID_ARGUMENT;
  if (!std::isfinite(dt) || !std::isfinite(y_lo)) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const long long n = (long long)nx * ny;
  long long nb = (n + SB - 1) / SB;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(cylindrical_source_kernel, dim3((unsigned)nb), dim3(SB), 0, ctx->stream, U,
                     lay->row_stride, lay->comp_stride, nx, ny, y_lo, ctx->grid.dy, dt,
                     ctx->grid.gamma - 1.0);
  EC_HIP_TRY(ctx, hipGetLastError());
  return EC_OK;
}
