// ec_metrics.hip -- whole-field diagnostics as one-pass device reductions:
// the admissibility check of the IMEX loop, the Gresho error and kinetic-
// energy metrics and the shock-bubble metrics.  Citations are relative to
// amrex/apps/euler_compare/src/euler_compare/.
//
// Each is one read of the state (32 B per cell; the shock-bubble schlieren
// also reads the density of the four neighbours, mostly from cache) reduced
// per workgroup (wavefront shuffles, then the waves in order) into a partial
// record; the host folds the partials in workgroup order.  Minima, maxima and
// counts are therefore exact; sums differ from the reference's serial order
// by rounding only.
#include <cmath>
#include <cstring>
#include <vector>

#include "ec_context.hpp"
#include "ec_ghost.hpp"
#include "ec_state.hpp"

namespace ec {

namespace {

constexpr int MB = 256;
constexpr int kMaxBlocks = 1024;
enum Op { OP_MIN = 0, OP_MAX = 1, OP_SUM = 2 };

__host__ __device__ __forceinline__ double combine(int op, double a, double b)
{
  if (op == OP_MIN) return b < a ? b : a;
  if (op == OP_MAX) return a < b ? b : a;
  return a + b;
}

// Reduces v over the workgroup and stores it as partial blockIdx.x.
template <int K>
__device__ __forceinline__ void fold_block(double (&v)[K], const int (&ops)[K], double* __restrict__ part)
{
  __shared__ double sh[MB / 64][K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    for (int o = 32; o > 0; o >>= 1) v[k] = combine(ops[k], v[k], __shfl_xor(v[k], o));
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) sh[threadIdx.x >> 6][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < K; ++k) {
      double r = sh[0][k];
      for (int w = 1; w < MB / 64; ++w) r = combine(ops[k], r, sh[w][k]);
      part[(long long)blockIdx.x * K + k] = r;
    }
  }
}


// compute_admissibility_diagnostics (numerics/diagnostic_metrics_timestep.hpp:49-102)
constexpr int KA = 7;
const int kAdmOpsHost[KA] = {OP_MIN, OP_MIN, OP_MIN, OP_SUM, OP_SUM, OP_SUM, OP_SUM};

__global__ void __launch_bounds__(MB) admissibility_kernel(const double* __restrict__ U, long long rs,
                                                            long long cs, int nx, int ny, double gamma,
                                                            double* __restrict__ part)
{
  const int ops[KA] = {OP_MIN, OP_MIN, OP_MIN, OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  double v[KA] = {INFINITY, INFINITY, INFINITY, 0.0, 0.0, 0.0, 0.0};
  const long long n = (long long)nx * ny;
  for (long long q = (long long)blockIdx.x * MB + threadIdx.x; q < n; q += (long long)gridDim.x * MB) {
    const long long k = (q / nx) * rs + q % nx;
    const double r = U[k], mx = U[cs + k], my = U[2 * cs + k], E = U[3 * cs + k];
    const bool cons_bad = !isfinite(r) || !isfinite(mx) || !isfinite(my) || !isfinite(E);
    const bool density_bad = !cons_bad && r <= 0.0;
    double ie = -INFINITY, p = -INFINITY;
    bool thermo_bad = false;
    if (!cons_bad && !density_bad) {
      const double kinetic = 0.5 * (mx * mx + my * my) / r;
      ie = E - kinetic;
      p = (gamma - 1.0) * ie;
      thermo_bad = !isfinite(kinetic) || !isfinite(ie) || !isfinite(p);
    }
    const bool bad = cons_bad || thermo_bad;
    const bool p_bad = !bad && !density_bad && p <= 0.0;
    const bool ie_bad = !bad && !density_bad && ie <= 0.0;
    v[0] = smin(v[0], bad ? -INFINITY : r);
    v[1] = smin(v[1], (bad || density_bad) ? -INFINITY : p);
    v[2] = smin(v[2], (bad || density_bad) ? -INFINITY : ie);
    v[3] += bad ? 1.0 : 0.0;
    v[4] += density_bad ? 1.0 : 0.0;
    v[5] += p_bad ? 1.0 : 0.0;
    v[6] += ie_bad ? 1.0 : 0.0;
  }
  fold_block(v, ops, part);
}

// compute_gresho_metrics (numerics/diagnostic_metrics_timestep.hpp:147-225)
constexpr int KG = 7;

struct GreshoArgs {
  int nx, ny;
  double xlo, ylo, dx, dy, gamma, p0, cell_vol;
  double cx, cy, rho, vscale;
};

// gresho_state (problems/problem_state_functions.hpp:2-30)
__device__ Prim gresho_exact(const GreshoArgs& A, double x, double y)
{
  const double ddx = x - A.cx, ddy = y - A.cy;
  const double r = sqrt(ddx * ddx + ddy * ddy);
  double tf = 0.0;
  double pf = -2.0 + 4.0 * log(2.0);
  if (r < 0.2) {
    tf = 5.0 * r;
    pf = 12.5 * r * r;
  } else if (r < 0.4) {
    tf = 2.0 - 5.0 * r;
    pf = 12.5 * r * r + 4.0 * (1.0 - 5.0 * r - log(0.2) + log(r));
  }
  const double vt = A.vscale * tf;
  const double p = A.p0 + A.rho * A.vscale * A.vscale * pf;
  if (r < 1.0e-14) return Prim{A.rho, 0.0, 0.0, p};
  return Prim{A.rho, -vt * ddy / r, vt * ddx / r, p};
}

__global__ void __launch_bounds__(MB) gresho_metrics_kernel(const double* __restrict__ U, long long rs,
                                                             long long cs, GreshoArgs A,
                                                             double* __restrict__ part)
{
  const int ops[KG] = {OP_SUM, OP_SUM, OP_SUM, OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  double v[KG] = {0, 0, 0, 0, 0, 0, 0};
  const long long n = (long long)A.nx * A.ny;
  for (long long q = (long long)blockIdx.x * MB + threadIdx.x; q < n; q += (long long)gridDim.x * MB) {
    const int i = (int)(q % A.nx), j = (int)(q / A.nx);
    const long long k = (long long)j * rs + i;
    const double x = A.xlo + ((double)i + 0.5) * A.dx;
    const double y = A.ylo + ((double)j + 0.5) * A.dy;
    const Prim num = to_prim(Cons{U[k], U[cs + k], U[2 * cs + k], U[3 * cs + k]}, A.gamma);
    const Prim ex = gresho_exact(A, x, y);
    const double du = num.u - ex.u, dv = num.v - ex.v;
    v[0] += fabs(num.r - ex.r);
    v[1] += sqrt(du * du + dv * dv);
    v[2] += fabs(num.p - ex.p);
    v[3] += fabs((num.p - A.p0) - (ex.p - A.p0));
    v[4] += fabs(ex.p - A.p0);
    v[5] += 0.5 * ex.r * (ex.u * ex.u + ex.v * ex.v) * A.cell_vol;
    v[6] += 0.5 * num.r * (num.u * num.u + num.v * num.v) * A.cell_vol;
  }
  fold_block(v, ops, part);
}

// compute_shock_density_bubble_metrics (io/output_cases.hpp:329-407), the
// neighbour densities through the context's ghost rule.
constexpr int KS = 10;
const int kSbOpsHost[KS] = {OP_MIN, OP_MAX, OP_MIN, OP_MAX, OP_MIN, OP_SUM, OP_SUM, OP_SUM, OP_SUM, OP_MAX};

__global__ void __launch_bounds__(MB) shock_bubble_kernel(const double* __restrict__ U, long long rs,
                                                           long long cs, GridParams g, double xlo, double ylo,
                                                           double threshold, double* __restrict__ part)
{
  const int ops[KS] = {OP_MIN, OP_MAX, OP_MIN, OP_MAX, OP_MIN, OP_SUM, OP_SUM, OP_SUM, OP_SUM, OP_MAX};
  // schlieren_max starts at 0 (io/output_cases.hpp:273)
  double v[KS] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double cell_vol = g.dx * g.dy;
  const long long n = (long long)g.nx * g.ny;
  for (long long q = (long long)blockIdx.x * MB + threadIdx.x; q < n; q += (long long)gridDim.x * MB) {
    const int i = (int)(q % g.nx), j = (int)(q / g.nx);
    const long long k = (long long)j * rs + i;
    const Cons u{U[k], U[cs + k], U[2 * cs + k], U[3 * cs + k]};
    const Prim w = to_prim(u, g.gamma);
    const double kinetic = u.r > 0.0 ? 0.5 * (u.mx * u.mx + u.my * u.my) / u.r : NAN;
    const double internal = u.E - kinetic;
    const bool bad = !isfinite(u.r) || !isfinite(u.mx) || !isfinite(u.my) || !isfinite(u.E) ||
                     !isfinite(w.p) || !isfinite(internal);
    if (bad) {
      v[5] += 1.0;
      continue;
    }
    v[0] = smin(v[0], w.r);
    v[1] = smax(v[1], w.r);
    v[2] = smin(v[2], w.p);
    v[3] = smax(v[3], w.p);
    v[4] = smin(v[4], internal);
    const double rho_x = (load_cell(U, rs, cs, g, i + 1, j).r - load_cell(U, rs, cs, g, i - 1, j).r) / (2.0 * g.dx);
    const double rho_y = (load_cell(U, rs, cs, g, i, j + 1).r - load_cell(U, rs, cs, g, i, j - 1).r) / (2.0 * g.dy);
    const double grad = sqrt(rho_x * rho_x + rho_y * rho_y);
    if (isfinite(grad)) v[9] = smax(v[9], grad);
    if (w.r < threshold) {
      const double x = xlo + ((double)i + 0.5) * g.dx;
      const double y = ylo + ((double)j + 0.5) * g.dy;
      v[6] += cell_vol;
      v[7] += x * cell_vol;
      v[8] += y * cell_vol;
    }
  }
  fold_block(v, ops, part);
}

int blocks(long long n)
{
  long long b = (n + MB - 1) / MB;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return b < 1 ? 1 : (int)b;
}

constexpr int kMaxK = 16;

// Launches through `launch` into the context's partial buffer, then folds
// the nb partials of K values on the host in workgroup order.
template <int K, typename L>
ec_status reduce(ec_ctx* ctx, int nb, const int (&ops)[K], double (&out)[K], const L& launch)
{
  static_assert(K <= kMaxK, "partial record too long");
  if (!ctx->metrics_part) {
    if (hipMalloc(&ctx->metrics_part, sizeof(double) * kMaxK * kMaxBlocks) != hipSuccess) {
      return EC_ERR_OUT_OF_MEMORY;
    }
    if (hipHostMalloc(&ctx->metrics_part_host, sizeof(double) * kMaxK * kMaxBlocks) != hipSuccess) {
      return EC_ERR_OUT_OF_MEMORY;
    }
  }
  launch(ctx->metrics_part);
  EC_HIP_TRY(ctx, hipGetLastError());
  EC_HIP_TRY(ctx, hipMemcpyAsync(ctx->metrics_part_host, ctx->metrics_part, sizeof(double) * K * nb,
                                 hipMemcpyDeviceToHost, ctx->stream));
  EC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const double* h = ctx->metrics_part_host;
  for (int k = 0; k < K; ++k) {
    double r = h[k];
    for (int b = 1; b < nb; ++b) r = combine(ops[k], r, h[(size_t)b * K + k]);
    out[k] = r;
  }
  return EC_OK;
}

bool layout_ok(const ec_ctx* ctx, const ec_layout* l)
{
  return l && l->row_stride >= ctx->grid.nx && l->comp_stride >= l->row_stride * ctx->grid.ny;
}

}  // namespace

}  // namespace ec

using namespace ec;

extern "C" ec_status ec_admissibility(ec_ctx* ctx, const double* U, const ec_layout* lay,
                                      ec_admissibility_diag* out)
{
  if (!ctx || !U || !out || !layout_ok(ctx, lay)) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int nb = blocks((long long)ctx->grid.nx * ctx->grid.ny);
  double v[KA];
  const ec_status st = reduce<KA>(ctx, nb, kAdmOpsHost, v, [&](double* part) {
    hipLaunchKernelGGL(admissibility_kernel, dim3(nb), dim3(MB), 0, ctx->stream, U, lay->row_stride,
                       lay->comp_stride, ctx->grid.nx, ctx->grid.ny, ctx->grid.gamma, part);
  });
  if (st != EC_OK) return st;
  out->rho_min = v[0];
  out->pressure_min = v[1];
  out->internal_energy_min = v[2];
  out->nonfinite_count = (int)v[3];
  out->density_failure_count = (int)v[4];
  out->pressure_failure_count = (int)v[5];
  out->internal_energy_failure_count = (int)v[6];
  return EC_OK;
}

extern "C" ec_status ec_gresho_metrics(ec_ctx* ctx, const double* U, const ec_layout* lay,
                                       const ec_problem_desc* pd, ec_gresho_metrics_out* out)
{
  if (!ctx || !U || !out || !pd || !layout_ok(ctx, lay)) return EC_ERR_INVALID_ARGUMENT;
  if (pd->nx != ctx->grid.nx || pd->ny != ctx->grid.ny) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  GreshoArgs A;
  A.nx = pd->nx;
  A.ny = pd->ny;
  A.xlo = pd->prob_lo[0];
  A.ylo = pd->prob_lo[1];
  A.dx = (pd->prob_hi[0] - pd->prob_lo[0]) / pd->nx;
  A.dy = (pd->prob_hi[1] - pd->prob_lo[1]) / pd->ny;
  A.gamma = pd->gamma;
  A.p0 = pd->pressure;
  A.cell_vol = A.dx * A.dy;
  A.cx = pd->vortex_center[0];
  A.cy = pd->vortex_center[1];
  A.rho = pd->density_outer;
  // core/state_primitives.hpp:207-215
  const double c0 = std::sqrt(pd->gamma * pd->pressure / pd->density_outer);
  A.vscale = pd->mach * c0;
  const int nb = blocks((long long)A.nx * A.ny);
  double v[KG];
  const int ops[KG] = {OP_SUM, OP_SUM, OP_SUM, OP_SUM, OP_SUM, OP_SUM, OP_SUM};
  const ec_status st = reduce<KG>(ctx, nb, ops, v, [&](double* part) {
    hipLaunchKernelGGL(gresho_metrics_kernel, dim3(nb), dim3(MB), 0, ctx->stream, U, lay->row_stride,
                       lay->comp_stride, A, part);
  });
  if (st != EC_OK) return st;
  const double inv_count = 1.0 / (double)(pd->nx * pd->ny);
  out->density_l1_error = v[0] * inv_count;
  out->velocity_l1_error = v[1] * inv_count;
  out->pressure_l1_error = v[2] * inv_count;
  out->pressure_perturbation_l1_error = v[3] * inv_count;
  out->pressure_perturbation_l1_relative_error = v[4] > 1.0e-30 ? v[3] / v[4] : NAN;
  out->kinetic_energy_initial = v[5];
  out->kinetic_energy_final = v[6];
  out->kinetic_energy_ratio = v[5] > 0.0 ? v[6] / v[5] : NAN;
  out->reference_sound_speed = c0;
  out->velocity_scale = A.vscale;
  return EC_OK;
}

extern "C" ec_status ec_shock_bubble_metrics(ec_ctx* ctx, const double* U, const ec_layout* lay,
                                             const double* prob_lo, double threshold,
                                             ec_shock_bubble_metrics_out* out)
{
  if (!ctx || !U || !out || !prob_lo || !layout_ok(ctx, lay)) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int nb = blocks((long long)ctx->grid.nx * ctx->grid.ny);
  double v[KS];
  const ec_status st = reduce<KS>(ctx, nb, kSbOpsHost, v, [&](double* part) {
    hipLaunchKernelGGL(shock_bubble_kernel, dim3(nb), dim3(MB), 0, ctx->stream, U, lay->row_stride,
                       lay->comp_stride, ctx->grid, prob_lo[0], prob_lo[1], threshold, part);
  });
  if (st != EC_OK) return st;
  out->rho_min = v[0];
  out->rho_max = v[1];
  out->pressure_min = v[2];
  out->pressure_max = v[3];
  out->internal_energy_min = v[4];
  out->nonfinite_count = (int)v[5];
  out->density_bubble_area_threshold = threshold;
  out->density_bubble_area = v[6];
  out->density_bubble_centroid_x = v[6] > 0.0 ? v[7] / v[6] : NAN;
  out->density_bubble_centroid_y = v[6] > 0.0 ? v[8] / v[6] : NAN;
  out->schlieren_max = v[9] < 0.0 ? 0.0 : v[9];
  return EC_OK;
}
