// ec_direct.hip -- the direct BDLTV20 T1/S2 step on the device (the driver
// behind every IMEX row of the reference's reports:
// amrex/apps/euler_compare/src/euler_compare/imex/
// imex_bdltv20_direct_driver.hpp:44-779).
//
// The reference runs it on the host: cell loops over std::vector copies of
// the state, an Eigen triplet assembly of the pressure rows with the
// exact-Dirichlet boundary pressures folded into the right-hand side, and a
// host GMRES.  Here the state stays in HBM and each stage is one kernel:
//   star       epsilon-scaled conserved minmod + material Lax-Friedrichs
//              predictor (:176-265), boundary states at the step's start time
//   correct    lagged momentum m* - dt grad p / (2 eps dx) (:301-321), x only
//              for the Riemann strips
//   faces      face enthalpy (eq. 46, zero-denominator guard 1e-12) on every
//              face from the corrected cells (:322-383)
//   rhs        eps (E* - K - dt div(h m*)) minus the Dirichlet columns of the
//              row (:449-494)
//   solve      the IMEX module's matrix-free operator (diagonal eps/(g-1),
//              cy = 0 for the x-only row) and its FP64 or mixed GMRES
//   closeout   E = p/(g-1) + eps |m|^2/(2 rho) (:558-598)
// Boundary states sample the problem's exact solution at the domain edge
// (:101-131): Gresho from a host-evaluated table, Toro from the exact Riemann
// solution's host constants (the device evaluates pow only inside a fan).
#include <cmath>
#include <cstring>
#include <vector>

#include "ec_context.hpp"
#include "ec_ghost.hpp"
#include "ec_imex_api.hpp"
#include "ec_problems.hpp"
#include "ec_state.hpp"

namespace ec {

void toro_constants(ExactBc& e, const prob::P4& L, const prob::P4& R, double gamma);

namespace {

constexpr int DB = 256;

// Boundary description of one direct case.
struct DirectBc {
  ExactBc e;          // EXACT_TORO constants, EXACT_RING Gresho table, or EXACT_NONE
  double eps;
  double xlo, xhi, ylo, yhi;
};

struct DirectArgs {
  int nx, ny, px, py, x_only;
  double dx, dy, gamma, eps;
  const DirectBc* bc;  // device
  int* counts;         // device: [0] guards, [1] nonfinite, [2] nonpositive
};

struct DirectWork {
  DirectBc* bc = nullptr;
  double* ring = nullptr;
  int* counts = nullptr;
  int* counts_host = nullptr;
  double* part = nullptr;
  double* part_host = nullptr;
  int nparts = 0;
  int dcase = -1;
  int px = 1, py = 1, x_only = 0;
  double eps = 1.0;
};

__host__ __device__ __forceinline__ Cons eps_cons(const Prim& q, double gamma, double eps)
{
  // core/state_primitives.hpp:130-139
  const double kinetic = 0.5 * q.r * (q.u * q.u + q.v * q.v);
  return Cons{q.r, q.r * q.u, q.r * q.v, q.p / (gamma - 1.0) + eps * kinetic};
}

__host__ __device__ __forceinline__ Prim eps_prim(const Cons& u, double gamma, double eps)
{
  // core/state_primitives.hpp:119-128
  const double inv_rho = 1.0 / u.r;
  const double ux = u.mx * inv_rho;
  const double uy = u.my * inv_rho;
  const double kinetic = 0.5 * u.r * (ux * ux + uy * uy);
  return Prim{u.r, ux, uy, (gamma - 1.0) * (u.E - eps * kinetic)};
}

// exact_boundary_state (:101-131) of an outside index.
__device__ Cons boundary_state(const DirectArgs& A, int i, int j, double t)
{
  const DirectBc& d = *A.bc;
  if (d.e.kind == EXACT_RING) {
    const int ic = i < -2 ? -2 : (i > A.nx + 1 ? A.nx + 1 : i);
    const int jc = j < -2 ? -2 : (j > A.ny + 1 ? A.ny + 1 : j);
    const double* r = d.e.ring + 4 * ring_index(ic, jc, A.nx, A.ny);
    return Cons{r[0], r[1], r[2], r[3]};
  }
  // Toro: the solution depends on x only
  const int ic = i < 0 ? 0 : (i > A.nx - 1 ? A.nx - 1 : i);
  double x = d.e.xlo + ((double)ic + 0.5) * d.e.dx;
  if (i < 0) x = d.xlo;
  else if (i >= A.nx) x = d.xhi;
  return eps_cons(toro_sample(d.e, x, t), A.gamma, A.eps);
}

__device__ __forceinline__ double boundary_pressure(const DirectArgs& A, int i, int j, double t)
{
  return eps_prim(boundary_state(A, i, j, t), A.gamma, A.eps).p;
}

// Wrap a periodic axis; true if (i, j) is then a cell of the grid.
__device__ __forceinline__ bool inside(const DirectArgs& A, int& i, int& j)
{
  if (A.px) {
    const int m = i % A.nx;
    i = m < 0 ? m + A.nx : m;
  }
  if (A.py) {
    const int m = j % A.ny;
    j = m < 0 ? m + A.ny : m;
  }
  return i >= 0 && i < A.nx && j >= 0 && j < A.ny;
}

// old_cell_from / cell_state_from (:176-187, :275-288) on a layout
__device__ __forceinline__ Cons cell_at(const DirectArgs& A, const double* __restrict__ U, long long rs,
                                        long long cs, int i, int j, double t)
{
  int ii = i, jj = j;
  if (!inside(A, ii, jj)) return boundary_state(A, i, j, t);
  const long long k = (long long)jj * rs + ii;
  return Cons{U[k], U[cs + k], U[2 * cs + k], U[3 * cs + k]};
}

__device__ __forceinline__ double pressure_at(const DirectArgs& A, const double* __restrict__ P, int i,
                                              int j, double t)
{
  int ii = i, jj = j;
  if (!inside(A, ii, jj)) return boundary_pressure(A, i, j, t);
  return P[(long long)jj * A.nx + ii];
}

// Corrected cell: the star density and energy with the lagged momentum.
__device__ __forceinline__ Cons corrected_at(const DirectArgs& A, const double* __restrict__ S,
                                             const double* __restrict__ M, int i, int j, double t)
{
  int ii = i, jj = j;
  if (!inside(A, ii, jj)) return boundary_state(A, i, j, t);
  const long long n = (long long)A.nx * A.ny, k = (long long)jj * A.nx + ii;
  return Cons{S[k], M[k], M[n + k], S[3 * n + k]};
}

__device__ __forceinline__ double mm(double a, double b)
{
  if (a * b <= 0.0) return 0.0;
  return fabs(a) < fabs(b) ? a : b;
}

// reconstructed_old (:188-203)
__device__ Cons recon(const DirectArgs& A, const double* U, long long rs, long long cs, int i, int j,
                      int axis, double side, double t)
{
  const Cons um = cell_at(A, U, rs, cs, axis == 0 ? i - 1 : i, axis == 0 ? j : j - 1, t);
  const Cons uc = cell_at(A, U, rs, cs, i, j, t);
  const Cons up = cell_at(A, U, rs, cs, axis == 0 ? i + 1 : i, axis == 0 ? j : j + 1, t);
  return Cons{uc.r + side * mm(uc.r - um.r, up.r - uc.r), uc.mx + side * mm(uc.mx - um.mx, up.mx - uc.mx),
              uc.my + side * mm(uc.my - um.my, up.my - uc.my), uc.E + side * mm(uc.E - um.E, up.E - uc.E)};
}

// explicit_flux (:204-233)
__device__ Cons lf(const DirectArgs& A, const Cons& lu, const Cons& ru, int axis)
{
  const Prim l = eps_prim(lu, A.gamma, A.eps), r = eps_prim(ru, A.gamma, A.eps);
  const double mx_l = l.r * l.u, my_l = l.r * l.v, mx_r = r.r * r.u, my_r = r.r * r.v;
  const double k_l = A.eps * 0.5 * l.r * (l.u * l.u + l.v * l.v);
  const double k_r = A.eps * 0.5 * r.r * (r.u * r.u + r.v * r.v);
  const double un_l = axis == 0 ? l.u : l.v, un_r = axis == 0 ? r.u : r.v;
  const Cons fl{axis == 0 ? mx_l : my_l, mx_l * un_l, my_l * un_l, k_l * un_l};
  const Cons fr{axis == 0 ? mx_r : my_r, mx_r * un_r, my_r * un_r, k_r * un_r};
  const double s = smax(fabs(un_l), fabs(un_r));
  const double hs = 0.5 * s;
  return Cons{0.5 * (fl.r + fr.r) - hs * (ru.r - lu.r), 0.5 * (fl.mx + fr.mx) - hs * (ru.mx - lu.mx),
              0.5 * (fl.my + fr.my) - hs * (ru.my - lu.my), 0.5 * (fl.E + fr.E) - hs * (ru.E - lu.E)};
}

// Star state and the old pressure (the first Picard initial guess).
__global__ void __launch_bounds__(DB) direct_star_kernel(DirectArgs A, const double* __restrict__ U,
                                                         long long rs, long long cs, double t, double dt,
                                                         double* __restrict__ S, double* __restrict__ P)
{
  const long long n = (long long)A.nx * A.ny;
  for (long long q = (long long)blockIdx.x * DB + threadIdx.x; q < n; q += (long long)gridDim.x * DB) {
    const int i = (int)(q % A.nx), j = (int)(q / A.nx);
    const Cons fxl = lf(A, recon(A, U, rs, cs, i - 1, j, 0, 0.5, t), recon(A, U, rs, cs, i, j, 0, -0.5, t), 0);
    const Cons fxr = lf(A, recon(A, U, rs, cs, i, j, 0, 0.5, t), recon(A, U, rs, cs, i + 1, j, 0, -0.5, t), 0);
    const Cons fyb = lf(A, recon(A, U, rs, cs, i, j - 1, 1, 0.5, t), recon(A, U, rs, cs, i, j, 1, -0.5, t), 1);
    const Cons fyt = lf(A, recon(A, U, rs, cs, i, j, 1, 0.5, t), recon(A, U, rs, cs, i, j + 1, 1, -0.5, t), 1);
    const long long k = (long long)j * rs + i;
    const Cons o{U[k], U[cs + k], U[2 * cs + k], U[3 * cs + k]};
    S[q] = o.r - dt * ((fxr.r - fxl.r) / A.dx + (fyt.r - fyb.r) / A.dy);
    S[n + q] = o.mx - dt * ((fxr.mx - fxl.mx) / A.dx + (fyt.mx - fyb.mx) / A.dy);
    S[2 * n + q] = o.my - dt * ((fxr.my - fxl.my) / A.dx + (fyt.my - fyb.my) / A.dy);
    S[3 * n + q] = o.E - dt * ((fxr.E - fxl.E) / A.dx + (fyt.E - fyb.E) / A.dy);
    P[q] = eps_prim(o, A.gamma, A.eps).p;
  }
}

// Lagged momentum (:301-321): the star momentum, corrected by the pressure
// gradient when use_p.
__global__ void __launch_bounds__(DB) direct_correct_kernel(DirectArgs A, const double* __restrict__ S,
                                                            const double* __restrict__ P, int use_p,
                                                            double dt, double t1, double* __restrict__ M)
{
  const long long n = (long long)A.nx * A.ny;
  for (long long q = (long long)blockIdx.x * DB + threadIdx.x; q < n; q += (long long)gridDim.x * DB) {
    const int i = (int)(q % A.nx), j = (int)(q / A.nx);
    double mx = S[n + q], my = S[2 * n + q];
    if (use_p) {
      const double pe = pressure_at(A, P, i + 1, j, t1), pw = pressure_at(A, P, i - 1, j, t1);
      mx -= dt * (pe - pw) / (2.0 * A.eps * A.dx);
      if (!A.x_only) {
        const double pn = pressure_at(A, P, i, j + 1, t1), ps = pressure_at(A, P, i, j - 1, t1);
        my -= dt * (pn - ps) / (2.0 * A.eps * A.dy);
      }
    }
    M[q] = mx;
    M[n + q] = my;
  }
}

// bdltv20_eq46_face_enthalpy
__device__ __forceinline__ double face_enthalpy(double rl, double pl, double ql, double rr, double pr,
                                                double qr, double gamma, int& guard)
{
  const double hl = gamma * pl / ((gamma - 1.0) * rl);
  const double hr = gamma * pr / ((gamma - 1.0) * rr);
  const double denom = ql + qr;
  const double scale = smax(1.0, fabs(ql) + fabs(qr));
  guard = fabs(denom) <= 1.0e-12 * scale ? 1 : 0;
  return guard ? 0.5 * (hl + hr) : (hl * ql + hr * qr) / denom;
}

// h on every face (:322-383).  Face f of row j lies between cells f-1 and f;
// on a periodic axis faces 0 and nx are one face (counted once).
__global__ void __launch_bounds__(DB) direct_faces_kernel(DirectArgs A, const double* __restrict__ S,
                                                          const double* __restrict__ M,
                                                          const double* __restrict__ P, double t1,
                                                          double* __restrict__ hx, double* __restrict__ hy,
                                                          float* __restrict__ hx32, float* __restrict__ hy32)
{
  const long long nxf = (long long)(A.nx + 1) * A.ny, nyf = (long long)A.nx * (A.ny + 1);
  int guards = 0, bad = 0;
  for (long long q = (long long)blockIdx.x * DB + threadIdx.x; q < nxf + nyf; q += (long long)gridDim.x * DB) {
    double h, m;
    int gd;
    bool dup;
    if (q < nxf) {
      const int f = (int)(q % (A.nx + 1)), j = (int)(q / (A.nx + 1));
      const Cons l = corrected_at(A, S, M, f - 1, j, t1), r = corrected_at(A, S, M, f, j, t1);
      const double pl = pressure_at(A, P, f - 1, j, t1), pr = pressure_at(A, P, f, j, t1);
      h = face_enthalpy(l.r, pl, l.mx, r.r, pr, r.mx, A.gamma, gd);
      m = 0.5 * (l.mx + r.mx);
      dup = A.px && f == A.nx;
      hx[q] = h;
      if (hx32) hx32[q] = (float)h;
    } else {
      const long long s = q - nxf;
      const int i = (int)(s % A.nx), f = (int)(s / A.nx);
      const Cons b = corrected_at(A, S, M, i, f - 1, t1), tp = corrected_at(A, S, M, i, f, t1);
      const double pb = pressure_at(A, P, i, f - 1, t1), pt = pressure_at(A, P, i, f, t1);
      h = face_enthalpy(b.r, pb, b.my, tp.r, pt, tp.my, A.gamma, gd);
      m = 0.5 * (b.my + tp.my);
      dup = A.py && f == A.ny;
      hy[s] = h;
      if (hy32) hy32[s] = (float)h;
    }
    if (!dup) {
      guards += gd;
      if (!isfinite(h) || !isfinite(m)) ++bad;
    }
  }
  if (guards) atomicAdd(&A.counts[0], guards);
  if (bad) atomicAdd(&A.counts[1], bad);
}

// Row right-hand side (:449-494).
__global__ void __launch_bounds__(DB) direct_rhs_kernel(DirectArgs A, const double* __restrict__ S,
                                                        const double* __restrict__ M,
                                                        const double* __restrict__ hx,
                                                        const double* __restrict__ hy, double dt, double t1,
                                                        double* __restrict__ b)
{
  const long long n = (long long)A.nx * A.ny;
  const double eos = A.eps / (A.gamma - 1.0);
  const double cx = 0.25 * dt * dt / (A.dx * A.dx);
  const double cy = A.x_only ? 0.0 : 0.25 * dt * dt / (A.dy * A.dy);
  int bad = 0;
  for (long long q = (long long)blockIdx.x * DB + threadIdx.x; q < n; q += (long long)gridDim.x * DB) {
    const int i = (int)(q % A.nx), j = (int)(q / A.nx);
    const double he = hx[(long long)j * (A.nx + 1) + i + 1], hw = hx[(long long)j * (A.nx + 1) + i];
    const double hn = hy[(long long)(j + 1) * A.nx + i], hs = hy[(long long)j * A.nx + i];
    const double rho = S[q], smx = S[n + q], smy = S[2 * n + q], E = S[3 * n + q];
    const double kinetic = A.eps * 0.5 * (M[q] * M[q] + M[n + q] * M[n + q]) / rho;
    const Cons se = corrected_at(A, S, S + n, i + 1, j, t1), sw = corrected_at(A, S, S + n, i - 1, j, t1);
    const Cons sn = corrected_at(A, S, S + n, i, j + 1, t1), ss = corrected_at(A, S, S + n, i, j - 1, t1);
    const double div = (he * se.mx + (he - hw) * smx - hw * sw.mx) / (2.0 * A.dx) +
                       (hn * sn.my + (hn - hs) * smy - hs * ss.my) / (2.0 * A.dy);
    double r = A.eps * (E - kinetic - dt * div);
    const int ti[9] = {i + 2, i + 1, i, i - 1, i - 2, i, i, i, i};
    const int tj[9] = {j, j, j, j, j, j + 2, j + 1, j - 1, j - 2};
    const double tv[9] = {-cx * he, -cx * (he - hw), eos + cx * (he + hw) + cy * (hn + hs),
                          cx * (he - hw), -cx * hw, -cy * hn, -cy * (hn - hs), cy * (hn - hs), -cy * hs};
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      if (tv[e] == 0.0) continue;
      int ci = ti[e], cj = tj[e];
      if (!inside(A, ci, cj)) r -= tv[e] * boundary_pressure(A, ti[e], tj[e], t1);
    }
    b[q] = r;
    if (!isfinite(r) || !(rho > 0.0)) ++bad;
  }
  if (bad) atomicAdd(&A.counts[1], bad);
}

// Closeout (:568-598) into the caller's layout.
__global__ void __launch_bounds__(DB) direct_close_kernel(DirectArgs A, const double* __restrict__ S,
                                                          const double* __restrict__ M,
                                                          const double* __restrict__ P,
                                                          double* __restrict__ U, long long rs, long long cs)
{
  const long long n = (long long)A.nx * A.ny;
  const double gm1 = A.gamma - 1.0;
  int bad = 0, nonpos = 0;
  for (long long q = (long long)blockIdx.x * DB + threadIdx.x; q < n; q += (long long)gridDim.x * DB) {
    const int i = (int)(q % A.nx), j = (int)(q / A.nx);
    Cons u{S[q], M[q], M[n + q], 0.0};
    const double kinetic = A.eps * 0.5 * (u.mx * u.mx + u.my * u.my) / u.r;
    u.E = P[q] / gm1 + kinetic;
    const Prim w = eps_prim(u, A.gamma, A.eps);
    if (!isfinite(w.r) || !isfinite(w.u) || !isfinite(w.v) || !isfinite(w.p)) ++bad;
    if (!(w.r > 0.0) || !(w.p > 0.0)) ++nonpos;
    const long long k = (long long)j * rs + i;
    U[k] = u.r;
    U[cs + k] = u.mx;
    U[2 * cs + k] = u.my;
    U[3 * cs + k] = u.E;
  }
  if (bad) atomicAdd(&A.counts[1], bad);
  if (nonpos) atomicAdd(&A.counts[2], nonpos);
}

// Material rate 2 |u| / min(dx, dy) of compute_bdltv20_paper_t1_s2_dt
// (:8-41): one partial maximum per workgroup.
__global__ void __launch_bounds__(DB) direct_rate_kernel(DirectArgs A, const double* __restrict__ U,
                                                         long long rs, long long cs, double* __restrict__ part)
{
  __shared__ double red[DB / 64];
  const long long n = (long long)A.nx * A.ny;
  const double min_dx = smin(A.dx, A.dy);
  double rate = 0.0;
  for (long long q = (long long)blockIdx.x * DB + threadIdx.x; q < n; q += (long long)gridDim.x * DB) {
    const int i = (int)(q % A.nx), j = (int)(q / A.nx);
    const long long k = (long long)j * rs + i;
    const Prim w = eps_prim(Cons{U[k], U[cs + k], U[2 * cs + k], U[3 * cs + k]}, A.gamma, A.eps);
    if (isfinite(w.u) && isfinite(w.v)) rate = smax(rate, 2.0 * sqrt(w.u * w.u + w.v * w.v) / min_dx);
  }
  for (int o = 32; o > 0; o >>= 1) rate = smax(rate, __shfl_xor(rate, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rate;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = red[0];
    for (int w = 1; w < DB / 64; ++w) r = smax(r, red[w]);
    part[blockIdx.x] = r;
  }
}

int blocks_for(long long n)
{
  long long b = (n + DB - 1) / DB;
  if (b > 4096) b = 4096;
  return b < 1 ? 1 : (int)b;
}

DirectWork* work(ec_ctx* ctx) { return static_cast<DirectWork*>(ctx->direct); }

DirectArgs args(ec_ctx* ctx)
{
  DirectWork* w = work(ctx);
  DirectArgs A;
  A.nx = ctx->grid.nx;
  A.ny = ctx->grid.ny;
  A.px = w->px;
  A.py = w->py;
  A.x_only = w->x_only;
  A.dx = ctx->grid.dx;
  A.dy = ctx->grid.dy;
  A.gamma = ctx->grid.gamma;
  A.eps = w->eps;
  A.bc = w->bc;
  A.counts = w->counts;
  return A;
}

}  // namespace

void direct_release(ec_ctx* ctx)
{
  DirectWork* w = work(ctx);
  if (!w) return;
  if (w->bc) (void)hipFree(w->bc);
  if (w->ring) (void)hipFree(w->ring);
  if (w->counts) (void)hipFree(w->counts);
  if (w->part) (void)hipFree(w->part);
  if (w->counts_host) (void)hipHostFree(w->counts_host);
  if (w->part_host) (void)hipHostFree(w->part_host);
  delete w;
  ctx->direct = nullptr;
}

}  // namespace ec

using namespace ec;

extern "C" ec_status ec_direct_setup(ec_ctx* ctx, int direct_case, const ec_problem_desc* pd,
                                     double epsilon)
{
  if (!ctx || !pd || !(epsilon > 0.0) || !std::isfinite(epsilon)) return EC_ERR_INVALID_ARGUMENT;
  if (direct_case < EC_DIRECT_TORO_X_EXACT_Y_PERIODIC || direct_case > EC_DIRECT_ADVECTION_BLOB_PERIODIC) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  if (pd->nx != ctx->grid.nx || pd->ny != ctx->grid.ny || pd->gamma != ctx->grid.gamma) {
    ctx->last_error = "ec_direct_setup: problem grid or gamma differs from the context's";
    return EC_ERR_INVALID_ARGUMENT;
  }
  const bool toro = direct_case == EC_DIRECT_TORO_X_EXACT_Y_PERIODIC || direct_case == EC_DIRECT_TORO_XY_EXACT;
  const int want_kind = toro ? EC_PROBLEM_TORO
                             : direct_case == EC_DIRECT_GRESHO_EXACT
                                   ? EC_PROBLEM_GRESHO
                                   : direct_case == EC_DIRECT_ISENTROPIC_VORTEX_PERIODIC
                                         ? EC_PROBLEM_ISENTROPIC_VORTEX
                                         : EC_PROBLEM_ADVECTION_BLOB;
  if (pd->kind != want_kind) {
    ctx->last_error = "ec_direct_setup: the problem does not match the case";
    return EC_ERR_INVALID_ARGUMENT;
  }
  prob::P4 tl{}, tr{};
  if (toro && !prob::toro_state(pd->toro_test, tl, tr)) {
    ctx->last_error = "ec_direct_setup: unknown Toro test id";
    return EC_ERR_INVALID_ARGUMENT;
  }
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  EC_HIP_TRY(ctx, ctx->stream ? hipStreamSynchronize(ctx->stream) : hipDeviceSynchronize());
  direct_release(ctx);
  DirectWork* w = new DirectWork();
  ctx->direct = w;
  w->dcase = direct_case;
  w->eps = epsilon;
  // the periodicity each case requires (core/config_reader.hpp:228-250)
  w->px = (direct_case == EC_DIRECT_ISENTROPIC_VORTEX_PERIODIC ||
           direct_case == EC_DIRECT_ADVECTION_BLOB_PERIODIC) ? 1 : 0;
  w->py = (direct_case == EC_DIRECT_TORO_XY_EXACT || direct_case == EC_DIRECT_GRESHO_EXACT) ? 0 : 1;
  w->x_only = toro ? 1 : 0;
  const long long n = (long long)pd->nx * pd->ny;
  w->nparts = blocks_for(n);
  if (hipMalloc(&w->counts, 4 * sizeof(int)) != hipSuccess ||
      hipHostMalloc(&w->counts_host, 4 * sizeof(int)) != hipSuccess ||
      hipMalloc(&w->part, w->nparts * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&w->part_host, w->nparts * sizeof(double)) != hipSuccess ||
      hipMalloc(&w->bc, sizeof(DirectBc)) != hipSuccess) {
    direct_release(ctx);
    return EC_ERR_OUT_OF_MEMORY;
  }
  DirectBc d;
  std::memset(&d, 0, sizeof(d));
  d.eps = epsilon;
  d.xlo = pd->prob_lo[0];
  d.xhi = pd->prob_hi[0];
  d.ylo = pd->prob_lo[1];
  d.yhi = pd->prob_hi[1];
  d.e.nx = pd->nx;
  d.e.ny = pd->ny;
  d.e.xlo = pd->prob_lo[0];
  d.e.ylo = pd->prob_lo[1];
  d.e.dx = (pd->prob_hi[0] - pd->prob_lo[0]) / pd->nx;
  d.e.dy = (pd->prob_hi[1] - pd->prob_lo[1]) / pd->ny;
  d.e.gamma = pd->gamma;
  d.e.kind = EXACT_NONE;
  if (toro) {
    d.e.kind = EXACT_TORO;
    const double l[4] = {tl.r, tl.u, tl.v, tl.p}, r[4] = {tr.r, tr.u, tr.v, tr.p};
    for (int c = 0; c < 4; ++c) {
      d.e.ql[c] = l[c];
      d.e.qr[c] = r[c];
    }
    d.e.x0 = pd->riemann_interface_x;
    toro_constants(d.e, tl, tr, pd->gamma);
  } else if (direct_case == EC_DIRECT_GRESHO_EXACT) {
    // the Gresho state at each outside index's boundary point, epsilon-scaled
    d.e.kind = EXACT_RING;
    std::vector<double> ring(4 * ring_cells(pd->nx, pd->ny));
    for (int j = -2; j < pd->ny + 2; ++j) {
      for (int i = -2; i < pd->nx + 2; ++i) {
        if (i >= 0 && i < pd->nx && j >= 0 && j < pd->ny) continue;
        const int ic = i < 0 ? 0 : (i > pd->nx - 1 ? pd->nx - 1 : i);
        const int jc = j < 0 ? 0 : (j > pd->ny - 1 ? pd->ny - 1 : j);
        double x = d.e.xlo + ((double)ic + 0.5) * d.e.dx;
        double y = d.e.ylo + ((double)jc + 0.5) * d.e.dy;
        if (i < 0) x = d.xlo;
        else if (i >= pd->nx) x = d.xhi;
        if (j < 0) y = d.ylo;
        else if (j >= pd->ny) y = d.yhi;
        const prob::P4 q = prob::gresho_state(x, y, *pd);
        const Cons u = eps_cons(Prim{q.r, q.u, q.v, q.p}, pd->gamma, epsilon);
        double* dst = &ring[4 * ring_index(i, j, pd->nx, pd->ny)];
        dst[0] = u.r;
        dst[1] = u.mx;
        dst[2] = u.my;
        dst[3] = u.E;
      }
    }
    if (hipMalloc(&w->ring, sizeof(double) * ring.size()) != hipSuccess) {
      direct_release(ctx);
      return EC_ERR_OUT_OF_MEMORY;
    }
    EC_HIP_TRY(ctx, hipMemcpy(w->ring, ring.data(), sizeof(double) * ring.size(), hipMemcpyHostToDevice));
    d.e.ring = w->ring;
  }
  EC_HIP_TRY(ctx, hipMemcpy(w->bc, &d, sizeof(DirectBc), hipMemcpyHostToDevice));
  return EC_OK;
}

extern "C" ec_status ec_direct_dt(ec_ctx* ctx, const double* U, const ec_layout* lay, double imex_cfl,
                                  double fixed_dt, double time, double stop_time, double* dt)
{
  if (!ctx || !U || !lay || !dt || !work(ctx)) return EC_ERR_INVALID_ARGUMENT;
  if (lay->row_stride < ctx->grid.nx || lay->comp_stride < lay->row_stride * ctx->grid.ny) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  if (fixed_dt > 0.0) {
    *dt = smin(fixed_dt, stop_time - time);
    return EC_OK;
  }
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  DirectWork* w = work(ctx);
  const DirectArgs A = args(ctx);
  hipLaunchKernelGGL(direct_rate_kernel, dim3(w->nparts), dim3(DB), 0, ctx->stream, A, U, lay->row_stride,
                     lay->comp_stride, w->part);
  EC_HIP_TRY(ctx, hipGetLastError());
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->part_host, w->part, w->nparts * sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->stream));
  EC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  double rate = 0.0;
  for (int b = 0; b < w->nparts; ++b) rate = smax(rate, w->part_host[b]);
  const double remaining = stop_time - time;
  *dt = !(rate > 0.0) ? remaining : smin(imex_cfl / rate, remaining);
  return EC_OK;
}

extern "C" ec_status ec_direct_step(ec_ctx* ctx, double* U, const ec_layout* lay, double time, double dt,
                                    const ec_imex_opts* opts, ec_direct_diag* diag, double* p_out)
{
  if (!ctx || !U || !lay || !opts || !diag || !work(ctx)) return EC_ERR_INVALID_ARGUMENT;
  if (lay->row_stride < ctx->grid.nx || lay->comp_stride < lay->row_stride * ctx->grid.ny) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  if (opts->picard_iterations < 1 || opts->restart < 1 || opts->restart > 64 || !(opts->tol > 0.0) ||
      opts->max_iter < 0 ||
      (opts->precision != EC_IMEX_PRECISION_FP64 && opts->precision != EC_IMEX_PRECISION_MIXED &&
       opts->precision != EC_IMEX_PRECISION_FP64_FLEXIBLE) ||
      !(dt > 0.0) || !std::isfinite(dt) || !std::isfinite(time)) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  ImexBuffers B;
  ec_status st = imex_buffers(ctx, opts->restart, &B);
  if (st != EC_OK) return st;
  DirectWork* w = work(ctx);
  const DirectArgs A = args(ctx);
  const long long n = (long long)A.nx * A.ny;
  const long long nf = (long long)(A.nx + 1) * A.ny + (long long)A.nx * (A.ny + 1);
  const int nb = blocks_for(n), nbf = blocks_for(nf);
  const bool mixed = opts->precision == EC_IMEX_PRECISION_MIXED;
  const double t1 = time + dt;
  std::memset(diag, 0, sizeof *diag);
  EC_HIP_TRY(ctx, hipMemsetAsync(w->counts, 0, 4 * sizeof(int), ctx->stream));
  hipLaunchKernelGGL(direct_star_kernel, dim3(nb), dim3(DB), 0, ctx->stream, A, U, lay->row_stride,
                     lay->comp_stride, time, dt, B.star, B.p);
  // the operator's sides: wrapped columns, or Dirichlet columns held in b
  const int bc[4] = {A.px ? BC_PERIODIC : BC_FIXED, A.px ? BC_PERIODIC : BC_FIXED,
                     A.py ? BC_PERIODIC : BC_FIXED, A.py ? BC_PERIODIC : BC_FIXED};
  const double cx = 0.25 * dt * dt / (A.dx * A.dx);
  const double cy = A.x_only ? 0.0 : 0.25 * dt * dt / (A.dy * A.dy);
  const double diag_coef = A.eps / (A.gamma - 1.0);
  for (int it = 0; it < opts->picard_iterations; ++it) {
    hipLaunchKernelGGL(direct_correct_kernel, dim3(nb), dim3(DB), 0, ctx->stream, A, B.star, B.p,
                       it > 0 ? 1 : 0, dt, t1, B.mom);
    hipLaunchKernelGGL(direct_faces_kernel, dim3(nbf), dim3(DB), 0, ctx->stream, A, B.star, B.mom, B.p, t1,
                       B.hx, B.hy, mixed ? B.hx32 : nullptr, mixed ? B.hy32 : nullptr);
    hipLaunchKernelGGL(direct_rhs_kernel, dim3(nb), dim3(DB), 0, ctx->stream, A, B.star, B.mom, B.hx, B.hy,
                       dt, t1, B.b);
    EC_HIP_TRY(ctx, hipGetLastError());
    PressureSolve ps;
    if ((st = imex_solve_pressure(ctx, bc, cx, cy, diag_coef, *opts, &ps)) != EC_OK) return st;
    const bool ok = !ps.bad && std::isfinite(ps.rinf) && (ps.rinf <= opts->tol || ps.rel <= opts->tol);
    if (!ok || ps.breakdown) ++diag->solver_failures;
    diag->iterations += ps.iterations;
    if (ps.iterations > diag->iterations_max) diag->iterations_max = ps.iterations;
    diag->residual_linf_max = smax(diag->residual_linf_max, ps.rinf);
    diag->relative_residual_linf_max = smax(diag->relative_residual_linf_max, ps.rel);
    // the reference continues with the solver's result either way
    EC_HIP_TRY(ctx, hipMemcpyAsync(B.p, B.x, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  }
  hipLaunchKernelGGL(direct_correct_kernel, dim3(nb), dim3(DB), 0, ctx->stream, A, B.star, B.p, 1, dt, t1,
                     B.mom);
  hipLaunchKernelGGL(direct_faces_kernel, dim3(nbf), dim3(DB), 0, ctx->stream, A, B.star, B.mom, B.p, t1,
                     B.hx, B.hy, nullptr, nullptr);
  hipLaunchKernelGGL(direct_close_kernel, dim3(nb), dim3(DB), 0, ctx->stream, A, B.star, B.mom, B.p, U,
                     lay->row_stride, lay->comp_stride);
  if (p_out) {
    EC_HIP_TRY(ctx, hipMemcpyAsync(p_out, B.p, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  }
  EC_HIP_TRY(ctx, hipGetLastError());
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->counts_host, w->counts, 4 * sizeof(int), hipMemcpyDeviceToHost,
                                 ctx->stream));
  EC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  diag->guard_count = w->counts_host[0];
  diag->nonfinite_count = w->counts_host[1];
  diag->nonpositive_count = w->counts_host[2];
  return EC_OK;
}
