// ec_imex.hip -- IMEX T1/S2 BDLTV20 source-map Picard step on the device.
//
// The reference (imex/imex_source_map_step.hpp:300-616) copies the star state
// to the host, assembles 11-term pressure rows into an Eigen sparse matrix
// (building label strings per term) and runs an unrestarted host GMRES.  Here
// every stage stays in HBM:
//   star       explicit Toro-Vazquez predictor, conserved minmod + material
//              LF flux (imex_source_map_helpers.hpp:2-68)
//   faces      momentum-weighted face enthalpy h on every x- and y-face
//              (imex_source_map_helpers.hpp:359-402), stored once per face
//   rhs        row right-hand side (imex_source_map_helpers.hpp:545-566)
//   operator   matrix-free, differences-first: for the x part
//                p_i/(g-1) - cx { h_e[(p_{i+2}-p_i)+(p_{i+1}-p_{i-1})]
//                                - h_w[(p_i-p_{i-2})+(p_{i+1}-p_{i-1})] }
//              which is the reference row (:533-543) re-associated so that a
//              constant pressure cancels exactly (it removes the FP64
//              cancellation floor of the assembled row at low Mach)
//   GMRES(m)   FP64 form: restarted, CGS2 orthogonalisation with fused
//              multi-dots, Givens least squares and early exit on device (one
//              host synchronisation per restart cycle).  Mixed form: FP64
//              flexible outer GMRES whose preconditioner is an FP32 inner
//              GMRES(m) with one reduction per Arnoldi step (FGMRES-based
//              iterative refinement, one synchronisation per outer cycle)
//   refresh    m = m* - dt grad_c p (imex_source_map_helpers.hpp:249-268)
//   closeout   E = E* - dt div(h m) and the new state
//              (imex_source_map_step.hpp:75-276)
#include <cmath>
#include <cstring>
#include <vector>

#include <cstddef>

#include "ec_context.hpp"
#include "ec_wave.hpp"
#include "ec_ghost.hpp"
#include "ec_state.hpp"
#include "ec_tile.hpp"
#include "ec_imex_api.hpp"

namespace ec {

constexpr int RB = 256;        // threads per reduction workgroup
constexpr int MAXK = 64;       // largest GMRES restart length

// ------------------------------------------------------------ device control
struct GmCtl {
  int done, kk, breakdown, bad;
  int mlim;                    // iterations allowed in this cycle
  int inner_its;               // flexible outer solver: FP32 iterations of this cycle
  int inner_mlim;              // flexible outer solver: the inner restart length
  int inner_budget;            // flexible outer solver: inner iterations this cycle may take
  double thr, est, beta, inv_norm;
  double rinf, r2;             // true residual norms
  double inner_thr;            // flexible outer solver: inner relative tolerance
  double dots1[MAXK], dots2[MAXK], nrm2;
  unsigned int counter;
  int floor_count, nonfinite, guard_count, mom_bad;
  double coupled_res, mat_max, aco_max;
  // closeout evaluation (imex/imex_source_map_step.hpp:75-258): the finite
  // star density's minimum (negated), the record of CloseCol, the eq78
  // divergence maximum and the count of cells it covers
  double star_rho_neg_max, star_pad;
  double close[8];
  double eq78, eq78_count;
};

template <typename T>
struct OpArgs;

struct ImexWork {
  int nx = 0, ny = 0, m = 0;
  double* star = nullptr;   // 4n
  double* p = nullptr;      // n  (current Picard pressure)
  double* x = nullptr;      // n  (GMRES iterate, FP64)
  double* mom = nullptr;    // 2n lagged momentum
  double* hx = nullptr;     // (nx+1)*ny face enthalpy
  double* hy = nullptr;     // nx*(ny+1)
  double* fx = nullptr;     // face energy flux (closeout)
  double* fy = nullptr;
  float* hx32 = nullptr;
  float* hy32 = nullptr;
  double* b = nullptr;      // n
  double* r = nullptr;      // n
  double* V = nullptr;      // (m+1) n
  float* V32 = nullptr;     // (m+1) n
  double* w = nullptr;
  float* w32 = nullptr;
  double* R = nullptr;      // (m+1) x m Hessenberg / triangular factor
  double* g = nullptr;      // 2(m+1): running entries, then the rotated ones
  double* cs = nullptr;
  double* sn = nullptr;
  double* y = nullptr;
  double* partials = nullptr;  // reduction partials
  double* npart = nullptr;     // |w|^2 partials of the Arnoldi norm pass
  int nparts = 0;
  // flexible outer solver of the mixed form (FP64), allocated on first use
  double* Vo = nullptr;      // (MO+1) n outer basis
  double* Zo = nullptr;      // MO n preconditioned directions
  double* Ro = nullptr;      // (MO+1) x MO
  double* go = nullptr;      // 2(MO+1)
  double* cso = nullptr;
  double* sno = nullptr;
  double* yo = nullptr;
  GmCtl* ctlo = nullptr;
  GmCtl* ctlo_host = nullptr;
  std::vector<hipGraphExec_t> exec_outer;    // MO step graphs, then the closing graph (FP32 inner)
  std::vector<hipGraphExec_t> exec_outer64;  // the same with the FP64 inner solver
  // one-reduction cycle (FP32 inner solver)
  double* opart = nullptr;   // tile partials, 2 x otiles x KP (ping-pong by Arnoldi step)
  double* otot = nullptr;    // their totals (KP) on grids of many tiles
  double* hraw = nullptr;    // unrotated Hessenberg columns, (m+1) x m
  double* inv_nu = nullptr;  // 1 / |u_q|, m+1
  int otiles = 0;
  GmCtl* ctl = nullptr;
  GmCtl* ctl_host = nullptr;
  // Restart cycles run as hipGraphs (one per precision) on a private stream;
  // the operator arguments they read live in device memory.
  OpArgs<double>* dA64 = nullptr;
  OpArgs<float>* dA32 = nullptr;
  OpArgs<double>* hA64 = nullptr;   // pinned staging of the two blocks
  OpArgs<float>* hA32 = nullptr;
  hipStream_t gs = nullptr;
  hipEvent_t ev = nullptr;
  hipGraphExec_t exec64 = nullptr;
};

void imex_release(ec_ctx* ctx)
{
  ImexWork* w = static_cast<ImexWork*>(ctx->imex);
  if (!w) return;
  void* ptrs[] = {w->star, w->p,   w->x,    w->mom, w->hx,       w->hy,      w->fx, w->fy,
                  w->hx32, w->hy32, w->b,   w->r,   w->V,        w->V32,     w->w,  w->w32,
                  w->R,    w->g,    w->cs,   w->sn,  w->y,        w->partials, w->ctl,
                  w->npart, w->opart,  w->otot,  w->hraw,   w->inv_nu,
                  w->Vo,   w->Zo,   w->Ro,   w->go,  w->cso,      w->sno,     w->yo, w->ctlo,
                  w->dA64, w->dA32};
  for (void* p : ptrs) {
    if (p) (void)hipFree(p);
  }
  if (w->ctl_host) (void)hipHostFree(w->ctl_host);
  if (w->ctlo_host) (void)hipHostFree(w->ctlo_host);
  for (hipGraphExec_t e : w->exec_outer) {
    if (e) (void)hipGraphExecDestroy(e);
  }
  for (hipGraphExec_t e : w->exec_outer64) {
    if (e) (void)hipGraphExecDestroy(e);
  }
  if (w->hA64) (void)hipHostFree(w->hA64);
  if (w->hA32) (void)hipHostFree(w->hA32);
  if (w->exec64) (void)hipGraphExecDestroy(w->exec64);
  if (w->ev) (void)hipEventDestroy(w->ev);
  if (w->gs) (void)hipStreamDestroy(w->gs);
  delete w;
  ctx->imex = nullptr;
}

// -------------------------------------------------------- boundary mapping
// The IMEX rules per side (imex_source_map_helpers.hpp:79-115, 169-245,
// 293-320): states and momenta are read wrap/clamp, with a FIXED side giving
// the inflow state and a REFLECT side the first row with the normal momentum
// negated; pressures are read wrap or even-reflected, with a FIXED side giving
// its Dirichlet value.
struct ImexGrid {
  int nx, ny;
  int bc[4];
  double fixed[4][4];
  double fixed_p[4];
  double gamma, dx, dy, inv_dx, inv_dy, dt;
};

__device__ __forceinline__ int fixed_side(const ImexGrid& G, int i, int j)
{
  if (i < 0 && G.bc[0] == BC_FIXED) return 0;
  if (i >= G.nx && G.bc[1] == BC_FIXED) return 1;
  if (j < 0 && G.bc[2] == BC_FIXED) return 2;
  if (j >= G.ny && G.bc[3] == BC_FIXED) return 3;
  return -1;
}

__device__ __forceinline__ int wrapi(int i, int n)
{
  const int m = i % n;
  return m < 0 ? m + n : m;
}

__device__ __forceinline__ int clampi(int i, int n) { return i < 0 ? 0 : (i > n - 1 ? n - 1 : i); }

__device__ __forceinline__ long long state_index(const ImexGrid& G, int i, int j, bool& nmx, bool& nmy)
{
  nmx = nmy = false;
  int ii = G.bc[0] == BC_PERIODIC ? wrapi(i, G.nx) : clampi(i, G.nx);
  int jj = G.bc[2] == BC_PERIODIC ? wrapi(j, G.ny) : clampi(j, G.ny);
  if (j < 0 && G.bc[2] == BC_REFLECT) { jj = 0; nmy = true; }
  if (j >= G.ny && G.bc[3] == BC_REFLECT) { jj = G.ny - 1; nmy = true; }
  if (i < 0 && G.bc[0] == BC_REFLECT) { ii = 0; nmx = true; }
  if (i >= G.nx && G.bc[1] == BC_REFLECT) { ii = G.nx - 1; nmx = true; }
  return (long long)jj * G.nx + ii;
}

__device__ __forceinline__ Cons star_at(const ImexGrid& G, const double* S, int i, int j)
{
  const int fs = fixed_side(G, i, j);
  if (fs >= 0) return Cons{G.fixed[fs][0], G.fixed[fs][1], G.fixed[fs][2], G.fixed[fs][3]};
  bool nmx, nmy;
  const long long k = state_index(G, i, j, nmx, nmy);
  const long long n = (long long)G.nx * G.ny;
  Cons u{S[k], S[n + k], S[2 * n + k], S[3 * n + k]};
  if (nmx) u.mx = -u.mx;
  if (nmy) u.my = -u.my;
  return u;
}

__device__ __forceinline__ void mom_at(const ImexGrid& G, const double* M, int i, int j, double& mx,
                                       double& my)
{
  const int fs = fixed_side(G, i, j);
  if (fs >= 0) {
    mx = G.fixed[fs][1];
    my = G.fixed[fs][2];
    return;
  }
  bool nmx, nmy;
  const long long k = state_index(G, i, j, nmx, nmy);
  const long long n = (long long)G.nx * G.ny;
  mx = nmx ? -M[k] : M[k];
  my = nmy ? -M[n + k] : M[n + k];
}

// Even reflection (source_map_index, imex_source_map_helpers.hpp:169-188).
__device__ __forceinline__ int pmap1(int i, int n, bool periodic)
{
  if (periodic) return wrapi(i, n);
  if (i < 0) {
    const int m = -i - 1;
    return m < n - 1 ? m : n - 1;
  }
  if (i >= n) {
    const int m = n - (i - n + 1);
    return m > 0 ? m : 0;
  }
  return i;
}

__device__ __forceinline__ long long pcol(const ImexGrid& G, int i, int j)
{
  return (long long)pmap1(j, G.ny, G.bc[2] == BC_PERIODIC) * G.nx +
         pmap1(i, G.nx, G.bc[0] == BC_PERIODIC);
}

__device__ __forceinline__ double p_at(const ImexGrid& G, const double* P, int i, int j)
{
  const int fs = fixed_side(G, i, j);
  if (fs >= 0) return G.fixed_p[fs];
  return P[pcol(G, i, j)];
}

// Operator read of x at (i,j): mapped column, or 0 at a FIXED side (the
// Dirichlet part lives in the right-hand side).
template <typename T>
__device__ __forceinline__ T x_at(const ImexGrid& G, const T* X, int i, int j)
{
  if (fixed_side(G, i, j) >= 0) return T(0);
  return X[pcol(G, i, j)];
}

// ------------------------------------------------------------ reductions
template <int K>
struct DSum {
  double v[K];
};

// Sum-reduce K doubles per workgroup into partials[block*K + k] (plain
// stores; a finalize kernel launched afterwards sums them, so no
// device-scope fence is ever needed).
template <int K>
__device__ void reduce_sum_store(DSum<K> s, double* partials, int kactive)
{
  __shared__ double sh[RB / 64][K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k < kactive) {
      const double v = wave_sum_lane63(s.v[k]);
      if (lane == 63) sh[wave][k] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < kactive) {
    double v = 0.0;
    for (int w = 0; w < RB / 64; ++w) v += sh[w][threadIdx.x];
    partials[(long long)blockIdx.x * K + threadIdx.x] = v;
  }
}

// Workgroup sums of the k products prod(q) (q < k) of this thread's row,
// stored as partials[block*MAXK + q].  One row per thread, so nothing is
// carried across rows and the accumulators do not occupy registers.
template <typename F>
__device__ __forceinline__ void rowdots_store(const F& prod, int k, double* partials)
{
  __shared__ double sh[RB / 64][MAXK];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // Eight products (and their basis loads) in flight at a time; each summed
  // over the wavefront with DPP moves (total in lane 63).
  for (int q0 = 0; q0 < k; q0 += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = q0 + u < k ? prod(q0 + u) : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = wave_sum_lane63(v[u]);
    if (lane == 63) {
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (q0 + u < k) sh[wave][q0 + u] = v[u];
    }
  }
  __syncthreads();
  if (threadIdx.x < k) {
    double v = sh[0][threadIdx.x];
    for (int w = 1; w < RB / 64; ++w) v += sh[w][threadIdx.x];
    partials[(long long)blockIdx.x * MAXK + threadIdx.x] = v;
  }
}

// Per-workgroup (max, sum) of one value each, stored as partials[2b], [2b+1].
__device__ void reduce_maxsum_store(double mx, double sm, double* partials)
{
  __shared__ double shm[RB / 64], shs[RB / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mx = smax(mx, __shfl_xor(mx, off));
    sm += __shfl_xor(sm, off);
  }
  if (lane == 0) {
    shm[wave] = mx;
    shs[wave] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = shm[0], b = shs[0];
    for (int w = 1; w < RB / 64; ++w) {
      a = smax(a, shm[w]);
      b += shs[w];
    }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = b;
  }
}

// combine = 0: (max, sum); 1: (max, max).  One workgroup; the sum runs in a
// fixed order (run-to-run reproducible).
__global__ void __launch_bounds__(RB) finalize_pair_kernel(const double* partials, int nparts,
                                                           int combine, double* out_a, double* out_b)
{
  __shared__ double sha[RB / 64], shb[RB / 64];
  double a = -INFINITY, b = combine ? -INFINITY : 0.0;
  for (int q = threadIdx.x; q < nparts; q += RB) {
    a = smax(a, partials[2 * q]);
    b = combine ? smax(b, partials[2 * q + 1]) : b + partials[2 * q + 1];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a = smax(a, __shfl_xor(a, off));
    const double ob = __shfl_xor(b, off);
    b = combine ? smax(b, ob) : b + ob;
  }
  if (lane == 0) {
    sha[wave] = a;
    shb[wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < RB / 64; ++w) {
      a = smax(a, sha[w]);
      b = combine ? smax(b, shb[w]) : b + shb[w];
    }
    *out_a = a;
    *out_b = b;
  }
}

// The closeout's record per workgroup: maxima and minima over the finite
// cells (evaluate_source_map_final, imex/imex_source_map_step.hpp:151-166).
enum CloseCol {
  CL_RES = 0,        // max |p/(g-1) - E + K|
  CL_RHO_MIN,        // min rho
  CL_P_MIN,          // min solved p
  CL_P_MAX,          // max solved p
  CL_EOS_MIN,        // min (g-1)(E - K)
  CL_INTERNAL_MIN,   // min E - K
  CL_MOM_CORR,       // max |m - m*|
  CL_E_DELTA,        // max |E - E*|
  CLOSE_K
};
__host__ __device__ constexpr bool close_is_min(int c)
{
  return c == CL_RHO_MIN || c == CL_P_MIN || c == CL_EOS_MIN || c == CL_INTERNAL_MIN;
}

__device__ __forceinline__ void close_record_init(double (&v)[CLOSE_K])
{
#pragma unroll
  for (int c = 0; c < CLOSE_K; ++c) v[c] = close_is_min(c) ? INFINITY : -INFINITY;
  v[CL_RES] = 0.0;
  v[CL_MOM_CORR] = 0.0;
  v[CL_E_DELTA] = 0.0;
}

__device__ __forceinline__ double close_fold(int c, double a, double b)
{
  return close_is_min(c) ? smin(a, b) : smax(a, b);
}

// Folds the record over the workgroup (wave shuffles, then the waves in
// order) and stores it as partials[CLOSE_K * blockIdx.x + c].
__device__ void close_record_store(double (&v)[CLOSE_K], double* partials)
{
  __shared__ double sh[RB / 64][CLOSE_K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < CLOSE_K; ++c)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[c] = close_fold(c, v[c], __shfl_xor(v[c], off));
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < CLOSE_K; ++c) sh[wave][c] = v[c];
  __syncthreads();
  if (threadIdx.x < CLOSE_K) {
    const int c = threadIdx.x;
    double a = sh[0][c];
    for (int w = 1; w < RB / 64; ++w) a = close_fold(c, a, sh[w][c]);
    partials[(long long)CLOSE_K * blockIdx.x + c] = a;
  }
}

// One workgroup folds the closeout records of nparts workgroups into out.
__global__ void __launch_bounds__(RB) finalize_close_kernel(const double* partials, int nparts, double* out)
{
  double v[CLOSE_K];
  close_record_init(v);
  for (int q = threadIdx.x; q < nparts; q += RB)
#pragma unroll
    for (int c = 0; c < CLOSE_K; ++c) v[c] = close_fold(c, v[c], partials[(long long)CLOSE_K * q + c]);
  __shared__ double sh[RB / 64][CLOSE_K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < CLOSE_K; ++c)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v[c] = close_fold(c, v[c], __shfl_xor(v[c], off));
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < CLOSE_K; ++c) sh[wave][c] = v[c];
  __syncthreads();
  if (threadIdx.x < CLOSE_K) {
    const int c = threadIdx.x;
    double a = sh[0][c];
    for (int w = 1; w < RB / 64; ++w) a = close_fold(c, a, sh[w][c]);
    out[c] = a;
  }
}

// ------------------------------------------------------------ star update
// reconstruction_fluxes.hpp:26-52 and :462-515; the explicit ghost rule of
// the state is the same as the explicit step's.
__device__ __forceinline__ Prim cons_minmod_state(const double* U, long long rs, long long cs,
                                                  const GridParams& g, int i, int j, int axis,
                                                  double half_side)
{
  const Cons c = load_cell(U, rs, cs, g, i, j);
  const int di = axis == 0 ? 1 : 0, dj = axis == 0 ? 0 : 1;
  const Cons m = load_cell(U, rs, cs, g, i - di, j - dj);
  const Cons p = load_cell(U, rs, cs, g, i + di, j + dj);
  const Cons r{c.r + half_side * minmod(c.r - m.r, p.r - c.r),
               c.mx + half_side * minmod(c.mx - m.mx, p.mx - c.mx),
               c.my + half_side * minmod(c.my - m.my, p.my - c.my),
               c.E + half_side * minmod(c.E - m.E, p.E - c.E)};
  return to_prim(r, g.gamma);
}

__device__ __forceinline__ Cons lf_flux(const Prim& L, const Prim& R, int axis, double gamma,
                                        int acoustic)
{
  const Cons ul = to_cons(L, gamma), ur = to_cons(R, gamma);
  auto pflux = [&](const Prim& q) {
    const double mx = q.r * q.u, my = q.r * q.v;
    const double kin = 0.5 * q.r * (q.u * q.u + q.v * q.v);
    const double un = axis == 0 ? q.u : q.v;
    return Cons{axis == 0 ? mx : my, mx * un, my * un, kin * un};
  };
  const Cons fl = pflux(L), fr = pflux(R);
  const double unl = normal_vel(L, axis), unr = normal_vel(R, axis);
  double s = smax(fabs(unl), fabs(unr));
  if (acoustic) s = smax(fabs(unl) + sound_speed(L, gamma), fabs(unr) + sound_speed(R, gamma));
  const double hs = 0.5 * s;
  return Cons{0.5 * (fl.r + fr.r) - hs * (ur.r - ul.r), 0.5 * (fl.mx + fr.mx) - hs * (ur.mx - ul.mx),
              0.5 * (fl.my + fr.my) - hs * (ur.my - ul.my), 0.5 * (fl.E + fr.E) - hs * (ur.E - ul.E)};
}

__global__ void __launch_bounds__(256) imex_star_kernel(const double* __restrict__ U, long long rs, long long cs, GridParams g,
                                 double dt, int acoustic, double* __restrict__ S)
{
  const long long n = (long long)g.nx * g.ny;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % g.nx), j = (int)(t / g.nx);
    Cons f[4];
    f[0] = lf_flux(cons_minmod_state(U, rs, cs, g, i - 1, j, 0, 0.5),
                   cons_minmod_state(U, rs, cs, g, i, j, 0, -0.5), 0, g.gamma, acoustic);
    f[1] = lf_flux(cons_minmod_state(U, rs, cs, g, i, j, 0, 0.5),
                   cons_minmod_state(U, rs, cs, g, i + 1, j, 0, -0.5), 0, g.gamma, acoustic);
    f[2] = lf_flux(cons_minmod_state(U, rs, cs, g, i, j - 1, 1, 0.5),
                   cons_minmod_state(U, rs, cs, g, i, j, 1, -0.5), 1, g.gamma, acoustic);
    f[3] = lf_flux(cons_minmod_state(U, rs, cs, g, i, j, 1, 0.5),
                   cons_minmod_state(U, rs, cs, g, i, j + 1, 1, -0.5), 1, g.gamma, acoustic);
    const long long k = (long long)j * rs + i;
    const double o[4] = {U[k], U[cs + k], U[2 * cs + k], U[3 * cs + k]};
    const double a[4] = {f[0].r, f[0].mx, f[0].my, f[0].E};
    const double b[4] = {f[1].r, f[1].mx, f[1].my, f[1].E};
    const double c[4] = {f[2].r, f[2].mx, f[2].my, f[2].E};
    const double d[4] = {f[3].r, f[3].mx, f[3].my, f[3].E};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      S[q * n + t] = o[q] - dt * ((b[q] - a[q]) * g.inv_dx + (d[q] - c[q]) * g.inv_dy);
    }
  }
}

// Initial pressure from the old state, lagged momentum from the star state,
// density floor count (imex_source_map_step.hpp:418-451).
__global__ void __launch_bounds__(256) imex_init_kernel(const double* __restrict__ U, long long rs, long long cs, double gamma,
                                 int nx, int ny, const double* __restrict__ S, double* __restrict__ P,
                                 double* __restrict__ M, double* partials, GmCtl* ctl)
{
  const long long n = (long long)nx * ny;
  double neg = -INFINITY;  // -min of the finite star densities
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % nx), j = (int)(t / nx);
    const long long k = (long long)j * rs + i;
    P[t] = to_prim(Cons{U[k], U[cs + k], U[2 * cs + k], U[3 * cs + k]}, gamma).p;
    M[t] = S[n + t];
    M[n + t] = S[2 * n + t];
    const double rho = S[t];
    if (!isfinite(rho) || rho <= 1.0e-14) atomicAdd(&ctl->floor_count, 1);
    if (isfinite(rho)) neg = smax(neg, -rho);
  }
  reduce_maxsum_store(neg, 0.0, partials);
}

// Face enthalpy (and, for the closeout, energy flux) on every face
// (imex_source_map_helpers.hpp:359-402).  x-face f of row j lies between
// cells (f-1,j) and (f,j).
__device__ __forceinline__ void face_h(const ImexGrid& G, const double* S, const double* M,
                                       const double* P, int il, int jl, int ir, int jr, int axis,
                                       double& hh, double& flux, int& guard)
{
  const double gm1 = G.gamma - 1.0;
  const Cons L = star_at(G, S, il, jl), R = star_at(G, S, ir, jr);
  double mlx, mly, mrx, mry;
  mom_at(G, M, il, jl, mlx, mly);
  mom_at(G, M, ir, jr, mrx, mry);
  const double pl = p_at(G, P, il, jl), pr = p_at(G, P, ir, jr);
  const double hl = G.gamma * pl / (gm1 * L.r);
  const double hr = G.gamma * pr / (gm1 * R.r);
  const double ql = axis == 0 ? mlx : mly;
  const double qr = axis == 0 ? mrx : mry;
  const double denom = ql + qr;
  const double scale = smax(smax(fabs(ql), fabs(qr)), 1.0);
  guard = fabs(denom) <= 1.0e-14 * scale ? 1 : 0;
  hh = guard ? 0.5 * (hl + hr) : (hl * ql + hr * qr) / denom;
  flux = hh * (0.5 * (ql + qr));
}

__global__ void __launch_bounds__(256) imex_faces_kernel(ImexGrid G, const double* __restrict__ S, const double* __restrict__ M,
                                  const double* __restrict__ P, double* __restrict__ hx,
                                  double* __restrict__ hy, float* __restrict__ hx32,
                                  float* __restrict__ hy32, double* __restrict__ fx,
                                  double* __restrict__ fy, GmCtl* ctl)
{
  const long long nxf = (long long)(G.nx + 1) * G.ny;
  const long long nyf = (long long)G.nx * (G.ny + 1);
  int guards = 0, bad = 0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nxf + nyf;
       t += (long long)gridDim.x * blockDim.x) {
    double h, f;
    int gd;
    if (t < nxf) {
      const int fi = (int)(t % (G.nx + 1)), j = (int)(t / (G.nx + 1));
      face_h(G, S, M, P, fi - 1, j, fi, j, 0, h, f, gd);
      hx[t] = h;
      if (hx32) hx32[t] = (float)h;
      if (fx) fx[t] = f;
    } else {
      const long long s = t - nxf;
      const int i = (int)(s % G.nx), fj = (int)(s / G.nx);
      face_h(G, S, M, P, i, fj - 1, i, fj, 1, h, f, gd);
      hy[s] = h;
      if (hy32) hy32[s] = (float)h;
      if (fy) fy[s] = f;
    }
    guards += gd;
    if (!isfinite(h) || !isfinite(f)) ++bad;
  }
  if (guards) atomicAdd(&ctl->guard_count, guards);
  if (bad) atomicAdd(&ctl->nonfinite, bad);
}

// Right-hand side of each pressure row (imex_source_map_helpers.hpp:545-566),
// with the Dirichlet parts of FIXED-side terms added first as the reference
// does while it builds the row (:488-507).
__global__ void __launch_bounds__(256) imex_rhs_kernel(ImexGrid G, const double* __restrict__ S, const double* __restrict__ M,
                                const double* __restrict__ hx, const double* __restrict__ hy,
                                double* __restrict__ b, GmCtl* ctl)
{
  const long long n = (long long)G.nx * G.ny;
  const double cx = 0.25 * G.dt * G.dt * G.inv_dx * G.inv_dx;
  const double cy = 0.25 * G.dt * G.dt * G.inv_dy * G.inv_dy;
  int bad = 0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % G.nx), j = (int)(t / G.nx);
    const double he = hx[(long long)j * (G.nx + 1) + i + 1], hw = hx[(long long)j * (G.nx + 1) + i];
    const double hn = hy[(long long)(j + 1) * G.nx + i], hs = hy[(long long)j * G.nx + i];
    double r = 0.0;
    const int ti[10] = {i + 2, i + 1, i - 1, i - 2, i, i, i, i, i, i};
    const int tj[10] = {j, j, j, j, j + 2, j + 1, j - 1, j - 2, j, j};
    const double tv[8] = {-cx * he, -cx * (he - hw), cx * (he - hw), -cx * hw,
                          -cy * hn, -cy * (hn - hs), cy * (hn - hs), -cy * hs};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int fs = fixed_side(G, ti[q], tj[q]);
      if (fs >= 0) r += -tv[q] * G.fixed_p[fs];
    }
    double mx, my;
    mom_at(G, M, i, j, mx, my);
    const double rho = S[t], E = S[3 * n + t];
    const double kinetic = 0.5 * (mx * mx + my * my) / rho;
    const Cons ep = star_at(G, S, i + 1, j), em = star_at(G, S, i - 1, j);
    const Cons np = star_at(G, S, i, j + 1), nm = star_at(G, S, i, j - 1);
    r += E;
    r += -kinetic;
    r += -0.5 * G.dt * G.inv_dx * he * ep.mx;
    r += -0.5 * G.dt * G.inv_dx * (he - hw) * S[n + t];
    r += 0.5 * G.dt * G.inv_dx * hw * em.mx;
    r += -0.5 * G.dt * G.inv_dy * hn * np.my;
    r += -0.5 * G.dt * G.inv_dy * (hn - hs) * S[2 * n + t];
    r += 0.5 * G.dt * G.inv_dy * hs * nm.my;
    b[t] = r;
    if (!isfinite(r)) ++bad;
  }
  if (bad) atomicAdd(&ctl->nonfinite, bad);
}

// ------------------------------------------------------------ operator
template <typename T>
struct OpArgs {
  ImexGrid G;
  const T* hx;
  const T* hy;
  T cx, cy, diag;
};

// Row (i,j) of the operator applied to the vector whose entries get(k)
// returns (k a column index).
template <typename T, typename Get>
__device__ __forceinline__ T apply_row_g(const OpArgs<T>& A, const Get& get, int i, int j)
{
  const ImexGrid& G = A.G;
  // Rows at least 2 cells from every side read their stencil directly (the
  // boundary mapping is the identity there).
  const bool inner = i >= 2 && j >= 2 && i < G.nx - 2 && j < G.ny - 2;
  auto at = [&](int ii, int jj) -> T {
    if (inner) return get((long long)jj * G.nx + ii);
    return fixed_side(G, ii, jj) >= 0 ? T(0) : get(pcol(G, ii, jj));
  };
  const long long k = (long long)j * G.nx + i;
  const T pc = get(k);
  const T he = A.hx[(long long)j * (G.nx + 1) + i + 1], hw = A.hx[(long long)j * (G.nx + 1) + i];
  const T hn = A.hy[(long long)(j + 1) * G.nx + i], hs = A.hy[(long long)j * G.nx + i];
  const T xe = at(i + 1, j), xw = at(i - 1, j);
  const T xee = at(i + 2, j), xww = at(i - 2, j);
  const T yn = at(i, j + 1), ys = at(i, j - 1);
  const T ynn = at(i, j + 2), yss = at(i, j - 2);
  const T cxd = xe - xw, cyd = yn - ys;
  const T xpart = he * ((xee - pc) + cxd) - hw * ((pc - xww) + cxd);
  const T ypart = hn * ((ynn - pc) + cyd) - hs * ((pc - yss) + cyd);
  return A.diag * pc - A.cx * xpart - A.cy * ypart;
}

template <typename T>
__device__ __forceinline__ T apply_row(const OpArgs<T>& A, const T* __restrict__ x, int i, int j)
{
  return apply_row_g(A, [x](long long k) -> T { return x[k]; }, i, j);
}

template <typename T>
__global__ void matvec_kernel(OpArgs<T> A, const T* __restrict__ x, T* __restrict__ y, const GmCtl* ctl)
{
  if (ctl && ctl->done) return;
  const long long n = (long long)A.G.nx * A.G.ny;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    y[t] = apply_row(A, x, (int)(t % A.G.nx), (int)(t / A.G.nx));
  }
}

// r = b - A x (FP64), with |r|_inf and |r|_2^2 reduced into ctl.
__global__ void __launch_bounds__(256) residual_kernel(const OpArgs<double>* __restrict__ Ap, const double* __restrict__ x,
                                const double* __restrict__ b, double* __restrict__ r,
                                double* partials, GmCtl* ctl)
{
  const OpArgs<double> A = *Ap;
  const long long n = (long long)A.G.nx * A.G.ny;
  double mx = 0.0, s2 = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const double v = b[t] - apply_row(A, x, (int)(t % A.G.nx), (int)(t / A.G.nx));
    r[t] = v;
    const double a = fabs(v);
    mx = (mx < a || isnan(a)) ? a : mx;
    s2 += v * v;
  }
  (void)ctl;
  reduce_maxsum_store(mx, s2, partials);
}

// ------------------------------------------------------------ GMRES pieces
// w -= V h1 (h1 = dots1), then dots2[i] = V_i . w (CGS pass 2, fused).
// One row per thread (grid = ceil(n / RB)).
template <typename T>
__global__ void __launch_bounds__(RB) update_dot_kernel(const T* __restrict__ V, long long n, int k,
                                                        T* __restrict__ w, double* partials, GmCtl* ctl)
{
  if (ctl->done) return;
  __shared__ double h[MAXK];
  if (threadIdx.x < k) h[threadIdx.x] = ctl->dots1[threadIdx.x];
  __syncthreads();
  const long long t = (long long)blockIdx.x * RB + threadIdx.x;
  double wd = 0.0;
  if (t < n) {
    double wt = (double)w[t];
    for (int q0 = 0; q0 < k; q0 += 8) {   // eight basis loads in flight
      double vb[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) vb[u] = q0 + u < k ? (double)V[(q0 + u) * n + t] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (q0 + u < k) wt -= h[q0 + u] * vb[u];
    }
    const T wn = (T)wt;
    w[t] = wn;
    wd = (double)wn;
  }
  rowdots_store([&](int q) { return t < n ? (double)V[q * n + t] * wd : 0.0; }, k, partials);
}

// w -= V h (h = dots2 after a second Gram-Schmidt pass, dots1 with a single
// pass) and |w|^2.
template <typename T>
__global__ void __launch_bounds__(RB) update_norm_kernel(const T* __restrict__ V, long long n, int k,
                                                         T* __restrict__ w, double* partials, GmCtl* ctl,
                                                         int pass)
{
  if (ctl->done) return;
  __shared__ double h[MAXK];
  if (threadIdx.x < k) h[threadIdx.x] = pass == 1 ? ctl->dots1[threadIdx.x] : ctl->dots2[threadIdx.x];
  __syncthreads();
  DSum<1> s;
  s.v[0] = 0.0;
  for (long long t = (long long)blockIdx.x * RB + threadIdx.x; t < n; t += (long long)gridDim.x * RB) {
    double wt = (double)w[t];
    for (int q0 = 0; q0 < k; q0 += 8) {   // eight basis loads in flight
      double vb[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) vb[u] = q0 + u < k ? (double)V[(q0 + u) * n + t] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (q0 + u < k) wt -= h[q0 + u] * vb[u];
    }
    const T wn = (T)wt;
    w[t] = wn;
    s.v[0] += (double)wn * (double)wn;
  }
  reduce_sum_store<1>(s, partials, 1);
}

// Fixed-order workgroup sum (valid in thread 0).
__device__ __forceinline__ double block_sum_fixed(double v, double* sh)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_sum_lane63(v);
  if (lane == 63) sh[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    v = sh[0];
    for (int w = 1; w < RB / 64; ++w) v += sh[w];
  }
  return v;
}

// out[q] = sum over blocks of partials[b*K + q]; one workgroup per q, each
// summing in a fixed order (run-to-run reproducible).
__global__ void __launch_bounds__(RB) finalize_dots_kernel(const double* partials, int nparts, int K,
                                                           double* out, const GmCtl* ctl)
{
  __shared__ double sh[RB / 64];
  if (ctl && ctl->done) return;
  const int q = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nparts; b += RB) v += partials[(long long)b * K + q];
  v = block_sum_fixed(v, sh);
  if (threadIdx.x == 0) out[q] = v;
}

// Column j's Givens step (thread 0, on the LDS-staged column and rotations):
// the column's earlier rotations, the new rotation, the residual estimate
// and the early-exit decision, from |w|^2 = v.  With `write` the results are
// committed (R column, rotation, g, control fields); every caller computes
// the same values, so only one workgroup needs to write.
struct GivensOut {
  double inv_norm;
  int done;
};

// The control values column j's Givens step reads, loaded together (one
// memory round trip) and ahead of the chain that needs them.
struct GivensIn {
  double beta, thr, gj;
  int bad, breakdown, mlim;
};

__device__ __forceinline__ GivensIn givens_in(const GmCtl* ctl, const double* g, int j)
{
  return GivensIn{ctl->beta, ctl->thr, g[j], ctl->bad, ctl->breakdown, ctl->mlim};
}

__device__ GivensOut givens_step(double v, double* col, const double* c, const double* sv,
                                 const GivensIn& in, double* R, double* g, double* cs, double* sn,
                                 int j, int m, GmCtl* ctl, bool write)
{
  // breakdown tolerance 100 eps max(beta, 1) (imex_pressure_helpers.hpp:167-169)
  const double breakdown_tol = 2.220446049250313e-16 * 100.0 * smax(in.beta, 1.0);
  const double hn = sqrt(v);
  // Rotation i mixes entries i and i+1; the rotated entry i+1 is the next
  // rotation's first operand, so it is carried in a register and the loads
  // of the column stay off the dependency chain.
  double carry = col[0];
  for (int i = 0; i < j; ++i) {
    const double a = carry, b = col[i + 1];
    col[i] = c[i] * a + sv[i] * b;
    carry = -sv[i] * a + c[i] * b;
  }
  const double a = carry, b = hn;
  const double rr = hypot(a, b);
  const double cj = rr == 0.0 ? 1.0 : a / rr;
  const double sj = rr == 0.0 ? 0.0 : b / rr;
  col[j] = rr;
  col[j + 1] = 0.0;
  const double gj = in.gj;
  const double gnext = -sj * gj;
  const double est = fabs(gnext);
  const int bad = (!isfinite(hn) || !isfinite(est) || !isfinite(rr)) ? 1 : in.bad;
  const int brk = hn <= breakdown_tol ? 1 : in.breakdown;
  const int done = (est <= in.thr || j + 1 >= in.mlim || bad || brk) ? 1 : 0;
  if (write) {
    cs[j] = cj;
    sn[j] = sj;
    g[j + 1] = gnext;   // the running last entry, read by column j + 1
    g[m + 1 + j] = cj * gj;  // the rotated entry, kept apart so g[j] stays readable
    ctl->nrm2 = v;
    ctl->est = est;
    ctl->kk = j + 1;
    ctl->inv_norm = 1.0 / hn;
    ctl->bad = bad;
    ctl->breakdown = brk;
    if (done) ctl->done = 1;
    double* rc = R + (long long)j * (m + 1);  // column-major (m+1) x m
    for (int i = 0; i <= j + 1; ++i) rc[i] = col[i];
  }
  return GivensOut{1.0 / hn, done};
}

__device__ GivensOut givens_step(double v, double* col, const double* c, const double* sv, double* R,
                                 double* g, double* cs, double* sn, int j, int m, GmCtl* ctl,
                                 bool write)
{
  return givens_step(v, col, c, sv, givens_in(ctl, g, j), R, g, cs, sn, j, m, ctl, write);
}

// Stages column j (first- plus, with two passes, second-pass dots) and the
// earlier rotations in LDS and sums |w|^2 from the norm partials (fixed
// order); result valid in thread 0 after the barrier inside.
__device__ double givens_stage(const double* npart, int nparts, const double* cs, const double* sn,
                               int j, const GmCtl* ctl, int two_pass, double* sh, double* col,
                               double* c, double* sv)
{
  double v = 0.0;
  for (int b = threadIdx.x; b < nparts; b += RB) v += npart[b];
  for (int i = threadIdx.x; i <= j; i += RB) {
    col[i] = two_pass ? ctl->dots1[i] + ctl->dots2[i] : ctl->dots1[i];
    if (i < j) {
      c[i] = cs[i];
      sv[i] = sn[i];
    }
  }
  return block_sum_fixed(v, sh);  // its barrier also publishes col, c, sv
}

// The Givens step of the cycle's last column (every other column's step runs
// inside the next Arnoldi kernel).
__global__ void __launch_bounds__(RB) norm_givens_kernel(const double* npart, int nparts, double* R,
                                                         double* g, double* cs, double* sn, int j,
                                                         int m, GmCtl* ctl, int two_pass)
{
  __shared__ double sh[RB / 64];
  __shared__ double col[MAXK + 1], c[MAXK], sv[MAXK];
  if (ctl->done) return;
  const double v = givens_stage(npart, nparts, cs, sn, j, ctl, two_pass, sh, col, c, sv);
  if (threadIdx.x == 0) (void)givens_step(v, col, c, sv, R, g, cs, sn, j, m, ctl, true);
}

// y = R^{-1} g for the kk x kk upper triangle, by the 64 lanes of one
// wavefront: column-oriented back substitution, one lane per row, the
// columns staged in LDS (Rs, MAXK x MAXK) and each solved entry broadcast
// from its lane; y_q times scale[q] when a scale is given (the one-reduction
// cycle's unnormalised basis).  g holds the rotated right-hand side at
// g[m + 1 + q].
__device__ void backsolve_wave(const double* R, const double* g, double* y, int m, int kk, GmCtl* ctl,
                               const double* scale, double* Rs)
{
  const int r = threadIdx.x & 63;
  const int rc = r < m ? r : m;  // clamped row: every load is in bounds
  for (int q0 = 0; q0 < kk; q0 += 8) {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = q0 + u < kk ? q0 + u : kk - 1;
      t[u] = R[(long long)q * (m + 1) + rc];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q0 + u < kk) Rs[(q0 + u) * MAXK + r] = r <= q0 + u ? t[u] : 0.0;
  }
  double s = g[m + 1 + rc];
  s = r < kk ? s : 0.0;
  __builtin_amdgcn_wave_barrier();
  double yr = 0.0;
  for (int i = kk - 1; i >= 0; --i) {
    const long long sb = __double_as_longlong(s);
    const int lo = __builtin_amdgcn_readlane((int)sb, i);
    const int hi = __builtin_amdgcn_readlane((int)(sb >> 32), i);
    const double yi = __longlong_as_double(((long long)hi << 32) | (unsigned int)lo) / Rs[i * MAXK + i];
    if (r == i) yr = yi;
    if (r < i) s -= Rs[i * MAXK + r] * yi;
  }
  if (r < kk) {
    if (!isfinite(yr)) ctl->bad = 1;
    y[r] = scale ? yr * scale[r] : yr;
  }
}

// ------------------------------------------- one-reduction GMRES cycle (FP32)
// Classical Gram-Schmidt with lagged normalisation: one global reduction per
// Arnoldi step (the one-synchronisation CGS of Swirydowicz, Langou,
// Ananthan, Yang and Thomas, "Low synchronization Gram-Schmidt and GMRES
// algorithms", Numer. Linear Algebra Appl. 28 (2021) e2343).  The basis is
// kept unnormalised, u_q = nu_q v_q with nu_q = |u_q| (nu_0 = 1: u_0 = r /
// beta), next to z = A u of the newest direction.  Kernel k:
//   0. (k > 0) sums the previous kernel's tile partials in a fixed order, in
//      every workgroup (beyond ONE_SUM_TILES tiles it reads the totals of a
//      fixed-order finalize kernel instead): nu_{k-1} = |u_{k-1}|,
//      h_{q,k-1} = (u_q . z_{k-1}) / (nu_q nu_{k-1}), and column k-2's
//      Givens step with its subdiagonal entry nu_{k-1} (workgroup 0 commits
//      H, the scales and the rotation); every workgroup exits on the same
//      early-exit decision;
//   1. forms u_k = z_{k-1} / nu_{k-1} - sum_q h_{q,k-1} u_q / nu_q on its
//      tile and the operator's cross-shaped 2-cell halo, in LDS (FP64
//      accumulation, stored in T), and stores the tile's cells;
//   2. applies the operator from LDS, z_k = A u_k;
//   3. reduces u_q . z_k (q <= k) and u_k . u_k per tile.
// The mathematics is that of CGS1 (the FP32 inner solver's
// orthogonalisation), in one kernel per Arnoldi step instead of three.
// z and the tile partials ping-pong between consecutive launches: a launch
// can have more workgroups than the chip holds, so workgroups of one launch
// run one after another, and none may overwrite what a later one reads.
constexpr int OTX = 128, OTY = 8, ONT = OTX * OTY;  // one cell per thread
constexpr int OHX = OTX + 4, OHY = OTY + 4, ONR = OHX * OHY;
constexpr int KP = MAXK + 2;  // partial entries per tile: MAXK + 1 dots and one norm
constexpr int ONE_SUM_TILES = 256;  // tiles up to which each workgroup sums the partials itself

static inline int onesync_tiles(int nx, int ny)
{
  return ((nx + OTX - 1) / OTX) * ((ny + OTY - 1) / OTY);
}

// tot[e] = sum over tiles of partials[b*KP + e] for e in [e0, e1), in a
// fixed order (lanes over tiles, then the wavefront): every workgroup gets
// the same bits.
__device__ __forceinline__ void sum_tile_entries(const double* partials, int nparts, int e0, int e1,
                                                 double* tot)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int e = e0 + wave; e < e1; e += ONT / 64) {
    double v = 0.0;
    for (int b = lane; b < nparts; b += 64) v += partials[(long long)b * KP + e];
    v = wave_sum_lane63(v);
    if (lane == 63) tot[e] = v;
  }
}

// Tile-local (li, lj) of halo slot h < NHALO: the two rows above the tile,
// the two below, then the two columns left and the two right.
constexpr int NHALO = 4 * OTX + 4 * OTY;
__device__ __forceinline__ void halo_slot(int h, int& li, int& lj)
{
  if (h < 4 * OTX) {
    const int r = h % (2 * OTX), below = h / (2 * OTX);
    li = 2 + r % OTX;
    lj = (below ? OTY + 2 : 0) + r / OTX;
  } else {
    const int s = h - 4 * OTX, r = s % (2 * OTY), right = s / (2 * OTY);
    li = (right ? OTX + 2 : 0) + (r & 1);
    lj = 2 + (r >> 1);
  }
}
static_assert(NHALO <= ONT, "one halo slot per thread at most");

// u_k at column c: z_{k-1} / nu_{k-1} - sum_q c_q u_q (FP64 accumulation),
// with the first QB basis values already loaded (uv) and the rest read here;
// for k = 0, z is u_0 itself.
template <typename T, int QB>
__device__ __forceinline__ T form_value(const T (&uv)[QB], T z, const T* __restrict__ U, long long n,
                                        long long c, const double* coef, double cz, int k)
{
  if (k == 0) return z;
  double sum = (double)z * cz;
#pragma unroll
  for (int q = 0; q < QB; ++q)
    if (q < k) sum -= coef[q] * (double)uv[q];
  for (int q0 = QB; q0 < k; q0 += 8) {  // basis vectors beyond the held ones
    T ub[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) ub[u] = U[(long long)(q0 + u < k ? q0 + u : 0) * n + c];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q0 + u < k) sum -= coef[q0 + u] * (double)ub[u];
  }
  return (T)sum;
}

template <typename T, int QB, bool SUMMED>
__global__ void __launch_bounds__(ONT) onesync_kernel(const OpArgs<T>* __restrict__ Ap, T* __restrict__ U,
                                                      T* __restrict__ Z, int k, int m,
                                                      const double* __restrict__ partials,
                                                      double* __restrict__ pout, int nparts,
                                                      GmCtl* ctl, double* hraw,
                                                      double* inv_nu, double* R, double* g, double* cs,
                                                      double* sn)
{
  __shared__ T tile[ONR];
  __shared__ double tot[KP];
  __shared__ double coef[MAXK];
  __shared__ double col[MAXK + 1], gc[MAXK], gsv[MAXK];
  __shared__ double sh[ONT / 64][KP];
  __shared__ int s_done;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int GIVENS_TID = ONT - 64;  // a lane of the last wave, which has no halo slot
  static_assert(GIVENS_TID >= NHALO, "the Givens lane's wave forms no halo slot");
  // workgroup 0 may set ctl->done during this kernel: read it once per
  // workgroup so every wave takes the same exit
  if (tid == 0) s_done = ctl->done;
  __syncthreads();
  if (s_done) return;

  const OpArgs<T> A = *Ap;
  const int nx = A.G.nx, ny = A.G.ny;
  const long long n = (long long)nx * ny;
  const int tiles_x = (nx + OTX - 1) / OTX;
  const int tile_id = xcd_banded(blockIdx.x, gridDim.x);
  const int i0 = (tile_id % tiles_x) * OTX, j0 = (tile_id / tiles_x) * OTY;
  const int ci = tid % OTX, cj = tid / OTX;
  const int gi = i0 + ci, gj = j0 + cj;
  const bool own = gi < nx && gj < ny;
  const long long kown = (long long)gj * nx + gi;
  const bool apply = k < m;  // the last kernel only needs |u_m|

  // Column read for tile slot (li, lj): the cell itself, its mapped column
  // outside the domain, or false for a FIXED side's zero.
  auto slot_col = [&](int li, int lj, long long& c) -> bool {
    const int i = i0 + li - 2, j = j0 + lj - 2;
    c = 0;
    if (i >= 0 && j >= 0 && i < nx && j < ny) {
      c = (long long)j * nx + i;
      return true;
    }
    if (fixed_side(A.G, i, j) >= 0) return false;
    c = pcol(A.G, i, j);
    return true;
  };
  const int oslot = (cj + 2) * OHX + (ci + 2);
  long long co, ch = 0;
  const bool vo = slot_col(ci + 2, cj + 2, co);
  const bool has_h = tid < NHALO;
  int hslot = 0;
  bool vh = false;
  if (has_h) {
    int hli, hlj;
    halo_slot(tid, hli, hlj);
    hslot = hlj * OHX + hli;
    vh = slot_col(hli, hlj, ch);
  }

  // Every global read that does not depend on the reductions, issued up
  // front and unconditionally (indices clamped into the arrays, unused values
  // selected away), so that none waits for another.  The previous kernel's
  // partials go first: loads complete in order, so stage 0 can run while the
  // basis loads behind them are still in flight.
  // SUMMED: partials holds the previous kernel's k + 1 totals (a separate
  // fixed-order sum, for grids of many tiles); otherwise one partial per tile.
  double pv[2][4], tv = 0.0;
  if constexpr (SUMMED) {
    tv = partials[tid < KP ? tid : 0];
  } else {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int e = wave + 16 * r;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int b = lane + 64 * i;
        pv[r][i] = partials[(long long)(b < nparts ? b : 0) * KP + e];
      }
    }
  }
  // Givens inputs of column k-2 and the basis scales, ahead of the basis
  // loads too (thread i holds entry i)
  const int gi_ = tid < m ? tid : 0;  // clamped: every load is in bounds
  const double gcol = hraw[(long long)(k >= 2 ? k - 2 : 0) * (m + 1) + gi_];
  const double gcs = cs[gi_], gsn = sn[gi_];
  const double ginv = inv_nu[gi_];
  const GivensIn gin = givens_in(ctl, g, k >= 2 ? k - 2 : 0);
  const int hi = gi < nx ? gi : nx - 1, hj = gj < ny ? gj : ny - 1;
  const T he = A.hx[(long long)hj * (nx + 1) + hi + 1];
  const T hw = A.hx[(long long)hj * (nx + 1) + hi];
  const T hn = A.hy[(long long)(hj + 1) * nx + hi];
  const T hs = A.hy[(long long)hj * nx + hi];
  const T* zsrc = k == 0 ? U : Z + (long long)((k + 1) & 1) * n;  // u_0, or z_{k-1}
  const T zo = zsrc[co], zh = zsrc[ch];
  T uo[QB], uh[QB];
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const long long qn = (long long)(q < k ? q : 0) * n;
    uo[q] = U[qn + co];
    uh[q] = U[qn + ch];
  }

  // 0. the previous kernel's reductions; column k-2's Givens step runs in
  // one lane while the others form u_k
  double cz = 1.0;
  if (k > 0) {
    if constexpr (SUMMED) {
      if (tid <= k) tot[tid] = tv;
    } else {
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const int e = wave + 16 * r;
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) v += lane + 64 * i < nparts ? pv[r][i] : 0.0;
        if (nparts > 256) {  // more tiles than the held partials cover
          for (int b = lane + 256; b < nparts; b += 64) v += partials[(long long)b * KP + e];
        }
        v = wave_sum_lane63(v);
        if (lane == 63 && e <= k) tot[e] = v;
      }
      if (k >= 32) sum_tile_entries(partials, nparts, 32, k + 1, tot);
    }
    if (k >= 2 && tid <= k - 2) {
      col[tid] = gcol;
      if (tid < k - 2) {
        gc[tid] = gcs;
        gsv[tid] = gsn;
      }
    }
    __syncthreads();
    const double inv_prev = k == 1 ? 1.0 : 1.0 / sqrt(tot[k]);  // 1 / |u_{k-1}|
    if (tid < k) {  // k <= MAXK < ONT: one entry per thread
      const double iq = tid == k - 1 ? inv_prev : ginv;
      const double h = tot[tid] * iq * inv_prev;
      coef[tid] = h * iq;
      if (blockIdx.x == 0) hraw[(long long)(k - 1) * (m + 1) + tid] = h;
    }
    if (tid == 0 && blockIdx.x == 0) inv_nu[k - 1] = inv_prev;
    __syncthreads();
    if (tid == GIVENS_TID && k >= 2) {
      s_done = givens_step(tot[k], col, gc, gsv, gin, R, g, cs, sn, k - 2, m, ctl, blockIdx.x == 0).done;
    }
    cz = inv_prev;
  }

  // 1. u_k at the own and halo slots (FP64 accumulation, stored in T)
  if (has_h) tile[hslot] = vh ? form_value<T, QB>(uh, zh, U, n, ch, coef, cz, k) : T(0);
  tile[oslot] = vo ? form_value<T, QB>(uo, zo, U, n, co, coef, cz, k) : T(0);
  __syncthreads();
  if (s_done) return;

  // 2. z_k = A u_k (the operator row of apply_row_g, from LDS)
  double ud = 0.0, zd = 0.0;
  if (own) {
    const T pc = tile[oslot];
    if (apply) {
      const T xe = tile[oslot + 1], xw = tile[oslot - 1];
      const T xee = tile[oslot + 2], xww = tile[oslot - 2];
      const T yn = tile[oslot + OHX], ys = tile[oslot - OHX];
      const T ynn = tile[oslot + 2 * OHX], yss = tile[oslot - 2 * OHX];
      const T cxd = xe - xw, cyd = yn - ys;
      const T xpart = he * ((xee - pc) + cxd) - hw * ((pc - xww) + cxd);
      const T ypart = hn * ((ynn - pc) + cyd) - hs * ((pc - yss) + cyd);
      const T zt = A.diag * pc - A.cx * xpart - A.cy * ypart;
      Z[(long long)(k & 1) * n + kown] = zt;
      if (k > 0) U[(long long)k * n + kown] = pc;
      zd = (double)zt;
    }
    ud = (double)pc;
  }

  // 3. tile partials: u_q . z_k (q < k from the held values), u_k . z_k and
  // u_k . u_k, eight entries per wavefront butterfly
  const int ne = k + 2;
  auto tail = [&](int e) -> double { return e == k ? ud * zd : (e == k + 1 ? ud * ud : 0.0); };
  auto store8 = [&](int e0, const double (&v)[8]) {
    const double t8 = wave_sum8(v);
    const int e = e0 + (lane >> 3);
    if ((lane & 7) == 0 && e < ne) sh[wave][e] = t8;
  };
#pragma unroll
  for (int e0 = 0; e0 < QB; e0 += 8) {
    if (e0 < ne) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u;
        v[u] = !own ? 0.0 : e < k ? (apply ? (double)uo[e] * zd : 0.0) : tail(e);
      }
      store8(e0, v);
    }
  }
  for (int e0 = QB; e0 < ne; e0 += 8) {  // basis vectors beyond the held ones
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u;
      const double uq = (double)U[(long long)(e < k ? e : 0) * n + co];
      v[u] = !own ? 0.0 : e < k ? (apply ? uq * zd : 0.0) : tail(e);
    }
    store8(e0, v);
  }
  __syncthreads();
  if (tid < ne) {
    double v = sh[0][tid];
    for (int w = 1; w < ONT / 64; ++w) v += sh[w][tid];
    pout[(long long)blockIdx.x * KP + tid] = v;
  }
  if (k == 0 && blockIdx.x == 0 && tid == 0) inv_nu[0] = 1.0;  // u_0 = r / beta
}

// Closes a one-reduction cycle: |u_m| from the last kernel's partials and
// column m-1's Givens step (unless the cycle ended earlier), then wavefront 0
// solves for the cycle's coefficients, y_q / |u_q|.
__global__ void __launch_bounds__(ONT) onesync_close_kernel(const double* partials, int nparts, int m,
                                                            GmCtl* ctl, const double* hraw, double* R,
                                                            double* g, double* cs, double* sn,
                                                            double* y, const double* inv_nu)
{
  __shared__ double tot[KP];
  __shared__ double col[MAXK + 1], gc[MAXK], gsv[MAXK];
  __shared__ double Rs[MAXK * MAXK];
  __shared__ int s_done, s_kk;
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_done = ctl->done;
    s_kk = ctl->kk;
  }
  __syncthreads();
  if (!s_done) {
    sum_tile_entries(partials, nparts, m + 1, m + 2, tot);
    for (int i = tid; i <= m - 1; i += ONT) {
      col[i] = hraw[(long long)(m - 1) * (m + 1) + i];
      if (i < m - 1) {
        gc[i] = cs[i];
        gsv[i] = sn[i];
      }
    }
    __syncthreads();
    if (tid == 0) {
      (void)givens_step(tot[m + 1], col, gc, gsv, R, g, cs, sn, m - 1, m, ctl, true);
      s_kk = m;
    }
  }
  __syncthreads();  // also orders thread 0's R and g writes before wavefront 0's reads
  if (tid < 64) backsolve_wave(R, g, y, m, s_kk, ctl, inv_nu, Rs);
}

// ------------------------------------------------ mixed form: FGMRES-IR
// The mixed form is iterative refinement accelerated by a flexible outer
// Krylov solver in FP64 (FGMRES, Saad, SIAM J. Sci. Comput. 14 (1993)
// 461-469) whose preconditioner is the FP32 inner GMRES(m) (Arioli and
// Duff, "Using FGMRES to obtain backward stability in mixed precision",
// Electron. Trans. Numer. Anal. 33 (2009) 31-44).  Each outer step solves
// A z_j = v_j approximately in FP32; the outer basis, its orthogonalisation,
// the update x += Z y and the true residual are FP64, so the attainable
// accuracy is FP64's while the Krylov work is FP32.
constexpr int MO = 20;  // outer restart length

// A fresh inner cycle on the unit right-hand side v_j: u_0 = v_j in FP32,
// g_0 = 1 and the inner control block reset (done at once when the outer
// cycle is done).
template <typename T>
__global__ void __launch_bounds__(RB) inner_start_kernel(const double* __restrict__ v,
                                                         T* __restrict__ U, long long n, GmCtl* ci,
                                                         GmCtl* co, double* g)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // the iterations left of euler.imex_solver_max_iter: a spent budget
    // closes the outer cycle before this step (its kernels read co->done)
    const int left = co->inner_budget - co->inner_its;
    if (left <= 0) co->done = 1;
    ci->done = co->done;
    ci->kk = 0;
    ci->bad = 0;
    ci->breakdown = 0;
    // the requested restart (the graph holds the workspace's m), at most the budget left
    ci->mlim = left < co->inner_mlim ? left : co->inner_mlim;
    ci->beta = 1.0;
    ci->thr = co->inner_thr;
    g[0] = 1.0;
  }
  for (long long t = (long long)blockIdx.x * RB + threadIdx.x; t < n; t += (long long)gridDim.x * RB) {
    U[t] = (T)v[t];
  }
}

// Outer step j: y = A z_j (FP64) into w and the first Gram-Schmidt pass's
// dots V_q . y (q <= j).
__global__ void __launch_bounds__(RB) outer_arnoldi_kernel(const OpArgs<double>* __restrict__ Ap,
                                                           const double* __restrict__ V,
                                                           const double* __restrict__ z, int j,
                                                           double* __restrict__ w, double* partials,
                                                           const GmCtl* ctl)
{
  if (ctl->done) return;  // nothing writes ctl during this kernel
  const OpArgs<double> A = *Ap;
  const long long n = (long long)A.G.nx * A.G.ny;
  const long long t = (long long)blockIdx.x * RB + threadIdx.x;
  double yd = 0.0;
  if (t < n) {
    yd = apply_row(A, z, (int)(t % A.G.nx), (int)(t / A.G.nx));
    w[t] = yd;
  }
  rowdots_store([&](int q) { return t < n ? V[q * n + t] * yd : 0.0; }, j + 1, partials);
}

// V_{j+1} = w / |w|, and the inner iterations of step j counted into the
// outer control block.
__global__ void __launch_bounds__(RB) outer_next_kernel(const double* __restrict__ w,
                                                        double* __restrict__ vnext, long long n,
                                                        GmCtl* co, const GmCtl* ci)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) co->inner_its += ci->kk;
  if (co->done) return;
  const double s = co->inv_norm;
  const long long t = (long long)blockIdx.x * RB + threadIdx.x;
  if (t < n) vnext[t] = w[t] * s;
}

// Arnoldi iteration j, first half.  For j > 0 every workgroup first closes
// column j - 1 (its Givens step, from the norm partials of the previous
// iteration; workgroup 0 commits it) and exits if that ends the cycle.  Then
// v_j = w_prev / |w_prev| (v_0 is already in V) is formed on the fly and
// stored, y = A v_j, and the first Gram-Schmidt pass's dots V_q . y (q <= j)
// are accumulated in the same sweep, so the new column and the basis are each
// read once.
template <typename T>
__global__ void __launch_bounds__(RB) arnoldi_kernel(const OpArgs<T>* __restrict__ Ap, T* __restrict__ V,
                                                     int j, const T* __restrict__ wprev,
                                                     T* __restrict__ y, double* partials, GmCtl* ctl,
                                                     const double* npart, int nparts, double* R,
                                                     double* g, double* cs, double* sn, int m,
                                                     int two_pass)
{
  __shared__ double gsh[RB / 64];
  __shared__ double col[MAXK + 1], gc[MAXK], gsv[MAXK];
  __shared__ double s_inv;
  __shared__ int s_done;
  // workgroup 0 may set ctl->done during this kernel: read it once per
  // workgroup so every wave takes the same exit
  if (threadIdx.x == 0) s_done = ctl->done;
  __syncthreads();
  if (s_done) return;
  double s = 1.0;
  if (j > 0) {
    // Column j-1's Givens step, redundantly in every workgroup (workgroup 0
    // commits it), then v_j = w / |w| below.
    const double v = givens_stage(npart, nparts, cs, sn, j - 1, ctl, two_pass, gsh, col, gc, gsv);
    if (threadIdx.x == 0) {
      const GivensOut o = givens_step(v, col, gc, gsv, R, g, cs, sn, j - 1, m, ctl, blockIdx.x == 0);
      s_inv = o.inv_norm;
      s_done = o.done;
    }
    __syncthreads();
    if (s_done) return;
    s = s_inv;
  }
  const OpArgs<T> A = *Ap;
  const long long n = (long long)A.G.nx * A.G.ny;
  T* vj = V + (long long)j * n;
  const long long t = (long long)blockIdx.x * RB + threadIdx.x;
  double yd = 0.0, vd = 0.0;
  if (t < n) {
    const int i = (int)(t % A.G.nx), jj = (int)(t / A.G.nx);
    T yt, vt;
    if (j == 0) {
      yt = apply_row(A, vj, i, jj);
      vt = vj[t];
    } else {
      const auto get = [wprev, s](long long k) -> T { return (T)((double)wprev[k] * s); };
      yt = apply_row_g(A, get, i, jj);
      vt = get(t);
      vj[t] = vt;
    }
    y[t] = yt;
    yd = (double)yt;
    vd = (double)vt;
  }
  rowdots_store([&](int q) { return t >= n ? 0.0 : (q < j ? (double)V[q * n + t] : vd) * yd; }, j + 1,
                partials);
}

// Start a cycle: v0 = r / beta, g = beta e1.
template <typename T>
__global__ void __launch_bounds__(256) cycle_start_kernel(const double* __restrict__ r, T* __restrict__ v0, long long n,
                                   const GmCtl* ctl, double* g)
{
  const double beta = ctl->beta;
  const double s = 1.0 / beta;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    v0[t] = (T)(r[t] * s);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) g[0] = beta;
}

// y = R^{-1} g (backsolve_wave), then x += V y (x in FP64; V in T).
__global__ void __launch_bounds__(64) backsolve_kernel(const double* R, const double* g, double* y,
                                                       int m, GmCtl* ctl, const double* scale)
{
  __shared__ double Rs[MAXK * MAXK];
  backsolve_wave(R, g, y, m, ctl->kk, ctl, scale, Rs);
}

template <typename T, typename X>
__global__ void __launch_bounds__(256) xupdate_kernel(const T* __restrict__ V, long long n, const double* __restrict__ y,
                               X* __restrict__ x, double scale, const GmCtl* ctl, int assign)
{
  const int kk = ctl->kk;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q0 = 0; q0 < kk; q0 += 8) {
      double vb[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) vb[u] = q0 + u < kk ? (double)V[(q0 + u) * n + t] : 0.0;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (q0 + u < kk) s += y[q0 + u] * vb[u];
    }
    x[t] = assign ? (X)(scale * s) : (X)((double)x[t] + scale * s);
  }
}

// ------------------------------------------------------ refresh / closeout
// m = m* - dt grad_c p (imex_source_map_helpers.hpp:249-268, with the
// physical-boundary pressure rule).
__global__ void __launch_bounds__(256) momentum_refresh_kernel(ImexGrid G, const double* __restrict__ S,
                                        const double* __restrict__ P, double* __restrict__ M, GmCtl* ctl)
{
  const long long n = (long long)G.nx * G.ny;
  int bad = 0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % G.nx), j = (int)(t / G.nx);
    const double gx = 0.5 * (p_at(G, P, i + 1, j) - p_at(G, P, i - 1, j)) * G.inv_dx;
    const double gy = 0.5 * (p_at(G, P, i, j + 1) - p_at(G, P, i, j - 1)) * G.inv_dy;
    const double mx = S[n + t] - G.dt * gx, my = S[2 * n + t] - G.dt * gy;
    M[t] = mx;
    M[n + t] = my;
    if (!isfinite(mx) || !isfinite(my)) ++bad;
  }
  if (bad) atomicAdd(&ctl->mom_bad, bad);
}

// E = E* - dt div(h m), the new state, and the coupled residual
// p/(g-1) - E + K (imex_source_map_step.hpp:95-187).
__global__ void __launch_bounds__(256) closeout_kernel(ImexGrid G, const double* __restrict__ S, const double* __restrict__ M,
                                const double* __restrict__ P, const double* __restrict__ fx,
                                const double* __restrict__ fy, double* __restrict__ U, long long rs,
                                long long cs, double* partials, GmCtl* ctl)
{
  const long long n = (long long)G.nx * G.ny;
  const double gm1 = G.gamma - 1.0;
  double v[CLOSE_K];
  close_record_init(v);
  int bad = 0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % G.nx), j = (int)(t / G.nx);
    const double fe = fx[(long long)j * (G.nx + 1) + i + 1], fw = fx[(long long)j * (G.nx + 1) + i];
    const double fn = fy[(long long)(j + 1) * G.nx + i], fs = fy[(long long)j * G.nx + i];
    const double div = (fe - fw) * G.inv_dx + (fn - fs) * G.inv_dy;
    const double E = S[3 * n + t] - G.dt * div;
    const double mx = M[t], my = M[n + t], rho = S[t], p = P[t];
    const double kin = 0.5 * (mx * mx + my * my) / rho;
    const double internal = E - kin;
    const double eos = gm1 * internal;
    const double res = p / gm1 - E + kin;
    const long long k = (long long)j * rs + i;
    U[k] = rho;
    U[cs + k] = mx;
    U[2 * cs + k] = my;
    U[3 * cs + k] = E;
    if (!isfinite(rho) || !isfinite(mx) || !isfinite(my) || !isfinite(E) || !isfinite(p) || !isfinite(kin) ||
        !isfinite(internal) || !isfinite(eos) || !isfinite(res) || !isfinite(div)) {
      ++bad;
      continue;
    }
    const double dmx = mx - S[n + t], dmy = my - S[2 * n + t];
    v[CL_RES] = smax(v[CL_RES], fabs(res));
    v[CL_RHO_MIN] = smin(v[CL_RHO_MIN], rho);
    v[CL_P_MIN] = smin(v[CL_P_MIN], p);
    v[CL_P_MAX] = smax(v[CL_P_MAX], p);
    v[CL_EOS_MIN] = smin(v[CL_EOS_MIN], eos);
    v[CL_INTERNAL_MIN] = smin(v[CL_INTERNAL_MIN], internal);
    v[CL_MOM_CORR] = smax(v[CL_MOM_CORR], sqrt(dmx * dmx + dmy * dmy));
    v[CL_E_DELTA] = smax(v[CL_E_DELTA], fabs(E - S[3 * n + t]));
  }
  if (bad) atomicAdd(&ctl->nonfinite, bad);
  close_record_store(v, partials);
}

// The eq78 velocity divergence of the closed state (imex/
// imex_source_map_step.hpp:207-227): central differences of the velocities,
// the neighbours wrapped on a periodic axis and mirrored at a physical side
// (source_map_index, imex/imex_source_map_helpers.hpp:169-188); a velocity
// reads NaN where the density is zero or either value non-finite, and cells
// whose divergence is not finite are skipped.  Partials: (max |div|, count).
__device__ __forceinline__ int source_map_mirror(int i, int n, bool periodic)
{
  if (periodic) return ((i % n) + n) % n;
  if (i < 0) return min(-i - 1, n - 1);
  if (i >= n) return max(n - (i - n + 1), 0);
  return i;
}

__device__ __forceinline__ double closed_velocity(const ImexGrid& G, const double* U, long long rs, long long cs,
                                                  int i, int j, int axis)
{
  const int ii = source_map_mirror(i, G.nx, G.bc[0] == BC_PERIODIC);
  const int jj = source_map_mirror(j, G.ny, G.bc[2] == BC_PERIODIC);
  const long long k = (long long)jj * rs + ii;
  const double rho = U[k], m = U[(axis == 0 ? cs : 2 * cs) + k];
  if (!isfinite(rho) || rho == 0.0 || !isfinite(m)) return NAN;
  return m / rho;
}

__global__ void __launch_bounds__(256) eq78_kernel(ImexGrid G, const double* __restrict__ U, long long rs,
                                                   long long cs, double* partials)
{
  const long long n = (long long)G.nx * G.ny;
  const double den_scale = smax(G.inv_dx, G.inv_dy);
  double dmax = 0.0, count = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % G.nx), j = (int)(t / G.nx);
    const double ue = closed_velocity(G, U, rs, cs, i + 1, j, 0);
    const double uw = closed_velocity(G, U, rs, cs, i - 1, j, 0);
    const double vn = closed_velocity(G, U, rs, cs, i, j + 1, 1);
    const double vs = closed_velocity(G, U, rs, cs, i, j - 1, 1);
    const double div = 0.5 * (ue - uw) * G.inv_dx + 0.5 * (vn - vs) * G.inv_dy;
    if (!isfinite(div)) continue;
    // velocity_scale >= 1 and finite here (every velocity is finite when div is)
    const double scale = smax(smax(smax(smax(fabs(ue), fabs(uw)), fabs(vn)), fabs(vs)), 1.0);
    const double den = scale * den_scale;
    if (!isfinite(den) || den <= 0.0) continue;
    dmax = smax(dmax, fabs(div));
    count += 1.0;
  }
  reduce_maxsum_store(dmax, count, partials);
}

// compute_imex_dt reductions (diagnostic_metrics_timestep.hpp:262-326): the
// eq79 material rate 2|u|/min_dx and the acoustic rate.
__global__ void __launch_bounds__(256) imex_rates_kernel(const double* __restrict__ U, long long rs, long long cs, GridParams g,
                                  double* partials, GmCtl* ctl)
{
  const long long n = (long long)g.nx * g.ny;
  const double min_dx = smin(g.dx, g.dy);
  double mat = -INFINITY, aco = -INFINITY;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % g.nx), j = (int)(t / g.nx);
    const long long k = (long long)j * rs + i;
    const Prim q = to_prim(Cons{U[k], U[cs + k], U[2 * cs + k], U[3 * cs + k]}, g.gamma);
    const double c = sound_speed(q, g.gamma);
    const double m = 2.0 * sqrt(q.u * q.u + q.v * q.v) / min_dx;
    const double a = (fabs(q.u) + c) / g.dx + (fabs(q.v) + c) / g.dy;
    if (!isnan(m)) mat = smax(mat, m);
    if (!isnan(a)) aco = smax(aco, a);
  }
  __shared__ double sh[2][RB / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mat = smax(mat, __shfl_xor(mat, off));
    aco = smax(aco, __shfl_xor(aco, off));
  }
  if (lane == 0) {
    sh[0][wave] = mat;
    sh[1][wave] = aco;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = sh[0][0], b = sh[1][0];
    for (int w = 1; w < RB / 64; ++w) {
      a = smax(a, sh[0][w]);
      b = smax(b, sh[1][w]);
    }
    partials[2 * blockIdx.x] = a;
    partials[2 * blockIdx.x + 1] = b;
  }
  (void)ctl;
}

// ----------------------------------------------------------------- host
static int grid_blocks(long long n)
{
  const long long b = (n + RB - 1) / RB;  // one row per thread
  return b < 1 ? 1 : (int)b;
}

static ec_status imex_workspace(ec_ctx* ctx, int m)
{
  ImexWork* w = static_cast<ImexWork*>(ctx->imex);
  if (w && w->m >= m) return EC_OK;
  imex_release(ctx);
  w = new ImexWork();
  ctx->imex = w;
  w->nx = ctx->grid.nx;
  w->ny = ctx->grid.ny;
  w->m = m;
  const size_t n = (size_t)w->nx * w->ny;
  const size_t nxf = (size_t)(w->nx + 1) * w->ny, nyf = (size_t)w->nx * (w->ny + 1);
  w->nparts = grid_blocks(n);
  bool ok = true;
  auto A = [&](void** p, size_t bytes) { ok = ok && hipMalloc(p, bytes) == hipSuccess; };
  A((void**)&w->star, 4 * n * sizeof(double));
  A((void**)&w->p, n * sizeof(double));
  A((void**)&w->x, n * sizeof(double));
  A((void**)&w->mom, 2 * n * sizeof(double));
  A((void**)&w->hx, nxf * sizeof(double));
  A((void**)&w->hy, nyf * sizeof(double));
  A((void**)&w->fx, nxf * sizeof(double));
  A((void**)&w->fy, nyf * sizeof(double));
  A((void**)&w->hx32, nxf * sizeof(float));
  A((void**)&w->hy32, nyf * sizeof(float));
  A((void**)&w->b, n * sizeof(double));
  A((void**)&w->r, n * sizeof(double));
  A((void**)&w->V, (size_t)(m + 1) * n * sizeof(double));
  A((void**)&w->V32, (size_t)(m + 1) * n * sizeof(float));
  A((void**)&w->w, 2 * n * sizeof(double));   // two halves: Arnoldi ping-pong
  A((void**)&w->w32, 2 * n * sizeof(float));
  A((void**)&w->dA64, sizeof(OpArgs<double>));
  A((void**)&w->dA32, sizeof(OpArgs<float>));
  A((void**)&w->R, (size_t)(m + 1) * m * sizeof(double));
  A((void**)&w->g, (size_t)2 * (m + 1) * sizeof(double));
  A((void**)&w->cs, (size_t)m * sizeof(double));
  A((void**)&w->sn, (size_t)m * sizeof(double));
  A((void**)&w->y, (size_t)m * sizeof(double));
  A((void**)&w->partials, (size_t)w->nparts * MAXK * sizeof(double));
  A((void**)&w->npart, (size_t)w->nparts * sizeof(double));
  w->otiles = onesync_tiles(w->nx, w->ny);
  A((void**)&w->opart, (size_t)2 * w->otiles * KP * sizeof(double));  // two halves
  A((void**)&w->otot, (size_t)KP * sizeof(double));
  A((void**)&w->hraw, (size_t)(m + 1) * m * sizeof(double));
  A((void**)&w->inv_nu, (size_t)(m + 1) * sizeof(double));
  A((void**)&w->ctl, sizeof(GmCtl));
  ok = ok && hipHostMalloc((void**)&w->ctl_host, sizeof(GmCtl)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&w->hA64, sizeof(OpArgs<double>)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&w->hA32, sizeof(OpArgs<float>)) == hipSuccess;
  ok = ok && hipStreamCreateWithFlags(&w->gs, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&w->ev, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    imex_release(ctx);
    return EC_ERR_OUT_OF_MEMORY;
  }
  std::memset(w->ctl_host, 0, sizeof(GmCtl));
  if (hipMemcpy(w->ctl, w->ctl_host, sizeof(GmCtl), hipMemcpyHostToDevice) != hipSuccess) {
    imex_release(ctx);
    return EC_ERR_HIP;
  }
  return EC_OK;
}

struct HostCtl {
  ec_ctx* ctx;
  ImexWork* w;
  ec_status pull()
  {
    EC_HIP_TRY(ctx, hipMemcpyAsync(w->ctl_host, w->ctl, sizeof(GmCtl), hipMemcpyDeviceToHost,
                                   ctx->stream));
    EC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return EC_OK;
  }
  ec_status push()
  {
    EC_HIP_TRY(ctx, hipMemcpyAsync(w->ctl, w->ctl_host, sizeof(GmCtl), hipMemcpyHostToDevice,
                                   ctx->stream));
    return EC_OK;
  }
};

// Enqueues one restart cycle of GMRES(m) for precision T on stream st:
// v0 = r / beta, up to ctl->mlim Arnoldi iterations (early exit through
// ctl->done), the back-solve, x += V y (x in FP64), then the true residual
// r = b - A x and its norms, and finally the control block back to the host.
// Every value that changes between cycles or solves is read from device
// memory (ctl, the operator block), so the sequence is captured once into a
// hipGraph and replayed.  two_pass: classical Gram-Schmidt with
// re-orthogonalisation (CGS2, what the FP64 form uses); otherwise one pass
// (CGS1).  The mixed form's FP32 inner cycles use the one-reduction kernels
// instead (enqueue_outer_step).
template <typename T>
static void enqueue_cycle(ImexWork* w, hipStream_t st, const OpArgs<T>* dA, T* V, T* wv, int two_pass)
{
  const long long n = (long long)w->nx * w->ny;
  const int nb = grid_blocks(n);
  const int m = w->m;
  (void)hipMemcpyAsync(w->ctl, w->ctl_host, offsetof(GmCtl, dots1), hipMemcpyHostToDevice, st);
  hipLaunchKernelGGL((cycle_start_kernel<T>), dim3(nb), dim3(RB), 0, st, w->r, V, n, w->ctl, w->g);
  for (int j = 0; j < m; ++j) {
    T* wj = (j & 1) ? wv + n : wv;
    const T* wp = (j & 1) ? wv : wv + n;
    hipLaunchKernelGGL((arnoldi_kernel<T>), dim3(nb), dim3(RB), 0, st, dA, V, j, wp, wj, w->partials,
                       w->ctl, w->npart, nb, w->R, w->g, w->cs, w->sn, m, two_pass);
    hipLaunchKernelGGL(finalize_dots_kernel, dim3(j + 1), dim3(RB), 0, st, w->partials, nb, MAXK,
                       w->ctl->dots1, w->ctl);
    if (two_pass) {
      hipLaunchKernelGGL((update_dot_kernel<T>), dim3(nb), dim3(RB), 0, st, V, n, j + 1, wj,
                         w->partials, w->ctl);
      hipLaunchKernelGGL(finalize_dots_kernel, dim3(j + 1), dim3(RB), 0, st, w->partials, nb, MAXK,
                         w->ctl->dots2, w->ctl);
    }
    hipLaunchKernelGGL((update_norm_kernel<T>), dim3(nb), dim3(RB), 0, st, V, n, j + 1, wj,
                       w->npart, w->ctl, two_pass ? 2 : 1);
  }
  // the last column's Givens step (the others ran inside the next Arnoldi kernel)
  hipLaunchKernelGGL(norm_givens_kernel, dim3(1), dim3(RB), 0, st, w->npart, nb, w->R, w->g, w->cs,
                     w->sn, m - 1, m, w->ctl, two_pass);
  hipLaunchKernelGGL(backsolve_kernel, dim3(1), dim3(64), 0, st, w->R, w->g, w->y, m, w->ctl,
                     (const double*)nullptr);
  hipLaunchKernelGGL((xupdate_kernel<T, double>), dim3(nb), dim3(RB), 0, st, V, n, w->y, w->x, 1.0,
                     w->ctl, 0);
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(RB), 0, st, w->dA64, w->x, w->b, w->r,
                     w->partials, w->ctl);
  hipLaunchKernelGGL(finalize_pair_kernel, dim3(1), dim3(RB), 0, st, w->partials, nb, 0,
                     &w->ctl->rinf, &w->ctl->r2);
  (void)hipMemcpyAsync(w->ctl_host, w->ctl, sizeof(GmCtl), hipMemcpyDeviceToHost, st);
}

// The one-reduction kernel for inner step k: the smallest instantiation that
// holds the k basis values in registers (FP64 holds at most 16, which keeps
// a 1024-thread workgroup within its 128 registers per lane; the rest are
// read from memory).
template <typename T>
using OnesyncFn = void (*)(const OpArgs<T>*, T*, T*, int, int, const double*, double*, int, GmCtl*,
                           double*, double*, double*, double*, double*, double*);

template <typename T>
static OnesyncFn<T> onesync_for(int k, bool summed)
{
  if constexpr (sizeof(T) == 4) {
    return summed ? (k <= 8    ? onesync_kernel<float, 8, true>
                     : k <= 16 ? onesync_kernel<float, 16, true>
                     : k <= 24 ? onesync_kernel<float, 24, true>
                               : onesync_kernel<float, 32, true>)
                  : (k <= 8    ? onesync_kernel<float, 8, false>
                     : k <= 16 ? onesync_kernel<float, 16, false>
                     : k <= 24 ? onesync_kernel<float, 24, false>
                               : onesync_kernel<float, 32, false>);
  } else {
    return summed ? (k <= 8 ? onesync_kernel<double, 8, true> : onesync_kernel<double, 16, true>)
                  : (k <= 8 ? onesync_kernel<double, 8, false> : onesync_kernel<double, 16, false>);
  }
}

// Outer step j of the flexible form: the inner GMRES(m) solve of A z_j = v_j
// in precision T (FP32 for the mixed form, FP64 for its FP64 comparator;
// one-reduction kernels, unit right-hand side, relative tolerance inner_thr),
// z_j in FP64, then w = A z_j orthogonalised against V_0..V_j in FP64 (CGS2),
// column j's Givens step and V_{j+1} = w / |w|.  Every kernel returns at once
// when the outer cycle is done.
template <typename T>
static void enqueue_outer_step(ImexWork* w, hipStream_t st, int j)
{
  const long long n = (long long)w->nx * w->ny;
  const int nb = grid_blocks(n);
  const int m = w->m;
  double* vj = w->Vo + (long long)j * n;
  double* zj = w->Zo + (long long)j * n;
  T* Ub;
  T* Zb;
  const OpArgs<T>* dA;
  if constexpr (sizeof(T) == 4) {
    Ub = w->V32;
    Zb = w->w32;
    dA = w->dA32;
  } else {
    Ub = w->V;
    Zb = w->w;
    dA = w->dA64;
  }
  hipLaunchKernelGGL((inner_start_kernel<T>), dim3(nb), dim3(RB), 0, st, vj, Ub, n, w->ctl, w->ctlo,
                     w->g);
  for (int k = 0; k <= m; ++k) {
    // Up to ONE_SUM_TILES tiles every workgroup sums the previous kernel's
    // tile partials itself; beyond, that redundant sum grows as tiles^2, and
    // a fixed-order finalize after each kernel sums them once instead.
    const bool summed = w->otiles > ONE_SUM_TILES;
    auto kern = onesync_for<T>(k, summed);
    // the partials ping-pong like z: a workgroup that finishes early must not
    // overwrite what a later workgroup of the same launch still has to read
    const size_t ps = (size_t)w->otiles * KP;
    double* pout = w->opart + (k & 1) * ps;
    const double* pin = summed ? w->otot : w->opart + ((k + 1) & 1) * ps;
    hipLaunchKernelGGL(kern, dim3(w->otiles), dim3(ONT), 0, st, dA, Ub, Zb, k, m, pin, pout,
                       w->otiles, w->ctl, w->hraw, w->inv_nu, w->R, w->g, w->cs, w->sn);
    if (summed && k < m) {
      hipLaunchKernelGGL(finalize_dots_kernel, dim3(k + 2), dim3(RB), 0, st, (const double*)pout,
                         w->otiles, KP, w->otot, (const GmCtl*)w->ctl);
    }
  }
  hipLaunchKernelGGL(onesync_close_kernel, dim3(1), dim3(ONT), 0, st,
                     w->opart + (m & 1) * (size_t)w->otiles * KP, w->otiles, m, w->ctl,
                     w->hraw, w->R, w->g, w->cs, w->sn, w->y, (const double*)w->inv_nu);
  hipLaunchKernelGGL((xupdate_kernel<T, double>), dim3(nb), dim3(RB), 0, st, Ub, n, w->y, zj, 1.0,
                     w->ctl, 1);
  hipLaunchKernelGGL(outer_arnoldi_kernel, dim3(nb), dim3(RB), 0, st, w->dA64, w->Vo, zj, j, w->w,
                     w->partials, w->ctlo);
  hipLaunchKernelGGL(finalize_dots_kernel, dim3(j + 1), dim3(RB), 0, st, w->partials, nb, MAXK,
                     w->ctlo->dots1, w->ctlo);
  hipLaunchKernelGGL((update_dot_kernel<double>), dim3(nb), dim3(RB), 0, st, w->Vo, n, j + 1, w->w,
                     w->partials, w->ctlo);
  hipLaunchKernelGGL(finalize_dots_kernel, dim3(j + 1), dim3(RB), 0, st, w->partials, nb, MAXK,
                     w->ctlo->dots2, w->ctlo);
  hipLaunchKernelGGL((update_norm_kernel<double>), dim3(nb), dim3(RB), 0, st, w->Vo, n, j + 1, w->w,
                     w->npart, w->ctlo, 2);
  hipLaunchKernelGGL(norm_givens_kernel, dim3(1), dim3(RB), 0, st, w->npart, nb, w->Ro, w->go, w->cso,
                     w->sno, j, MO, w->ctlo, 1);
  hipLaunchKernelGGL(outer_next_kernel, dim3(nb), dim3(RB), 0, st, w->w, w->Vo + (long long)(j + 1) * n,
                     n, w->ctlo, w->ctl);
}

// Closes an outer cycle: y = R^{-1} g, x += Z y (FP64), the true residual
// and its norms, and the outer control block back to the host.
static void enqueue_outer_close(ImexWork* w, hipStream_t st)
{
  const long long n = (long long)w->nx * w->ny;
  const int nb = grid_blocks(n);
  hipLaunchKernelGGL(backsolve_kernel, dim3(1), dim3(64), 0, st, w->Ro, w->go, w->yo, MO, w->ctlo,
                     (const double*)nullptr);
  hipLaunchKernelGGL((xupdate_kernel<double, double>), dim3(nb), dim3(RB), 0, st, w->Zo, n, w->yo, w->x,
                     1.0, w->ctlo, 0);
  hipLaunchKernelGGL(residual_kernel, dim3(nb), dim3(RB), 0, st, w->dA64, w->x, w->b, w->r,
                     w->partials, w->ctlo);
  hipLaunchKernelGGL(finalize_pair_kernel, dim3(1), dim3(RB), 0, st, w->partials, nb, 0,
                     &w->ctlo->rinf, &w->ctlo->r2);
  (void)hipMemcpyAsync(w->ctlo_host, w->ctlo, sizeof(GmCtl), hipMemcpyDeviceToHost, st);
}

// Captures one restart cycle (enqueue(stream)) into a hipGraph, once.
template <typename F>
static ec_status cycle_graph(ec_ctx* ctx, ImexWork* w, hipGraphExec_t* exec, const F& enqueue)
{
  if (*exec) return EC_OK;
  hipGraph_t g = nullptr;
  EC_HIP_TRY(ctx, hipStreamBeginCapture(w->gs, hipStreamCaptureModeThreadLocal));
  enqueue(w->gs);
  const hipError_t e = hipStreamEndCapture(w->gs, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(ctx, e, "hipStreamEndCapture (GMRES cycle)");
  }
  const hipError_t e2 = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e2 != hipSuccess) {
    *exec = nullptr;
    return hip_fail(ctx, e2, "hipGraphInstantiate (GMRES cycle)");
  }
  return EC_OK;
}

struct SolveResult {
  int iterations = 0;
  int refinements = 0;
  bool breakdown = false;
  bool bad = false;
  double rinf = INFINITY, rel = INFINITY;
};

// r = b - A x with |r|_inf and |r|_2^2, outside any cycle (the initial guess).
static ec_status true_residual(ec_ctx* ctx, ImexWork* w, hipStream_t st, double* rinf, double* r2)
{
  const long long n = (long long)w->nx * w->ny;
  hipLaunchKernelGGL(residual_kernel, dim3(grid_blocks(n)), dim3(RB), 0, st, w->dA64, w->x, w->b,
                     w->r, w->partials, w->ctl);
  hipLaunchKernelGGL(finalize_pair_kernel, dim3(1), dim3(RB), 0, st, w->partials, grid_blocks(n), 0,
                     &w->ctl->rinf, &w->ctl->r2);
  EC_HIP_TRY(ctx, hipGetLastError());
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->ctl_host, w->ctl, sizeof(GmCtl), hipMemcpyDeviceToHost, st));
  EC_HIP_TRY(ctx, hipStreamSynchronize(st));
  *rinf = w->ctl_host->rinf;
  *r2 = w->ctl_host->r2;
  return EC_OK;
}

// The flexible form's FP64 outer buffers and step graphs for one inner
// precision (first use only).
static ec_status outer_setup(ec_ctx* ctx, ImexWork* w, bool fp64_inner)
{
  std::vector<hipGraphExec_t>& execs = fp64_inner ? w->exec_outer64 : w->exec_outer;
  if (w->Vo && (int)execs.size() == MO + 1) return EC_OK;
  const size_t n = (size_t)w->nx * w->ny;
  if (!w->Vo) {
    bool ok = true;
    auto A = [&](void** p, size_t bytes) { ok = ok && hipMalloc(p, bytes) == hipSuccess; };
    A((void**)&w->Vo, (size_t)(MO + 1) * n * sizeof(double));
    A((void**)&w->Zo, (size_t)MO * n * sizeof(double));
    A((void**)&w->Ro, (size_t)(MO + 1) * MO * sizeof(double));
    A((void**)&w->go, (size_t)2 * (MO + 1) * sizeof(double));
    A((void**)&w->cso, (size_t)MO * sizeof(double));
    A((void**)&w->sno, (size_t)MO * sizeof(double));
    A((void**)&w->yo, (size_t)MO * sizeof(double));
    A((void**)&w->ctlo, sizeof(GmCtl));
    ok = ok && hipHostMalloc((void**)&w->ctlo_host, sizeof(GmCtl)) == hipSuccess;
    if (!ok) return hip_fail(ctx, hipErrorOutOfMemory, "mixed-precision outer workspace");
    std::memset(w->ctlo_host, 0, sizeof(GmCtl));
    EC_HIP_TRY(ctx, hipMemcpy(w->ctlo, w->ctlo_host, sizeof(GmCtl), hipMemcpyHostToDevice));
  }
  execs.assign(MO + 1, nullptr);
  for (int j = 0; j <= MO; ++j) {
    const ec_status st = cycle_graph(ctx, w, &execs[j], [&](hipStream_t gs) {
      if (j < MO) {
        if (fp64_inner) enqueue_outer_step<double>(w, gs, j);
        else enqueue_outer_step<float>(w, gs, j);
      } else {
        enqueue_outer_close(w, gs);
      }
    });
    if (st != EC_OK) return st;
  }
  return EC_OK;
}

// Mixed form (FGMRES-IR): outer FGMRES(MO) cycles in FP64 from the current
// x, each outer step preconditioned by an FP32 GMRES(m) solve, until the
// reference's criterion holds.  Iterations are the FP32 inner iterations,
// refinements the outer steps.
static ec_status solve_mixed(ec_ctx* ctx, ImexWork* w, double target, double rhs_norm,
                             const ec_imex_opts& o, double rinf, double r2, SolveResult* out)
{
  const bool fp64_inner = o.precision == EC_IMEX_PRECISION_FP64_FLEXIBLE;
  ec_status st = outer_setup(ctx, w, fp64_inner);
  if (st != EC_OK) return st;
  const std::vector<hipGraphExec_t>& execs = fp64_inner ? w->exec_outer64 : w->exec_outer;
  const long long n = (long long)w->nx * w->ny;
  GmCtl& h = *w->ctlo_host;
  int total = 0;
  while (rinf > target && total < o.max_iter) {
    const double beta = std::sqrt(r2);
    if (!(beta > 0.0)) break;
    h.done = 0;
    h.kk = 0;
    h.bad = 0;
    h.breakdown = 0;
    h.mlim = MO;
    h.inner_its = 0;
    h.beta = beta;
    h.thr = target * (beta / rinf);  // as in the FP64 form
    h.inner_thr = o.inner_reduction;
    h.inner_mlim = std::min(o.restart, w->m);
    h.inner_budget = o.max_iter - total;
    EC_HIP_TRY(ctx, hipMemcpyAsync(w->ctlo, w->ctlo_host, offsetof(GmCtl, dots1),
                                   hipMemcpyHostToDevice, w->gs));
    hipLaunchKernelGGL((cycle_start_kernel<double>), dim3(grid_blocks(n)), dim3(RB), 0, w->gs, w->r,
                       w->Vo, n, w->ctlo, w->go);
    EC_HIP_TRY(ctx, hipGetLastError());
    for (int j = 0; j <= MO; ++j) EC_HIP_TRY(ctx, hipGraphLaunch(execs[j], w->gs));
    EC_HIP_TRY(ctx, hipStreamSynchronize(w->gs));
    const int kk = h.kk;
    total += h.inner_its;
    out->refinements += kk;
    out->breakdown = h.breakdown != 0;
    rinf = h.rinf;
    r2 = h.r2;
    out->rinf = rinf;
    out->rel = rinf / rhs_norm;
    if (h.bad || !std::isfinite(rinf)) {
      out->bad = true;
      break;
    }
    if (out->breakdown && rinf > target) break;
    if (kk == 0) break;
  }
  out->iterations = total;
  return EC_OK;
}

// Solves A x = b (x holds the initial guess) to the reference's criterion
// |r|_inf <= tol or |r|_inf / max(|b|_inf,1) <= tol
// (imex/imex_pressure_helpers.hpp:105-125).
static ec_status solve_pressure(ec_ctx* ctx, ImexWork* w, const OpArgs<double>& A64,
                                const OpArgs<float>& A32, double rhs_norm, const ec_imex_opts& o,
                                SolveResult* out)
{
  const int m = o.restart;
  const double target = o.tol * smax(1.0, rhs_norm);  // abs_ok or rel_ok
  const int max_it = o.max_iter;
  int total = 0;
  (void)A64;
  (void)A32;
  // The private stream starts after the work already queued on the context's
  // stream (the operator blocks are already in device memory).
  EC_HIP_TRY(ctx, hipEventRecord(w->ev, ctx->stream));
  EC_HIP_TRY(ctx, hipStreamWaitEvent(w->gs, w->ev, 0));
  double rinf, r2;
  ec_status st = true_residual(ctx, w, w->gs, &rinf, &r2);
  if (st != EC_OK) return st;
  auto record = [&]() {
    out->rinf = rinf;
    out->rel = rinf / rhs_norm;
  };
  record();
  if (!std::isfinite(rinf)) {
    out->bad = true;
    return EC_OK;
  }
  if (o.precision == EC_IMEX_PRECISION_MIXED || o.precision == EC_IMEX_PRECISION_FP64_FLEXIBLE) {
    return solve_mixed(ctx, w, target, rhs_norm, o, rinf, r2, out);
  }
  st = cycle_graph(ctx, w, &w->exec64, [&](hipStream_t gs) {
    enqueue_cycle<double>(w, gs, w->dA64, w->V, w->w, 1);
  });
  if (st != EC_OK) return st;
  GmCtl& h = *w->ctl_host;
  while (rinf > target && total < max_it) {
    const double beta = std::sqrt(r2);
    if (!(beta > 0.0)) break;
    h.done = 0;
    h.kk = 0;
    h.bad = 0;
    h.breakdown = 0;
    h.mlim = std::min(m, max_it - total);
    h.beta = beta;
    h.counter = 0;
    // 2-norm estimate threshold scaled by the |r|_2 / |r|_inf ratio of the
    // cycle's start residual; the true residual decides convergence.
    h.thr = target * (beta / rinf);
    EC_HIP_TRY(ctx, hipGraphLaunch(w->exec64, w->gs));
    EC_HIP_TRY(ctx, hipStreamSynchronize(w->gs));
    const int kk = h.kk;
    const bool bad = h.bad != 0, brk = h.breakdown != 0;
    total += kk;
    out->breakdown = brk;
    rinf = h.rinf;
    r2 = h.r2;
    record();
    if (bad || !std::isfinite(rinf)) {
      out->bad = true;
      break;
    }
    if (brk && rinf > target) break;
    if (kk == 0) break;
  }
  out->iterations = total;
  return EC_OK;
}

ec_status imex_buffers(ec_ctx* ctx, int restart, ImexBuffers* out)
{
  const ec_status st = imex_workspace(ctx, restart);
  if (st != EC_OK) return st;
  ImexWork* w = static_cast<ImexWork*>(ctx->imex);
  *out = ImexBuffers{w->nx, w->ny, w->star, w->p, w->x, w->mom, w->hx, w->hy, w->hx32, w->hy32, w->b};
  return EC_OK;
}

ec_status imex_solve_pressure(ec_ctx* ctx, const int bc[4], double cx, double cy, double diag,
                              const ec_imex_opts& o, PressureSolve* out)
{
  ec_status st = imex_workspace(ctx, o.restart);
  if (st != EC_OK) return st;
  ImexWork* w = static_cast<ImexWork*>(ctx->imex);
  const long long n = (long long)w->nx * w->ny;
  ImexGrid G{};
  G.nx = w->nx;
  G.ny = w->ny;
  for (int s = 0; s < 4; ++s) G.bc[s] = bc[s];
  G.gamma = ctx->grid.gamma;
  G.dx = ctx->grid.dx;
  G.dy = ctx->grid.dy;
  G.inv_dx = ctx->grid.inv_dx;
  G.inv_dy = ctx->grid.inv_dy;
  const OpArgs<double> A64{G, w->hx, w->hy, cx, cy, diag};
  const OpArgs<float> A32{G, w->hx32, w->hy32, (float)cx, (float)cy, (float)diag};
  *w->hA64 = A64;
  *w->hA32 = A32;
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->dA64, w->hA64, sizeof(A64), hipMemcpyHostToDevice, ctx->stream));
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->dA32, w->hA32, sizeof(A32), hipMemcpyHostToDevice, ctx->stream));
  // |b|_inf as the residual of x = 0, then the initial guess
  EC_HIP_TRY(ctx, hipMemsetAsync(w->x, 0, n * sizeof(double), ctx->stream));
  double binf, b2;
  if ((st = true_residual(ctx, w, ctx->stream, &binf, &b2)) != EC_OK) return st;
  *out = PressureSolve{};
  out->rhs_norm = smax(binf, 1.0);
  if (!std::isfinite(binf)) {
    out->bad = 1;
    out->rinf = out->rel = INFINITY;
    return EC_OK;
  }
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->x, w->p, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  SolveResult sr;
  if ((st = solve_pressure(ctx, w, A64, A32, out->rhs_norm, o, &sr)) != EC_OK) return st;
  out->iterations = sr.iterations;
  out->refinements = sr.refinements;
  out->bad = sr.bad ? 1 : 0;
  out->breakdown = sr.breakdown ? 1 : 0;
  out->rinf = sr.rinf;
  out->rel = sr.rel;
  return EC_OK;
}

}  // namespace ec

// ------------------------------------------------------------------ C-ABI
using namespace ec;

// The step's ghost rule.  Exact-Dirichlet sides read as outflow on a
// non-periodic axis and as periodic on a periodic one: the source-map step
// fills the state's ghosts with the periodic images and clamps the physical
// sides before the star update for every problem but the shock bubble
// (imex/imex_source_map_step.hpp:405-409, problems/boundary_conditions.hpp:
// 2-25), and its host grid reads valid cells only, wrapped or clamped
// (imex/imex_source_map_helpers.hpp:79-95,117-135), so the exact ghosts it
// writes after the star update (imex_source_map_helpers.hpp:67) and at the
// end of the step (imex_source_map_step.hpp:275) are never read.
static GridParams imex_ghost_rule(const ec_ctx* ctx)
{
  GridParams g = ctx->grid;
  for (int s = 0; s < 4; ++s)
    if (g.bc[s] == BC_EXACT) g.bc[s] = ctx->exact_plain_bc[s];
  g.exact = nullptr;
  g.time_ptr = nullptr;
  return g;
}

static ImexGrid imex_grid(const GridParams& g, double dt)
{
  ImexGrid G{};
  G.nx = g.nx;
  G.ny = g.ny;
  for (int s = 0; s < 4; ++s) {
    G.bc[s] = g.bc[s];
    for (int c = 0; c < 4; ++c) G.fixed[s][c] = g.fixed[s][c];
    const Prim q = to_prim(Cons{g.fixed[s][0], g.fixed[s][1], g.fixed[s][2], g.fixed[s][3]}, g.gamma);
    G.fixed_p[s] = g.bc[s] == BC_FIXED ? q.p : 0.0;
  }
  G.gamma = g.gamma;
  G.dx = g.dx;
  G.dy = g.dy;
  G.inv_dx = g.inv_dx;
  G.inv_dy = g.inv_dy;
  G.dt = dt;
  return G;
}

extern "C" ec_status ec_imex_default_opts(ec_imex_opts* o)
{
  if (!o) return EC_ERR_INVALID_ARGUMENT;
  // core/run_config.hpp:2-55 (imex_picard_iterations 2, solver tol 1e-10,
  // max_iter 100, material dissipation)
  o->picard_iterations = 2;
  o->tol = 1.0e-10;
  o->max_iter = 100;
  o->restart = 30;
  o->precision = EC_IMEX_PRECISION_FP64;
  o->predictor_dissipation = 0;
  o->inner_reduction = 1.0e-4;
  return EC_OK;
}

extern "C" ec_status ec_imex_step(ec_ctx* ctx, double* U, const ec_layout* lay, double dt,
                                  const ec_imex_opts* opts, ec_imex_diag* diag, double* p_out)
{
  if (!ctx || !U || !lay || !opts || !diag) return EC_ERR_INVALID_ARGUMENT;
  if (lay->row_stride < ctx->grid.nx || lay->comp_stride < lay->row_stride * ctx->grid.ny) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  if (opts->picard_iterations < 1 || opts->restart < 1 || opts->restart > MAXK ||
      !(opts->tol > 0.0) || opts->max_iter < 0 ||
      (opts->precision != EC_IMEX_PRECISION_FP64 && opts->precision != EC_IMEX_PRECISION_MIXED &&
       opts->precision != EC_IMEX_PRECISION_FP64_FLEXIBLE) ||
      !(dt > 0.0) || !std::isfinite(dt)) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  ec_status st = imex_workspace(ctx, opts->restart);
  if (st != EC_OK) return st;
  ImexWork* w = static_cast<ImexWork*>(ctx->imex);
  std::memset(diag, 0, sizeof *diag);
  diag->status = EC_IMEX_NOT_RUN;
  // core/imex_diagnostics.hpp:111-163 defaults, kept by the early returns
  diag->solved_pressure_min = diag->solved_pressure_max = NAN;
  diag->trial_density_min = diag->trial_pressure_min = diag->trial_internal_energy_min = NAN;
  diag->eq78_divergence_linf = NAN;
  const long long n = (long long)w->nx * w->ny;
  const int nb = grid_blocks(n);
  const long long nf = (long long)(w->nx + 1) * w->ny + (long long)w->nx * (w->ny + 1);
  const int nbf = grid_blocks(nf);
  GmCtl& h = *w->ctl_host;
  std::memset(&h, 0, sizeof(GmCtl));
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->ctl, &h, sizeof(GmCtl), hipMemcpyHostToDevice, ctx->stream));
  HostCtl hc{ctx, w};
  const GridParams g = imex_ghost_rule(ctx);
  const ImexGrid G = imex_grid(g, dt);

  // star state and initial fields
  hipLaunchKernelGGL(imex_star_kernel, dim3(nb), dim3(RB), 0, ctx->stream, U, lay->row_stride,
                     lay->comp_stride, g, dt, opts->predictor_dissipation, w->star);
  hipLaunchKernelGGL(imex_init_kernel, dim3(nb), dim3(RB), 0, ctx->stream, U, lay->row_stride,
                     lay->comp_stride, ctx->grid.gamma, w->nx, w->ny, w->star, w->p, w->mom, w->partials,
                     w->ctl);
  hipLaunchKernelGGL(finalize_pair_kernel, dim3(1), dim3(RB), 0, ctx->stream, w->partials, nb, 1,
                     &w->ctl->star_rho_neg_max, &w->ctl->star_pad);
  EC_HIP_TRY(ctx, hipGetLastError());
  if ((st = hc.pull()) != EC_OK) return st;
  // imex/imex_source_map_step.hpp:436-444: the finite star densities' minimum
  diag->trial_density_min = -h.star_rho_neg_max;
  diag->density_floor_count = h.floor_count;
  if (h.floor_count > 0) {
    diag->status = EC_IMEX_DENSITY_FLOOR;
    return EC_OK;
  }

  OpArgs<double> A64{G, w->hx, w->hy, 0.25 * dt * dt * G.inv_dx * G.inv_dx,
                     0.25 * dt * dt * G.inv_dy * G.inv_dy, 1.0 / (G.gamma - 1.0)};
  OpArgs<float> A32{G, w->hx32, w->hy32, (float)A64.cx, (float)A64.cy, (float)A64.diag};
  *w->hA64 = A64;
  *w->hA32 = A32;
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->dA64, w->hA64, sizeof(A64), hipMemcpyHostToDevice, ctx->stream));
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->dA32, w->hA32, sizeof(A32), hipMemcpyHostToDevice, ctx->stream));
  const bool mixed = opts->precision == EC_IMEX_PRECISION_MIXED;
  for (int it = 0; it < opts->picard_iterations; ++it) {
    hipLaunchKernelGGL(imex_faces_kernel, dim3(nbf), dim3(RB), 0, ctx->stream, G, w->star, w->mom,
                       w->p, w->hx, w->hy, mixed ? w->hx32 : nullptr, mixed ? w->hy32 : nullptr,
                       nullptr, nullptr, w->ctl);
    hipLaunchKernelGGL(imex_rhs_kernel, dim3(nb), dim3(RB), 0, ctx->stream, G, w->star, w->mom, w->hx,
                       w->hy, w->b, w->ctl);
    // |b|_inf through the residual reducer with x = 0 would need a second
    // buffer; reuse r = b - A*0 = b.
    EC_HIP_TRY(ctx, hipMemsetAsync(w->x, 0, n * sizeof(double), ctx->stream));
    double binf, b2;
    if ((st = true_residual(ctx, w, ctx->stream, &binf, &b2)) != EC_OK) return st;
    if (h.nonfinite > 0 || !std::isfinite(binf)) {
      diag->status = EC_IMEX_ROW_NONFINITE;
      diag->nonfinite_count = h.nonfinite;
      return EC_OK;
    }
    const double rhs_norm = smax(binf, 1.0);
    diag->rhs_norm = rhs_norm;
    EC_HIP_TRY(ctx, hipMemcpyAsync(w->x, w->p, n * sizeof(double), hipMemcpyDeviceToDevice,
                                   ctx->stream));
    SolveResult sr;
    if ((st = solve_pressure(ctx, w, A64, A32, rhs_norm, *opts, &sr)) != EC_OK) return st;
    diag->iterations += sr.iterations;
    diag->last_iterations = sr.iterations;
    diag->refinements = sr.refinements;
    diag->residual_linf = sr.rinf;
    diag->relative_residual_linf = sr.rel;
    const bool abs_ok = sr.rinf <= opts->tol, rel_ok = sr.rel <= opts->tol;
    if (sr.bad || !std::isfinite(sr.rinf)) {
      diag->status = EC_IMEX_RESIDUAL_NONFINITE;
      return EC_OK;
    }
    if (!abs_ok && !rel_ok) {
      diag->status = sr.breakdown ? EC_IMEX_BREAKDOWN : EC_IMEX_NOT_CONVERGED;
      return EC_OK;
    }
    diag->status = abs_ok ? EC_IMEX_CONVERGED_ABSOLUTE : EC_IMEX_CONVERGED_RELATIVE;
    EC_HIP_TRY(ctx, hipMemcpyAsync(w->p, w->x, n * sizeof(double), hipMemcpyDeviceToDevice,
                                   ctx->stream));
    hipLaunchKernelGGL(momentum_refresh_kernel, dim3(nb), dim3(RB), 0, ctx->stream, G, w->star, w->p,
                       w->mom, w->ctl);
    EC_HIP_TRY(ctx, hipGetLastError());
    if ((st = hc.pull()) != EC_OK) return st;
    if (h.mom_bad > 0) {
      diag->status = EC_IMEX_MOMENTUM_NONFINITE;
      diag->nonfinite_count = h.mom_bad;
      return EC_OK;
    }
    ++diag->picard_solves;
  }
  // closeout with the corrected momentum of the final pressure
  if ((st = hc.pull()) != EC_OK) return st;
  h.guard_count = 0;
  h.nonfinite = 0;
  EC_HIP_TRY(ctx, hipMemcpyAsync(w->ctl, &h, sizeof(GmCtl), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(imex_faces_kernel, dim3(nbf), dim3(RB), 0, ctx->stream, G, w->star, w->mom, w->p,
                     w->hx, w->hy, nullptr, nullptr, w->fx, w->fy, w->ctl);
  hipLaunchKernelGGL(closeout_kernel, dim3(nb), dim3(RB), 0, ctx->stream, G, w->star, w->mom, w->p,
                     w->fx, w->fy, U, lay->row_stride, lay->comp_stride, w->partials, w->ctl);
  hipLaunchKernelGGL(finalize_close_kernel, dim3(1), dim3(RB), 0, ctx->stream, w->partials, nb,
                     w->ctl->close);
  hipLaunchKernelGGL(eq78_kernel, dim3(nb), dim3(RB), 0, ctx->stream, G, U, lay->row_stride, lay->comp_stride,
                     w->partials);
  hipLaunchKernelGGL(finalize_pair_kernel, dim3(1), dim3(RB), 0, ctx->stream, w->partials, nb, 0,
                     &w->ctl->eq78, &w->ctl->eq78_count);
  if (p_out) {
    EC_HIP_TRY(ctx, hipMemcpyAsync(p_out, w->p, n * sizeof(double), hipMemcpyDeviceToDevice,
                                   ctx->stream));
  }
  EC_HIP_TRY(ctx, hipGetLastError());
  if ((st = hc.pull()) != EC_OK) return st;
  diag->guard_count = h.guard_count;
  diag->nonfinite_count = h.nonfinite;
  // imex/imex_source_map_step.hpp:229-257: the residual and the divergence
  // read infinite once a cell is not finite; extrema without a finite cell
  // read NaN
  auto finite_or_nan = [](double v) { return std::isfinite(v) ? v : NAN; };
  diag->coupled_residual_linf = h.nonfinite > 0 ? INFINITY : h.close[CL_RES];
  diag->trial_density_min = finite_or_nan(h.close[CL_RHO_MIN]);
  diag->solved_pressure_min = finite_or_nan(h.close[CL_P_MIN]);
  diag->solved_pressure_max = finite_or_nan(h.close[CL_P_MAX]);
  diag->trial_pressure_min = finite_or_nan(h.close[CL_EOS_MIN]);
  diag->trial_internal_energy_min = finite_or_nan(h.close[CL_INTERNAL_MIN]);
  diag->momentum_correction_linf = h.close[CL_MOM_CORR];
  diag->energy_correction_linf = h.close[CL_E_DELTA];
  diag->eq78_divergence_linf = (h.nonfinite > 0 || h.eq78_count == 0.0) ? INFINITY : h.eq78;
  return EC_OK;
}

extern "C" ec_status ec_imex_dt(ec_ctx* ctx, const double* U, const ec_layout* lay, double cfl,
                                double time, double stop_time, int acoustic_startup,
                                double acoustic_cfl_cap, double* dt, double* material_rate_max,
                                double* acoustic_rate_max)
{
  if (!ctx || !U || !lay || !dt) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  ec_status st = imex_workspace(ctx, 30);
  if (st != EC_OK) return st;
  ImexWork* w = static_cast<ImexWork*>(ctx->imex);
  const long long n = (long long)w->nx * w->ny;
  hipLaunchKernelGGL(imex_rates_kernel, dim3(grid_blocks(n)), dim3(RB), 0, ctx->stream, U,
                     lay->row_stride, lay->comp_stride, imex_ghost_rule(ctx), w->partials, w->ctl);
  hipLaunchKernelGGL(finalize_pair_kernel, dim3(1), dim3(RB), 0, ctx->stream, w->partials,
                     grid_blocks(n), 1, &w->ctl->mat_max, &w->ctl->aco_max);
  EC_HIP_TRY(ctx, hipGetLastError());
  HostCtl hc{ctx, w};
  if ((st = hc.pull()) != EC_OK) return st;
  const double mat = w->ctl_host->mat_max, aco = w->ctl_host->aco_max;
  // diagnostic_metrics_timestep.hpp:310-325
  double rate = mat;
  if (acoustic_startup != 0 && mat <= 1.0e-12 && aco > 0.0) rate = aco;
  double d = stop_time - time;
  if (rate > 0.0) d = cfl / rate;
  if (acoustic_cfl_cap > 0.0 && aco > 0.0) {
    const double lim = acoustic_cfl_cap / aco;
    if (lim < d) d = lim;
  }
  if (!std::isfinite(d) || d <= 0.0) d = stop_time - time;
  *dt = smin(d, stop_time - time);
  if (material_rate_max) *material_rate_max = mat;
  if (acoustic_rate_max) *acoustic_rate_max = aco;
  return EC_OK;
}
