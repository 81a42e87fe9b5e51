// ec_explicit.hip -- fused explicit O2 step and the CFL / diagnostics
// reduction for gfx950 (FP64 stencils, no MFMA).
//
// One workgroup owns a TX x TY tile of cells and stages, in LDS:
//   A. the tile plus a 2-cell halo as primitive variables, read through the
//      per-side ghost rule (the reference's separate ghost fill is fused in);
//   B. the limited primitive slopes of every cell whose face states are
//      needed (tile plus a 1-cell ring);
//   C. for HLLC-P, the pressure ratio of every face the 5-face sensor reads;
//   D. every face flux of the tile exactly once (the reference evaluates each
//      face twice and each reconstruction up to 6 times), kept in registers
//      and then written over the dead stage-A..C storage;
//   E. the conservative update, written to U_out, with an epilogue on U_out:
//      max CFL spectral rate, min rho, min p, non-finite and degenerate-star
//      counts, reduced per wavefront with cross-lane shuffles, per workgroup
//      in LDS, and across workgroups by the last workgroup to arrive.
// The last workgroup also advances the device-resident run control (time,
// step, status), so consecutive steps need no host round trip.
//
// Workgroups are numbered so that the 8 XCDs each receive a contiguous band
// of tile rows: neighbouring tiles then share halo rows through one L2.

#include <type_traits>

#include "ec_context.hpp"
#include "ec_ghost.hpp"
#include "ec_state.hpp"

namespace ec {

struct StepArgs {
  const double* __restrict__ uin;
  double* __restrict__ uout;
  long long rs, cs;
  GridParams g;
  double dt;        // used when device_dt == 0
  int device_dt;    // 1: dt from ctrl (device-resident loop)
  DevCtrl* ctrl;
  Partial* partials;
  double* dt_log;   // optional: dt of every device-loop step, indexed by step
  int tiles_x, nblocks;
};

// compute_euler_dt's closing arithmetic (diagnostic_metrics_timestep.hpp:252-259).
__host__ __device__ inline double euler_dt(double max_rate, double cfl, double time, double stop)
{
  double dt = stop - time;
  if (max_rate > 0.0) dt = cfl / max_rate;
  if (!isfinite(dt) || dt <= 0.0) dt = stop - time;
  return smin(dt, stop - time);
}

// Logical tile of workgroup b with the XCD-banded numbering (workgroups are
// dealt to the 8 XCDs round robin: XCD x receives b = x, x+8, x+16, ...).
__device__ __forceinline__ int xcd_banded(int b, int nb)
{
  const int per = nb >> 3, rem = nb & 7;
  const int xcd = b & 7, k = b >> 3;
  return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

struct Red {
  double rate, rmin, pmin;
  int bad, deg;
};

__device__ __forceinline__ Red red_identity()
{
  return Red{-INFINITY, INFINITY, INFINITY, 0, 0};
}

__device__ __forceinline__ void red_combine(Red& a, const Red& b)
{
  a.rate = smax(a.rate, b.rate);
  a.rmin = smin(a.rmin, b.rmin);
  a.pmin = smin(a.pmin, b.pmin);
  a.bad += b.bad;
  a.deg += b.deg;
}

// Diagnostics and CFL rate of one cell (diagnostic_metrics_timestep.hpp:30-38
// and :240-245).  A NaN rate is dropped, as the reference's sequential
// max(running, value) drops it.
__device__ __forceinline__ void red_cell(Red& r, const Cons& u, const GridParams& g)
{
  const Prim q = to_prim(u, g.gamma);
  const bool bad = !isfinite(q.r) || !isfinite(q.u) || !isfinite(q.v) || !isfinite(q.p) ||
                   !isfinite(u.E);
  r.rmin = smin(r.rmin, bad ? -INFINITY : q.r);
  r.pmin = smin(r.pmin, bad ? -INFINITY : q.p);
  r.bad += bad ? 1 : 0;
  const double c = sound_speed(q, g.gamma);
  const double rate = (fabs(q.u) + c) / g.dx + (fabs(q.v) + c) / g.dy;
  if (!isnan(rate)) r.rate = smax(r.rate, rate);
}

__device__ __forceinline__ Red wave_reduce(Red r)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    Red o;
    o.rate = __shfl_xor(r.rate, off);
    o.rmin = __shfl_xor(r.rmin, off);
    o.pmin = __shfl_xor(r.pmin, off);
    o.bad = __shfl_xor(r.bad, off);
    o.deg = __shfl_xor(r.deg, off);
    red_combine(r, o);
  }
  return r;
}

// Workgroup reduction; result valid in thread 0.
template <int NT>
__device__ __forceinline__ Red block_reduce(Red r, Red* scratch)
{
  r = wave_reduce(r);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) scratch[wave] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < NT / 64; ++w) red_combine(r, scratch[w]);
  }
  return r;
}

// Run control of a device-loop step, as every workgroup of the step kernel
// and its finalize kernel derive it from the control block.
__device__ __forceinline__ void loop_control(const DevCtrl* c, double& dt, bool& active, bool& dt_bad)
{
  active = c->status == 0 && c->step < c->max_step && c->time < c->stop_time - 1.0e-14;
  dt_bad = false;
  dt = 0.0;
  if (active) {
    dt = euler_dt(c->max_rate, c->cfl, c->time, c->stop_time);
    if (!isfinite(dt) || dt <= 0.0) {
      active = false;
      dt_bad = true;
    }
  }
}

// Reduces the per-workgroup partials of a step (or of a state reduction) and
// updates the run control.  One workgroup; launched after the producer, so
// the kernel boundary orders the partial stores before these loads and no
// device-scope fence is needed anywhere.  mode: 0 host-dt step, 1 device
// loop step, 2 state diagnostics only, 3 initial diagnostics of a run.
template <int NT>
__global__ void __launch_bounds__(NT) finalize_kernel(const Partial* partials, int nparts, DevCtrl* ctrl,
                                                      int mode, double* dt_log)
{
  __shared__ Red scratch[NT / 64];
  // Every operand of the combine is order independent (max, min, integer
  // sums), so the loads are batched: kLoads partials in flight per thread.
  constexpr int kLoads = 8;
  Red r = red_identity();
  for (int base = 0; base < nparts; base += NT * kLoads) {
    Partial q[kLoads];
#pragma unroll
    for (int u = 0; u < kLoads; ++u) {
      const int p = base + u * NT + (int)threadIdx.x;
      q[u] = p < nparts ? partials[p] : Partial{-INFINITY, INFINITY, INFINITY, 0, 0};
    }
#pragma unroll
    for (int u = 0; u < kLoads; ++u)
      red_combine(r, Red{q[u].rate, q[u].rho_min, q[u].p_min, q[u].bad, q[u].degenerate});
  }
  r = block_reduce<NT>(r, scratch);
  if (threadIdx.x != 0) return;
  DevCtrl& c = *ctrl;
  if (mode == 1) {
    double dt;
    bool active, dt_bad;
    loop_control(&c, dt, active, dt_bad);
    if (active) {
      if (dt_log) dt_log[c.step] = dt;
      c.time = c.time + dt;
      c.step += 1;
      c.dt_last = dt;
      c.dt_min = smin(c.dt_min, dt);
      c.dt_max = smax(c.dt_max, dt);
      c.dt_sum += dt;
      c.degenerate_total += r.deg;
      c.last_degenerate = r.deg;
      c.max_rate = r.rate;
      c.rho_min = r.rmin;
      c.p_min = r.pmin;
      c.nonfinite = r.bad;
      if (r.bad > 0) c.status = 3;
      else if (r.rmin <= 0.0) c.status = 4;
      else if (r.pmin <= 0.0) c.status = 5;
    } else if (dt_bad) {
      c.status = 2;
    }
  } else {
    c.last_degenerate = r.deg;
    c.max_rate = r.rate;
    c.rho_min = r.rmin;
    c.p_min = r.pmin;
    c.nonfinite = r.bad;
    if (mode == 3 && (r.bad > 0 || r.rmin <= 0.0 || r.pmin <= 0.0)) c.status = 1;
  }
}

__device__ __forceinline__ void store_partial(Partial* partials, int slot, const Red& r)
{
  if (threadIdx.x == 0) partials[slot] = Partial{r.rate, r.rmin, r.pmin, r.bad, r.deg};
}

template <int TX, int TY, int NT, int RIEMANN, int ORDER>
__global__ void __launch_bounds__(NT) explicit_step_kernel(StepArgs A)
{
  constexpr int RX = TX + 4, RY = TY + 4, NR = RX * RY;        // q: [-2,TX+1]x[-2,TY+1]
  constexpr int SX = TX + 2, SY = TY + 2, NS = SX * SY;        // slopes: [-1,TX]x[-1,TY]
  constexpr int XFX = TX + 1, NXF = XFX * TY;                  // flux x-faces [0,TX]x[0,TY-1]
  constexpr int NYF = TX * (TY + 1);                           // flux y-faces [0,TX-1]x[0,TY]
  constexpr int XRX = TX + 1, NXR = XRX * (TY + 2);            // x ratios [0,TX]x[-1,TY]
  constexpr int YRX = TX + 2, NYR = YRX * (TY + 1);            // y ratios [-1,TX]x[0,TY]
  constexpr bool HP = RIEMANN == 3;
  constexpr bool O2 = ORDER == 2;
  constexpr int OFF_Q = 0;
  constexpr int OFF_SX = OFF_Q + 4 * NR;
  constexpr int OFF_SY = OFF_SX + (O2 ? 4 * NS : 0);
  constexpr int OFF_RX = OFF_SY + (O2 ? 4 * NS : 0);
  constexpr int OFF_RY = OFF_RX + (HP ? NXR : 0);
  constexpr int END1 = OFF_RY + (HP ? NYR : 0);
  constexpr int OFF_FX = 0;
  constexpr int OFF_FY = 4 * NXF;
  constexpr int END2 = OFF_FY + 4 * NYF;
  constexpr int NLDS = END1 > END2 ? END1 : END2;
  __shared__ double lds[NLDS];
  __shared__ unsigned char degx[NXF];
  __shared__ unsigned char degy[NYF];
  __shared__ Red red_scratch[NT / 64];

  const GridParams& g = A.g;
  const int tid = threadIdx.x;
  const int tile = xcd_banded(blockIdx.x, A.nblocks);
  const int i0 = (tile % A.tiles_x) * TX;
  const int j0 = (tile / A.tiles_x) * TY;
  const double gamma = g.gamma;

  // Run control: every workgroup derives the same dt and activity.
  double dt = A.dt;
  bool active = true, dt_bad = false;
  if (A.device_dt) loop_control(A.ctrl, dt, active, dt_bad);
  (void)dt_bad;

  Red mine = red_identity();
  if (!active) {
    // The loop has stopped: carry the state through so the ping-pong pair
    // stays consistent.
    for (int t = tid; t < TX * TY; t += NT) {
      const int gi = i0 + t % TX, gj = j0 + t / TX;
      if (gi < g.nx && gj < g.ny) {
        const long long k = (long long)gj * A.rs + gi;
#pragma unroll
        for (int c = 0; c < 4; ++c) A.uout[c * A.cs + k] = A.uin[c * A.cs + k];
      }
    }
    store_partial(A.partials, tile, mine);
    return;
  }

  // A. primitive variables of the tile and its 2-cell halo.
  for (int t = tid; t < NR; t += NT) {
    const int li = t % RX, lj = t / RX;
    const Prim q = to_prim(load_cell(A.uin, A.rs, A.cs, g, i0 + li - 2, j0 + lj - 2), gamma);
    lds[OFF_Q + t] = q.r;
    lds[OFF_Q + NR + t] = q.u;
    lds[OFF_Q + 2 * NR + t] = q.v;
    lds[OFF_Q + 3 * NR + t] = q.p;
  }
  __syncthreads();

  // B. limited slopes (reconstruction_fluxes.hpp:2-12).
  if constexpr (O2) {
    for (int t = tid; t < NS; t += NT) {
      const int qc = (t / SX + 1) * RX + (t % SX + 1);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double* q = lds + OFF_Q + c * NR;
        const double vc = q[qc];
        lds[OFF_SX + c * NS + t] = limited_slope(vc - q[qc - 1], q[qc + 1] - vc, g.limiter);
        lds[OFF_SY + c * NS + t] = limited_slope(vc - q[qc - RX], q[qc + RX] - vc, g.limiter);
      }
    }
    __syncthreads();
  }

  // Face state of cell (ci,cj) (tile coordinates) on one side of an axis:
  // q + (0.5*side)*slope (reconstruction_fluxes.hpp:14-24).
  auto face_state = [&](int ci, int cj, int axis, double half_side) -> Prim {
    const int qi = (cj + 2) * RX + (ci + 2);
    Prim q{lds[OFF_Q + qi], lds[OFF_Q + NR + qi], lds[OFF_Q + 2 * NR + qi], lds[OFF_Q + 3 * NR + qi]};
    if constexpr (O2) {
      const int si = (cj + 1) * SX + (ci + 1);
      const double* s = lds + (axis == 0 ? OFF_SX : OFF_SY) + si;
      q.r = q.r + half_side * s[0];
      q.u = q.u + half_side * s[NS];
      q.v = q.v + half_side * s[2 * NS];
      q.p = q.p + half_side * s[3 * NS];
    }
    return q;
  };
  auto face_p = [&](int ci, int cj, int axis, double half_side) -> double {
    const int qi = (cj + 2) * RX + (ci + 2);
    double p = lds[OFF_Q + 3 * NR + qi];
    if constexpr (O2) {
      const int si = (cj + 1) * SX + (ci + 1);
      p = p + half_side * lds[(axis == 0 ? OFF_SX : OFF_SY) + 3 * NS + si];
    }
    return p;
  };

  // C. face pressure ratios for the HLLC-P sensor (reconstruction_fluxes.hpp:390-419).
  if constexpr (HP) {
    for (int t = tid; t < NXR + NYR; t += NT) {
      if (t < NXR) {
        const int fi = t % XRX, j = t / XRX - 1;
        lds[OFF_RX + t] = pressure_ratio(face_p(fi - 1, j, 0, 0.5), face_p(fi, j, 0, -0.5));
      } else {
        const int s = t - NXR;
        const int i = s % YRX - 1, fj = s / YRX;
        lds[OFF_RY + s] = pressure_ratio(face_p(i, fj - 1, 1, 0.5), face_p(i, fj, 1, -0.5));
      }
    }
    __syncthreads();
  }
  auto rx = [&](int fi, int j) { return lds[OFF_RX + (j + 1) * XRX + fi]; };
  auto ry = [&](int i, int fj) { return lds[OFF_RY + fj * YRX + (i + 1)]; };

  // D. every face flux of the tile once: all x-faces, then all y-faces, so
  // that the axis is a compile-time constant of each pass and the selects on
  // it fold away.
  const int lim_x = g.nx - i0, lim_y = g.ny - j0;  // cells of this tile inside the domain
  constexpr int NPX = (NXF + NT - 1) / NT, NPY = (NYF + NT - 1) / NT;
  auto face = [&](auto axis_c, int f, Cons& out, unsigned char& dflag) {
    constexpr int AX = decltype(axis_c)::value;
    int li, lj, ri, rj;
    bool needed;
    double sensor = 1.0;
    if constexpr (AX == 0) {
      const int fi = f % XFX, j = f / XFX;
      li = fi - 1; lj = j; ri = fi; rj = j;
      needed = fi <= lim_x && j < lim_y;
      if constexpr (HP) {
        double m = rx(fi, j);
        m = smin(m, ry(fi - 1, j));
        m = smin(m, ry(fi - 1, j + 1));
        m = smin(m, ry(fi, j));
        m = smin(m, ry(fi, j + 1));
        const double c = clamp01(m);
        sensor = c * c * c;
      }
    } else {
      const int i = f % TX, fj = f / TX;
      li = i; lj = fj - 1; ri = i; rj = fj;
      needed = i < lim_x && fj <= lim_y;
      if constexpr (HP) {
        double m = ry(i, fj);
        m = smin(m, rx(i, fj - 1));
        m = smin(m, rx(i + 1, fj - 1));
        m = smin(m, rx(i, fj));
        m = smin(m, rx(i + 1, fj));
        const double c = clamp01(m);
        sensor = c * c * c;
      }
    }
    dflag = 0;
    out = Cons{0.0, 0.0, 0.0, 0.0};
    if (needed) {
      const Prim L = face_state(li, lj, AX, 0.5);
      const Prim R = face_state(ri, rj, AX, -0.5);
      out = interface_flux(L, R, AX, gamma, RIEMANN, sensor);
      if constexpr (RIEMANN != 0) dflag = face_degenerate(L, R, AX, gamma, RIEMANN) ? 1 : 0;
    }
  };
  Cons fxv[NPX], fyv[NPY];
  unsigned char dfx[NPX], dfy[NPY];
#pragma unroll
  for (int p = 0; p < NPX; ++p) {
    const int f = tid + p * NT;
    if (f < NXF) face(std::integral_constant<int, 0>{}, f, fxv[p], dfx[p]);
  }
#pragma unroll
  for (int p = 0; p < NPY; ++p) {
    const int f = tid + p * NT;
    if (f < NYF) face(std::integral_constant<int, 1>{}, f, fyv[p], dfy[p]);
  }
  __syncthreads();  // stage A-C storage is dead from here on
#pragma unroll
  for (int p = 0; p < NPX; ++p) {
    const int f = tid + p * NT;
    if (f < NXF) {
      double* base = lds + OFF_FX + f;
      base[0] = fxv[p].r;
      base[NXF] = fxv[p].mx;
      base[2 * NXF] = fxv[p].my;
      base[3 * NXF] = fxv[p].E;
      degx[f] = dfx[p];
    }
  }
#pragma unroll
  for (int p = 0; p < NPY; ++p) {
    const int f = tid + p * NT;
    if (f < NYF) {
      double* base = lds + OFF_FY + f;
      base[0] = fyv[p].r;
      base[NYF] = fyv[p].mx;
      base[2 * NYF] = fyv[p].my;
      base[3 * NYF] = fyv[p].E;
      degy[f] = dfy[p];
    }
  }
  __syncthreads();

  // E. conservative update (explicit_update.hpp:90-96) and the epilogue.
  const double inv_dx = g.inv_dx, inv_dy = g.inv_dy;
  for (int t = tid; t < TX * TY; t += NT) {
    const int ci = t % TX, cj = t / TX;
    if (ci < lim_x && cj < lim_y) {
      const long long k = (long long)(j0 + cj) * A.rs + (i0 + ci);
      const int xl = cj * XFX + ci, yb = cj * TX + ci;
      double nu[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double fxl = lds[OFF_FX + c * NXF + xl], fxr = lds[OFF_FX + c * NXF + xl + 1];
        const double fyb = lds[OFF_FY + c * NYF + yb], fyt = lds[OFF_FY + c * NYF + yb + TX];
        const double old = A.uin[c * A.cs + k];
        nu[c] = old - dt * ((fxr - fxl) * inv_dx + (fyt - fyb) * inv_dy);
        A.uout[c * A.cs + k] = nu[c];
      }
      mine.deg += degx[xl] + degx[xl + 1] + degy[yb] + degy[yb + TX];
      red_cell(mine, Cons{nu[0], nu[1], nu[2], nu[3]}, g);
    }
  }
  mine = block_reduce<NT>(mine, red_scratch);
  store_partial(A.partials, tile, mine);
}

// Standalone max-rate + diagnostics reduction over a state (part (b) of the
// path as its own kernel; the step kernel fuses the same reduction).
template <int NT>
__global__ void __launch_bounds__(NT)
    state_reduce_kernel(const double* __restrict__ u, long long rs, long long cs, GridParams g,
                        DevCtrl* /*ctrl*/, Partial* partials, int /*mode*/)
{
  __shared__ Red red_scratch[NT / 64];
  Red mine = red_identity();
  const long long n = (long long)g.nx * g.ny;
  for (long long t = (long long)blockIdx.x * NT + threadIdx.x; t < n; t += (long long)gridDim.x * NT) {
    const int i = (int)(t % g.nx), j = (int)(t / g.nx);
    const long long k = (long long)j * rs + i;
    red_cell(mine, Cons{u[k], u[cs + k], u[2 * cs + k], u[3 * cs + k]}, g);
  }
  mine = block_reduce<NT>(mine, red_scratch);
  store_partial(partials, blockIdx.x, mine);
}

// ---------------------------------------------------------------- dispatch
constexpr int kFinalizeThreads = 1024;

struct TileShape {
  int tx, ty, nt;
};

using StepFn = void (*)(StepArgs, dim3, dim3, hipStream_t);

template <int TX, int TY, int NT, int R, int O>
void launch_step(StepArgs a, dim3 grid, dim3 block, hipStream_t s)
{
  hipLaunchKernelGGL((explicit_step_kernel<TX, TY, NT, R, O>), grid, block, 0, s, a);
}

template <int TX, int TY, int NT>
struct TileFns {
  static constexpr TileShape shape{TX, TY, NT};
  static StepFn get(int riemann, int order)
  {
    const int k = riemann * 2 + (order == 2 ? 1 : 0);
    static const StepFn fns[8] = {
        launch_step<TX, TY, NT, 0, 1>, launch_step<TX, TY, NT, 0, 2>,
        launch_step<TX, TY, NT, 1, 1>, launch_step<TX, TY, NT, 1, 2>,
        launch_step<TX, TY, NT, 2, 1>, launch_step<TX, TY, NT, 2, 2>,
        launch_step<TX, TY, NT, 3, 1>, launch_step<TX, TY, NT, 3, 2>};
    return fns[k];
  }
};

static const TileShape kShapes[] = {{16, 15, 256}, {32, 15, 512}, {32, 7, 256}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

static StepFn step_fn(int tile_id, int riemann, int order)
{
  switch (tile_id) {
    case 1: return TileFns<32, 15, 512>::get(riemann, order);
    case 2: return TileFns<32, 7, 256>::get(riemann, order);
    default: return TileFns<16, 15, 256>::get(riemann, order);
  }
}

int step_partials_needed(const ec_ctx* ctx, int tile_id)
{
  const TileShape s = kShapes[tile_id];
  const int tx = (ctx->grid.nx + s.tx - 1) / s.tx;
  const int ty = (ctx->grid.ny + s.ty - 1) / s.ty;
  return tx * ty;
}

ec_status launch_explicit_step(ec_ctx* ctx, const double* uin, double* uout, long long rs,
                               long long cs, double dt, int device_dt)
{
  const TileShape s = kShapes[ctx->tile_id];
  StepArgs a;
  a.uin = uin;
  a.uout = uout;
  a.rs = rs;
  a.cs = cs;
  a.g = ctx->grid;
  a.dt = dt;
  a.device_dt = device_dt;
  a.ctrl = ctx->ctrl;
  a.partials = ctx->partials;
  a.dt_log = device_dt ? ctx->dt_log : nullptr;
  a.tiles_x = (ctx->grid.nx + s.tx - 1) / s.tx;
  const int tiles_y = (ctx->grid.ny + s.ty - 1) / s.ty;
  a.nblocks = a.tiles_x * tiles_y;
  if (a.nblocks > ctx->max_partials) {
    ctx->last_error = "partials buffer too small for the tile shape";
    return EC_ERR_INVALID_ARGUMENT;
  }
  step_fn(ctx->tile_id, ctx->grid.riemann, ctx->grid.order)(a, dim3(a.nblocks), dim3(s.nt),
                                                             ctx->stream);
  hipLaunchKernelGGL((finalize_kernel<kFinalizeThreads>), dim3(1), dim3(kFinalizeThreads), 0,
                     ctx->stream, ctx->partials, a.nblocks, ctx->ctrl, device_dt ? 1 : 0, a.dt_log);
  EC_HIP_TRY(ctx, hipGetLastError());
  return EC_OK;
}

int reduce_blocks(const ec_ctx* ctx)
{
  const long long n = (long long)ctx->grid.nx * ctx->grid.ny;
  long long b = (n + 255) / 256;
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (int)b;
}

ec_status launch_state_reduce(ec_ctx* ctx, const double* u, long long rs, long long cs, int mode)
{
  const int nb = reduce_blocks(ctx);
  hipLaunchKernelGGL((state_reduce_kernel<256>), dim3(nb), dim3(256), 0, ctx->stream, u, rs, cs,
                     ctx->grid, ctx->ctrl, ctx->partials, mode);
  hipLaunchKernelGGL((finalize_kernel<kFinalizeThreads>), dim3(1), dim3(kFinalizeThreads), 0,
                     ctx->stream, ctx->partials, nb, ctx->ctrl, mode, nullptr);
  EC_HIP_TRY(ctx, hipGetLastError());
  return EC_OK;
}

int num_tile_shapes() { return kNumShapes; }
TileShape tile_shape(int id) { return kShapes[id]; }

}  // namespace ec

// ------------------------------------------------------------------ C-ABI
using namespace ec;

static bool layout_ok(const ec_ctx* ctx, const ec_layout* l)
{
  return l && l->row_stride >= ctx->grid.nx && l->comp_stride >= l->row_stride * ctx->grid.ny;
}

static ec_status read_diag(ec_ctx* ctx, ec_step_diag* diag)
{
  EC_HIP_TRY(ctx, hipMemcpyAsync(ctx->ctrl_host, ctx->ctrl, sizeof(DevCtrl),
                                 hipMemcpyDeviceToHost, ctx->stream));
  EC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  const DevCtrl& c = *ctx->ctrl_host;
  diag->rho_min = c.rho_min;
  diag->pressure_min = c.p_min;
  diag->nonfinite_count = c.nonfinite;
  diag->degenerate_count = c.last_degenerate;
  diag->max_rate = c.max_rate;
  return EC_OK;
}

extern "C" ec_status ec_explicit_step(ec_ctx* ctx, const double* U_in, double* U_out,
                                      const ec_layout* layout, double dt, ec_step_diag* diag)
{
  if (!ctx || !U_in || !U_out || U_in == U_out || !layout_ok(ctx, layout)) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const ec_status st =
      launch_explicit_step(ctx, U_in, U_out, layout->row_stride, layout->comp_stride, dt, 0);
  if (st != EC_OK) return st;
  return diag ? read_diag(ctx, diag) : EC_OK;
}

extern "C" ec_status ec_explicit_step_inplace(ec_ctx* ctx, double* U, const ec_layout* layout,
                                              double dt, ec_step_diag* diag)
{
  if (!ctx || !U || !layout_ok(ctx, layout)) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const long long nx = ctx->grid.nx, ny = ctx->grid.ny;
  const ec_status st = launch_explicit_step(ctx, U, ctx->scratch_state, nx, nx * ny, dt, 0);
  if (st != EC_OK) return st;
  // Copy back into the caller's layout, row by row per component.
  EC_HIP_TRY(ctx, hipMemcpy2DAsync(U, layout->row_stride * sizeof(double), ctx->scratch_state,
                                   nx * sizeof(double), nx * sizeof(double), ny,
                                   hipMemcpyDeviceToDevice, ctx->stream));
  for (int c = 1; c < 4; ++c) {
    EC_HIP_TRY(ctx, hipMemcpy2DAsync(U + c * layout->comp_stride, layout->row_stride * sizeof(double),
                                     ctx->scratch_state + c * nx * ny, nx * sizeof(double),
                                     nx * sizeof(double), ny, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return diag ? read_diag(ctx, diag) : EC_OK;
}

extern "C" ec_status ec_state_diagnostics(ec_ctx* ctx, const double* U, const ec_layout* layout,
                                          ec_step_diag* diag)
{
  if (!ctx || !U || !diag || !layout_ok(ctx, layout)) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  const ec_status st = launch_state_reduce(ctx, U, layout->row_stride, layout->comp_stride, 2);
  if (st != EC_OK) return st;
  const ec_status r = read_diag(ctx, diag);
  diag->degenerate_count = 0;
  return r;
}

extern "C" double ec_euler_dt(double max_rate, double cfl, double time, double stop_time)
{
  return euler_dt(max_rate, cfl, time, stop_time);
}

extern "C" ec_status ec_explicit_advance(ec_ctx* ctx, double* U_a, double* U_b,
                                         const ec_layout* layout, int nsteps, ec_run_state* io,
                                         int* result_in_b)
{
  if (!ctx || !U_a || !U_b || U_a == U_b || !io || !result_in_b || nsteps < 0 ||
      !layout_ok(ctx, layout)) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  DevCtrl& h = *ctx->ctrl_host;
  h.time = io->time;
  h.stop_time = io->stop_time;
  h.cfl = io->cfl;
  h.dt_last = io->dt_last;
  h.dt_min = io->step == 0 ? INFINITY : io->dt_min;
  h.dt_max = io->step == 0 ? 0.0 : io->dt_max;
  h.dt_sum = io->step == 0 ? 0.0 : io->dt_sum;
  h.rho_min = io->rho_min;
  h.p_min = io->pressure_min;
  h.max_rate = io->max_rate;
  h.degenerate_total = io->degenerate_total;
  h.step = io->step;
  h.max_step = io->max_step;
  h.nonfinite = io->nonfinite_count;
  h.status = io->status;
  h.last_degenerate = 0;
  h.counter = 0;
  EC_HIP_TRY(ctx, hipMemcpyAsync(ctx->ctrl, ctx->ctrl_host, sizeof(DevCtrl), hipMemcpyHostToDevice,
                                 ctx->stream));
  const long long rs = layout->row_stride, cs = layout->comp_stride;
  if (io->step == 0) {
    // Initial diagnostics and max rate (driver_main.hpp:216-222).
    const ec_status st = launch_state_reduce(ctx, U_a, rs, cs, 3);
    if (st != EC_OK) return st;
  }
  for (int s = 0; s < nsteps; ++s) {
    const double* in = (s & 1) ? U_b : U_a;
    double* out = (s & 1) ? U_a : U_b;
    const ec_status st = launch_explicit_step(ctx, in, out, rs, cs, 0.0, 1);
    if (st != EC_OK) return st;
  }
  EC_HIP_TRY(ctx, hipMemcpyAsync(ctx->ctrl_host, ctx->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost,
                                 ctx->stream));
  EC_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  io->time = h.time;
  io->dt_last = h.dt_last;
  io->dt_min = h.dt_min;
  io->dt_max = h.dt_max;
  io->dt_sum = h.dt_sum;
  io->rho_min = h.rho_min;
  io->pressure_min = h.p_min;
  io->max_rate = h.max_rate;
  io->degenerate_total = h.degenerate_total;
  io->step = h.step;
  io->nonfinite_count = h.nonfinite;
  io->status = h.status;
  *result_in_b = nsteps & 1;
  return EC_OK;
}

extern "C" int ec_explicit_tile_count(void) { return num_tile_shapes(); }

extern "C" ec_status ec_explicit_tile_shape(int tile_id, int* tx, int* ty, int* threads)
{
  if (tile_id < 0 || tile_id >= num_tile_shapes() || !tx || !ty || !threads) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  const TileShape s = tile_shape(tile_id);
  *tx = s.tx;
  *ty = s.ty;
  *threads = s.nt;
  return EC_OK;
}

extern "C" ec_status ec_set_explicit_tile(ec_ctx* ctx, int tile_id)
{
  if (!ctx || tile_id < 0 || tile_id >= num_tile_shapes()) return EC_ERR_INVALID_ARGUMENT;
  if (step_partials_needed(ctx, tile_id) > ctx->max_partials) return EC_ERR_INVALID_ARGUMENT;
  ctx->tile_id = tile_id;
  return EC_OK;
}
