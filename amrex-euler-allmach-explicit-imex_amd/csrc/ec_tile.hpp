// ec_tile.hpp -- one workgroup's explicit step of one TX x TY tile, and the
// reductions around it; shared by the single-grid step kernel and the
// ensemble kernels.
//
// The tile stages, in LDS:
//   A. the tile plus a 2-cell halo as primitive variables, read through the
//      per-side ghost rule (the reference's separate ghost fill is fused in);
//   B. the limited primitive slopes of every cell whose face states are
//      needed (tile plus a 1-cell ring);
//   C. for HLLC-P, the pressure ratio of every face the 5-face sensor reads;
//   D. every face flux of the tile exactly once (the reference evaluates each
//      face twice and each reconstruction up to 6 times), x-faces and then
//      y-faces, kept in registers and then written over the dead A..C storage;
//   E. the conservative update, written to U_out, with an epilogue on U_out:
//      max CFL spectral rate, min rho, min p, non-finite and degenerate-star
//      counts, stored per thread in LDS beside the fluxes and reduced by
//      wavefront 0 alone (four slots per lane, then one DPP tree).
// Stage A also checks that every state the tile reads lies in the magnitude
// window of ec_exact_ops.hpp.  When it does (always, for physical states of
// moderate size), stages C and D use the window operators of
// ec_flux_window.hpp, whose results equal the plain IEEE arithmetic bit for
// bit; otherwise they use the plain operators.
#pragma once

#include <type_traits>

#include "ec_context.hpp"
#include "ec_flux_window.hpp"
#include "ec_ghost.hpp"
#include "ec_state.hpp"
#include "ec_wave.hpp"

namespace ec {

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

// Wave priority inside the tile step.  The stages between barriers (loads
// and conversion, slopes, pressure ratios, the flux stores, the update) run
// at a raised priority and the two flux passes at the base priority: when a
// SIMD has a wave of each kind ready, it issues the barrier-bound stage first,
// so that workgroup reaches its own flux passes sooner and the flux work of
// the other workgroups fills the issue slots that the barriers leave
// (measured on the 2048^2 HLLC-P step: 197 -> 187 us per launch).
__device__ __forceinline__ void prio_raised() { __builtin_amdgcn_s_setprio(1); }
__device__ __forceinline__ void prio_base() { __builtin_amdgcn_s_setprio(0); }

struct Red {
  double rate, rmin, pmin;
  int bad, deg;
};

__device__ __forceinline__ Red red_identity()
{
  return Red{-INFINITY, INFINITY, INFINITY, 0, 0};
}

// The reduced values are never NaN (a NaN rate is dropped, a non-finite
// state's minima are -inf), so max and min are v_max_f64 / v_min_f64.
__device__ __forceinline__ void red_combine(Red& a, const Red& b)
{
  a.rate = fmax(a.rate, b.rate);
  a.rmin = fmin(a.rmin, b.rmin);
  a.pmin = fmin(a.pmin, b.pmin);
  a.bad += b.bad;
  a.deg += b.deg;
}

// Diagnostics and CFL rate of one cell (diagnostic_metrics_timestep.hpp:30-38
// and :240-245).  A NaN rate is dropped, as the reference's sequential
// max(running, value) drops it: fmax returns its non-NaN operand, and the
// running rate is never NaN.
__device__ __forceinline__ double cell_rate(double ux, double uy, double c, const GridParams& g)
{
  // division by a power-of-two cell size is the multiplication by its exact
  // reciprocal, bit for bit
  const double ax = fabs(ux) + c, ay = fabs(uy) + c;
  return (g.dx_pow2 ? ax * g.inv_dx : ax / g.dx) + (g.dy_pow2 ? ay * g.inv_dy : ay / g.dy);
}

__device__ __forceinline__ void red_cell(Red& r, const Cons& u, const GridParams& g)
{
  const Prim q = to_prim(u, g.gamma);
  const bool bad = !isfinite(q.r) || !isfinite(q.u) || !isfinite(q.v) || !isfinite(q.p) ||
                   !isfinite(u.E);
  r.rmin = fmin(r.rmin, bad ? -INFINITY : q.r);
  r.pmin = fmin(r.pmin, bad ? -INFINITY : q.p);
  r.bad += bad ? 1 : 0;
  r.rate = fmax(r.rate, cell_rate(q.u, q.v, sound_speed(q, g.gamma), g));
}

// red_cell with the window operators where every active lane's density (for
// 1/rho) and then pressure (for the sound speed) lie in the window: the same
// results, bit for bit.  Both choices are wavefront-uniform.
__device__ __forceinline__ void red_cell_w(Red& r, const Cons& u, const GridParams& g)
{
  if (EC_UNLIKELY(!__all(pos_in_window(u.r) ? 1 : 0))) {
    red_cell(r, u, g);
    return;
  }
  const Recip rr = recip(u.r);
  const double inv_rho = quot_pos(1.0, rr);
  const double ux = u.mx * inv_rho;
  const double uy = u.my * inv_rho;
  const double kinetic = 0.5 * u.r * (ux * ux + uy * uy);
  const double p = (g.gamma - 1.0) * (u.E - kinetic);
  // rho is finite inside the window
  const bool bad = !isfinite(ux) || !isfinite(uy) || !isfinite(p) || !isfinite(u.E);
  r.rmin = fmin(r.rmin, bad ? -INFINITY : u.r);
  r.pmin = fmin(r.pmin, bad ? -INFINITY : p);
  r.bad += bad ? 1 : 0;
  double c;
  if (EC_LIKELY(__all(pos_in_window(p) ? 1 : 0))) c = sqrt_core(quot_pos(g.gamma * p, rr));
  else c = sqrt(g.gamma * p / u.r);
  r.rate = fmax(r.rate, cell_rate(ux, uy, c, g));
}

// One step of the wavefront tree: combine with the value DPP control CTRL
// brings in (rows outside ROW_MASK bring the identity).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void red_dpp_step(Red& r)
{
  Red o;
  o.rate = dpp_f64_id<CTRL, ROW_MASK>(r.rate, -INFINITY);
  o.rmin = dpp_f64_id<CTRL, ROW_MASK>(r.rmin, INFINITY);
  o.pmin = dpp_f64_id<CTRL, ROW_MASK>(r.pmin, INFINITY);
  o.bad = dpp_i32<CTRL, ROW_MASK>(r.bad);
  o.deg = dpp_i32<CTRL, ROW_MASK>(r.deg);
  red_combine(r, o);
}

// Wavefront reduction (result in lane 63): quad swaps, half-row and row
// mirrors, then the row broadcasts of lanes 15 and 31, all DPP moves.  Every
// combine is order independent (max, min without NaN, integer sums), so the
// result equals any other reduction order's bit for bit.
__device__ __forceinline__ Red wave_reduce(Red r)
{
  red_dpp_step<0xB1, 0xF>(r);
  red_dpp_step<0x4E, 0xF>(r);
  red_dpp_step<0x141, 0xF>(r);
  red_dpp_step<0x140, 0xF>(r);
  red_dpp_step<0x142, 0xA>(r);
  red_dpp_step<0x143, 0xC>(r);
  return r;
}

// The wavefront tree of rate, rho_min, p_min and a packed count k (result in
// lane 63).  The quad swaps and mirrors give every lane a lane of its own
// row, so their moves need no old value; the row broadcasts keep the old
// value (the identity) in the rows they do not write.
__device__ __forceinline__ void small_tree(double& rate, double& rmin, double& pmin, int& k)
{
  auto full = [&](auto ctrl_c) {
    constexpr int CTRL = decltype(ctrl_c)::value;
    rate = fmax(rate, dpp_f64_all<CTRL>(rate));
    rmin = fmin(rmin, dpp_f64_all<CTRL>(rmin));
    pmin = fmin(pmin, dpp_f64_all<CTRL>(pmin));
    k += dpp_i32_all<CTRL>(k);
  };
  auto bcast = [&](auto ctrl_c, auto mask_c) {
    constexpr int CTRL = decltype(ctrl_c)::value, MASK = decltype(mask_c)::value;
    rate = fmax(rate, dpp_f64_id<CTRL, MASK>(rate, -INFINITY));
    rmin = fmin(rmin, dpp_f64_id<CTRL, MASK>(rmin, INFINITY));
    pmin = fmin(pmin, dpp_f64_id<CTRL, MASK>(pmin, INFINITY));
    k += dpp_i32<CTRL, MASK>(k);
  };
  full(std::integral_constant<int, 0xB1>{});
  full(std::integral_constant<int, 0x4E>{});
  full(std::integral_constant<int, 0x141>{});
  full(std::integral_constant<int, 0x140>{});
  bcast(std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xA>{});
  bcast(std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xC>{});
}

// wave_reduce for per-lane counts small enough to share one 32-bit word
// (bad < 2^15 and deg < 2^16 summed over the wavefront, as in a tile's
// epilogue): one integer move per step instead of two.
__device__ __forceinline__ Red wave_reduce_small(Red r)
{
  int k = (r.bad << 16) | r.deg;
  small_tree(r.rate, r.rmin, r.pmin, k);
  r.bad = k >> 16;
  r.deg = k & 0xffff;
  return r;
}

// A tile epilogue's workgroup reduction through LDS: every thread stores its
// values in `slots` (3 NT doubles, then NT packed counts bad << 16 | deg, with
// bad < 2^15 and deg < 2^16 over the workgroup), wavefront 0 combines NT / 64
// slots per lane and finishes with one wavefront tree.  The other wavefronts
// do no reduction work.  The result is valid in every lane of wavefront 0;
// the slots must stay untouched until wavefront 0 has returned.
template <int NT>
__device__ __forceinline__ Red tile_reduce(const Red& r, double* slots)
{
  static_assert(NT % 64 == 0, "whole wavefronts");
  int* ks = reinterpret_cast<int*>(slots + 3 * NT);
  const int t = threadIdx.x;
  slots[t] = r.rate;
  slots[NT + t] = r.rmin;
  slots[2 * NT + t] = r.pmin;
  ks[t] = (r.bad << 16) | r.deg;
  __syncthreads();
  Red out = r;
  if (t < 64) {
    double rate = slots[t], rmin = slots[NT + t], pmin = slots[2 * NT + t];
    int k = ks[t];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) {
      rate = fmax(rate, slots[64 * w + t]);
      rmin = fmin(rmin, slots[NT + 64 * w + t]);
      pmin = fmin(pmin, slots[2 * NT + 64 * w + t]);
      k += ks[64 * w + t];
    }
    small_tree(rate, rmin, pmin, k);
    out.rate = readlane_f64<63>(rate);
    out.rmin = readlane_f64<63>(rmin);
    out.pmin = readlane_f64<63>(pmin);
    k = __builtin_amdgcn_readlane(k, 63);
    out.bad = k >> 16;
    out.deg = k & 0xffff;
  }
  return out;
}

// Workgroup reduction; result valid in thread 0.  SMALL: the per-lane counts
// of a tile epilogue (see wave_reduce_small).
template <int NT, bool SMALL = false>
__device__ __forceinline__ Red block_reduce(Red r, Red* scratch)
{
  r = SMALL ? wave_reduce_small(r) : wave_reduce(r);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 63) scratch[wave] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    r = scratch[0];
    for (int w = 1; w < NT / 64; ++w) red_combine(r, scratch[w]);
  }
  return r;
}

// Run control of a device-loop step, as every workgroup of the step kernel
// and its finalize kernel derive it from the control block: whether the
// loop takes another step, and that step's dt (a dt that is not finite and
// positive ends the run instead).
__device__ __forceinline__ bool loop_active(const DevCtrl* c)
{
  return c->status == 0 && c->step < c->max_step && c->time < c->stop_time - 1.0e-14;
}

__device__ __forceinline__ bool dt_valid(double dt) { return isfinite(dt) && dt > 0.0; }

__device__ __forceinline__ void loop_control(const DevCtrl* c, double& dt, bool& active, bool& dt_bad)
{
  active = loop_active(c);
  dt_bad = false;
  dt = 0.0;
  if (active) {
    dt = euler_dt(c->max_rate, c->cfl, c->time, c->stop_time);
    if (!dt_valid(dt)) {
      active = false;
      dt_bad = true;
    }
  }
}

// Folds the reduction of one step (or of one state) into the run control:
// the explicit branch of the driver loop (app/driver_main.hpp:224-310).
// mode: 0 host-dt step, 1 device-loop step, 2 state diagnostics only,
// 3 initial diagnostics of a run.
__device__ __forceinline__ void ctrl_fold(DevCtrl& c, const Red& r, int mode, double* dt_log)
{
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

// SPLIT_SLOPES: see SPLIT below.  PREFETCH_OWN: each thread's own cell of
// stage E (one cell per thread when the tile has no more cells than threads)
// is loaded with stage A's loads, so a latency-bound caller (one tile per
// round) does not wait on a second round trip in stage E; it costs 8 VGPRs
// through stages A-D.
template <int TX, int TY, int NT, int RIEMANN, int ORDER, bool SPLIT_SLOPES = true,
          bool PREFETCH_OWN = false>
struct TileStep {
  static constexpr int RX = TX + 4, RY = TY + 4, NR = RX * RY;        // q: [-2,TX+1]x[-2,TY+1]
  static constexpr int SX = TX + 2, SY = TY + 2, NS = SX * SY;        // slopes: [-1,TX]x[-1,TY]
  static constexpr int XFX = TX + 1, NXF = XFX * TY;                  // flux x-faces [0,TX]x[0,TY-1]
  static constexpr int NYF = TX * (TY + 1);                           // flux y-faces [0,TX-1]x[0,TY]
  static constexpr int XRX = TX + 1, NXR = XRX * (TY + 2);            // x ratios [0,TX]x[-1,TY]
  static constexpr int YRX = TX + 2, NYR = YRX * (TY + 1);            // y ratios [-1,TX]x[0,TY]
  static constexpr bool HP = RIEMANN == 3;
  static constexpr bool O2 = ORDER == 2;
  // SPLIT: the slopes of rho, u and v are held for one axis at a time (the
  // x-slopes until the x-faces are done, then the y-slopes in their place);
  // both pressure slopes stay, for the sensor's ratios.  Five slope planes
  // instead of eight.
  static constexpr bool SPLIT = O2 && SPLIT_SLOPES;
  // YOFF: with every slope plane held (no barrier between the two face
  // passes) and room for both face sets, the y-faces go to the wavefronts
  // after the x-faces' (y-face f on thread YOFF + f), so that the two passes
  // run side by side on different wavefronts instead of one after the other
  // in the same threads; 0 otherwise.
  static constexpr int XW = (NXF + 63) / 64 * 64;
  static constexpr int YOFF = (!SPLIT && XW + NYF <= NT) ? XW : 0;
  static constexpr int OFF_Q = 0;
  static constexpr int OFF_S = OFF_Q + 4 * NR;
  static constexpr int NSP = O2 ? (SPLIT ? 5 : 8) : 0;
  static constexpr int OFF_SX = OFF_S;                      // unsplit: x-slopes, 4 planes
  static constexpr int OFF_SY = OFF_S + 4 * NS;             // unsplit: y-slopes, 4 planes
  static constexpr int OFF_RX = OFF_S + NSP * NS;
  static constexpr int OFF_RY = OFF_RX + (HP ? NXR : 0);
  static constexpr int END1 = OFF_RY + (HP ? NYR : 0);
  static constexpr int OFF_FX = 0;
  static constexpr int OFF_FY = 4 * NXF;
  static constexpr int END2 = OFF_FY + 4 * NYF;
  // the epilogue reduction's slots, beside the fluxes stage E reads
  static constexpr int OFF_RED = END2;
  static constexpr int END3 = OFF_RED + 3 * NT + (NT + 1) / 2;
  static constexpr int NLDS = END1 > END3 ? END1 : END3;

  struct Smem {
    double lds[NLDS];
    unsigned char degx[NXF];
    unsigned char degy[NYF];
    Red red_scratch[NT / 64];
    int wave_ok[NT / 64];
  };

  // Copies the tile through unchanged (a stopped run keeps its ping-pong
  // pair consistent).
  __device__ static void carry(const double* __restrict__ uin, double* __restrict__ uout,
                               long long rs, long long cs, const GridParams& g, int i0, int j0)
  {
    for (int t = threadIdx.x; t < TX * TY; t += NT) {
      const int gi = i0 + t % TX, gj = j0 + t / TX;
      if (gi < g.nx && gj < g.ny) {
        const long long k = (long long)gj * rs + gi;
#pragma unroll
        for (int c = 0; c < 4; ++c) uout[c * cs + k] = uin[c * cs + k];
      }
    }
  }

  // Advances the tile whose first cell is (i0, j0) by dt from uin to uout and
  // returns the tile's epilogue reduction (valid in thread 0).
  // diag = false skips the epilogue (CFL rate, diagnostics, degenerate
  // count): a replay with known dts needs only the state.
  static constexpr int NA = (NR + NT - 1) / NT;  // stage-A cells per thread
  static constexpr bool OWN = PREFETCH_OWN && TX * TY <= NT;

  // What stage A loads, held in registers between load() and run_loaded().
  struct Loaded {
    Cons raw[NA];
    CellSrc src[NA];
    double own[4];
    bool interior;
  };

  // Stage A's loads.  Each lane first finds where its cells come from
  // (directly on a tile whose halo lies inside the domain, where the ghost
  // rule is the identity; through the ghost rule otherwise), then every load
  // is issued in straight-line code: loads under per-lane branches would each
  // wait for the loads before them.  Lanes without a cell read a clamped one
  // and ignore it; fixed and exact ghost cells read cell 0 and ignore it.
  // Nothing waits for the loads here, so a caller can issue them early.
  __device__ static void load(Loaded& ld, const double* __restrict__ uin, long long rs, long long cs,
                              const GridParams& g, int i0, int j0)
  {
    const int tid = threadIdx.x;
    ld.interior = i0 >= 2 && j0 >= 2 && i0 + TX + 2 <= g.nx && j0 + TY + 2 <= g.ny;
    if constexpr (OWN) {
      const int ci = min(tid % TX, g.nx - i0 - 1), cj = min(tid / TX, min(TY, g.ny - j0) - 1);
      const long long k = (long long)(j0 + cj) * rs + (i0 + ci);
#pragma unroll
      for (int c = 0; c < 4; ++c) ld.own[c] = uin[c * cs + k];
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const int t = min(tid + a * NT, NR - 1);
      const int i = i0 + t % RX - 2, j = j0 + t / RX - 2;
      if (ld.interior) ld.src[a] = CellSrc{(long long)j * rs + i, 0, 0, false, false};
      else ld.src[a] = cell_src(g, rs, i, j);
      const long long k = ld.src[a].kind == 0 ? ld.src[a].k : 0;
      ld.raw[a] = Cons{uin[k], uin[cs + k], uin[2 * cs + k], uin[3 * cs + k]};
    }
  }

  __device__ static Red run(Smem& sm, const double* __restrict__ uin, double* __restrict__ uout,
                            long long rs, long long cs, const GridParams& g, int i0, int j0,
                            double dt, bool diag = true, const double* time_src = nullptr,
                            const double* dt_src = nullptr, const int* skip_src = nullptr)
  {
    Loaded ld;
    prio_raised();
    load(ld, uin, rs, cs, g, i0, j0);
    return run_loaded(sm, ld, uin, uout, rs, cs, g, i0, j0, dt, diag, time_src, dt_src, skip_src);
  }

  // The step of the tile from the values load() fetched from uin.
  // dt_src, skip_src (optional, LDS): a dt formed by one wavefront before
  // stage A's barrier and read in stage E, and a flag, read after that
  // barrier, that the step is not taken (the tile is carried through).
  __device__ static Red run_loaded(Smem& sm, const Loaded& ld, const double* __restrict__ uin,
                                   double* __restrict__ uout, long long rs, long long cs,
                                   const GridParams& g, int i0, int j0, double dt, bool diag = true,
                                   const double* time_src = nullptr, const double* dt_src = nullptr,
                                   const int* skip_src = nullptr)
  {
    prio_raised();
    double* lds = sm.lds;
    unsigned char* degx = sm.degx;
    unsigned char* degy = sm.degy;
    const int tid = threadIdx.x;
    const double gamma = g.gamma;
    Red mine = red_identity();
    const bool interior = ld.interior;
    bool ok = pos_in_window(gamma) && pos_in_window(gamma - 1.0);

    // A. primitive variables of the tile and its 2-cell halo, from the loads.
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      const int t = tid + a * NT;
      if (t < NR) {
        const Cons c = interior ? ld.raw[a]
                                : cell_value(ld.src[a], ld.raw[a], g, i0 + t % RX - 2, j0 + t / RX - 2, time_src);
        const Prim q = to_prim(c, gamma);
        lds[OFF_Q + t] = q.r;
        lds[OFF_Q + NR + t] = q.u;
        lds[OFF_Q + 2 * NR + t] = q.v;
        lds[OFF_Q + 3 * NR + t] = q.p;
        ok = ok & pos_in_window(q.r) & pos_in_window(q.p) & vel_in_window(q.u) & vel_in_window(q.v);
      }
    }
    {
      const int wave_all = __all(ok ? 1 : 0);
      if ((tid & 63) == 0) sm.wave_ok[tid >> 6] = wave_all;
    }
    __syncthreads();
    if (skip_src && EC_UNLIKELY(*skip_src != 0)) {
      carry(uin, uout, rs, cs, g, i0, j0);
      return red_identity();
    }
    bool win = true;
  #pragma unroll
    for (int w = 0; w < NT / 64; ++w) win = win && sm.wave_ok[w] != 0;

    // Slope planes: component c (0..3 = rho, u, v, p) along axis.
    auto slope_plane = [&](int axis, int c) -> int {
      if constexpr (SPLIT) return c < 3 ? OFF_S + c * NS : OFF_S + (3 + axis) * NS;
      return (axis == 0 ? OFF_SX : OFF_SY) + c * NS;
    };
    // B. limited slopes (reconstruction_fluxes.hpp:2-12).
    if constexpr (SPLIT) {
      // phase 1: every x-slope, and the y-slope of p
      auto slopes1 = [&](auto fast_c, auto half_c) {
        constexpr bool FAST = decltype(fast_c)::value;
        constexpr bool HALF = decltype(half_c)::value;
        constexpr double SC = HALF ? 0.5 : 1.0;
        for (int t = tid; t < NS; t += NT) {
          const int qc = (t / SX + 1) * RX + (t % SX + 1);
          double st[4][3], pd, pu;
  #pragma unroll
          for (int c = 0; c < 4; ++c) {
            const double* q = lds + OFF_Q + c * NR;
            st[c][0] = q[qc];
            st[c][1] = q[qc - 1];
            st[c][2] = q[qc + 1];
          }
          pd = lds[OFF_Q + 3 * NR + qc - RX];
          pu = lds[OFF_Q + 3 * NR + qc + RX];
  #pragma unroll
          for (int c = 0; c < 4; ++c) {
            const double vc = st[c][0];
            const double sx = FAST ? minmod_finite(vc - st[c][1], st[c][2] - vc)
                                   : limited_slope(vc - st[c][1], st[c][2] - vc, g.limiter);
            lds[OFF_S + c * NS + t] = SC * sx;
          }
          const double vc = st[3][0];
          const double sy = FAST ? minmod_finite(vc - pd, pu - vc) : limited_slope(vc - pd, pu - vc, g.limiter);
          lds[OFF_S + 4 * NS + t] = SC * sy;
        }
      };
      constexpr bool HW = true;
      if (EC_LIKELY(win && g.limiter == 0)) slopes1(std::true_type{}, std::bool_constant<HW>{});
      else if (win) slopes1(std::false_type{}, std::bool_constant<HW>{});
      else slopes1(std::false_type{}, std::false_type{});
      __syncthreads();
    } else if constexpr (O2) {
      // each cell's 5-point stencil of the 4 variables is read before any
      // slope is written, so the reads issue back to back
      // HALF: the window path stores 0.5 * slope, which is exact there
      // (every nonzero slope is at least 2^-153 in magnitude: a difference of
      // window values, or a limiter combination of such differences), so a
      // face state is q + h or q - h, one operation per component.
      auto slopes = [&](auto fast_c, auto half_c) {
        constexpr bool FAST = decltype(fast_c)::value;
        constexpr bool HALF = decltype(half_c)::value;
        constexpr double SC = HALF ? 0.5 : 1.0;
        for (int t = tid; t < NS; t += NT) {
          const int qc = (t / SX + 1) * RX + (t % SX + 1);
          double st[4][5];
  #pragma unroll
          for (int c = 0; c < 4; ++c) {
            const double* q = lds + OFF_Q + c * NR;
            st[c][0] = q[qc];
            st[c][1] = q[qc - 1];
            st[c][2] = q[qc + 1];
            st[c][3] = q[qc - RX];
            st[c][4] = q[qc + RX];
          }
  #pragma unroll
          for (int c = 0; c < 4; ++c) {
            const double vc = st[c][0];
            if constexpr (FAST) {
              lds[OFF_SX + c * NS + t] = SC * minmod_finite(vc - st[c][1], st[c][2] - vc);
              lds[OFF_SY + c * NS + t] = SC * minmod_finite(vc - st[c][3], st[c][4] - vc);
            } else {
              lds[OFF_SX + c * NS + t] = SC * limited_slope(vc - st[c][1], st[c][2] - vc, g.limiter);
              lds[OFF_SY + c * NS + t] = SC * limited_slope(vc - st[c][3], st[c][4] - vc, g.limiter);
            }
          }
        }
      };
      // minmod on finite differences when every state of the tile is inside
      // the window
      constexpr bool HW = true;
      if (EC_LIKELY(win && g.limiter == 0)) slopes(std::true_type{}, std::bool_constant<HW>{});
      else if (win) slopes(std::false_type{}, std::bool_constant<HW>{});
      else slopes(std::false_type{}, std::false_type{});
      __syncthreads();
    }

    // Face state of cell (ci,cj) (tile coordinates) on one side of an axis:
    // q + (0.5*side)*slope (reconstruction_fluxes.hpp:14-24).
    // HS: the slopes in LDS are halved (the window path), so half_side only
    // chooses between adding and subtracting them; (-0.5) * s == -(0.5 * s).
    auto face_state = [&](auto hs_c, int ci, int cj, int axis, double half_side) -> Prim {
      constexpr bool HS = decltype(hs_c)::value;
      const int qi = (cj + 2) * RX + (ci + 2);
      Prim q{lds[OFF_Q + qi], lds[OFF_Q + NR + qi], lds[OFF_Q + 2 * NR + qi], lds[OFF_Q + 3 * NR + qi]};
      if constexpr (O2) {
        const int si = (cj + 1) * SX + (ci + 1);
        const double sr = lds[slope_plane(axis, 0) + si], su = lds[slope_plane(axis, 1) + si];
        const double sv = lds[slope_plane(axis, 2) + si], sp = lds[slope_plane(axis, 3) + si];
        if constexpr (HS) {
          if (half_side > 0.0) {
            q.r = q.r + sr;
            q.u = q.u + su;
            q.v = q.v + sv;
            q.p = q.p + sp;
          } else {
            q.r = q.r - sr;
            q.u = q.u - su;
            q.v = q.v - sv;
            q.p = q.p - sp;
          }
        } else {
          q.r = q.r + half_side * sr;
          q.u = q.u + half_side * su;
          q.v = q.v + half_side * sv;
          q.p = q.p + half_side * sp;
        }
      }
      return q;
    };
    auto face_p = [&](auto hs_c, int ci, int cj, int axis, double half_side) -> double {
      constexpr bool HS = decltype(hs_c)::value;
      const int qi = (cj + 2) * RX + (ci + 2);
      double p = lds[OFF_Q + 3 * NR + qi];
      if constexpr (O2) {
        const int si = (cj + 1) * SX + (ci + 1);
        const double s = lds[slope_plane(axis, 3) + si];
        if constexpr (HS) p = half_side > 0.0 ? p + s : p - s;
        else p = p + half_side * s;
      }
      return p;
    };
    using HalfW = std::true_type;

    // C. face pressure ratios for the HLLC-P sensor (reconstruction_fluxes.hpp:390-419).
    // Inside the window both face pressures lie between neighbouring cell
    // pressures, so they are finite and at least 2^-40 > 1e-14: the floor is
    // the identity, and the ratio min/max lies in (0, 1] already.
    auto ratio_w = [](double pl, double pr) {
      return quot_pos(fmin(pl, pr), recip(fmax(pl, pr)));
    };
    if constexpr (HP) {
      if (EC_LIKELY(win)) {
        for (int t = tid; t < NXR + NYR; t += NT) {
          if (t < NXR) {
            const int fi = t % XRX, j = t / XRX - 1;
            lds[OFF_RX + t] = ratio_w(face_p(HalfW{}, fi - 1, j, 0, 0.5), face_p(HalfW{}, fi, j, 0, -0.5));
          } else {
            const int s = t - NXR;
            const int i = s % YRX - 1, fj = s / YRX;
            lds[OFF_RY + s] = ratio_w(face_p(HalfW{}, i, fj - 1, 1, 0.5), face_p(HalfW{}, i, fj, 1, -0.5));
          }
        }
      } else {
        for (int t = tid; t < NXR + NYR; t += NT) {
          if (t < NXR) {
            const int fi = t % XRX, j = t / XRX - 1;
            lds[OFF_RX + t] = pressure_ratio(face_p(std::false_type{}, fi - 1, j, 0, 0.5),
                                            face_p(std::false_type{}, fi, j, 0, -0.5));
          } else {
            const int s = t - NXR;
            const int i = s % YRX - 1, fj = s / YRX;
            lds[OFF_RY + s] = pressure_ratio(face_p(std::false_type{}, i, fj - 1, 1, 0.5),
                                            face_p(std::false_type{}, i, fj, 1, -0.5));
          }
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
    // W: the window operators (every state of the tile inside the window).
    const Recip rg = recip(gamma - 1.0);  // gamma - 1 lies in the window there
    auto face = [&](auto axis_c, auto w_c, int f, Cons& out, unsigned char& dflag) {
      constexpr int AX = decltype(axis_c)::value;
      constexpr bool W = decltype(w_c)::value;
      // inside the window the ratios are finite and in (0, 1]: min is fmin and
      // clamp01 the identity
      auto mn = [](double a, double b) { return W ? fmin(a, b) : smin(a, b); };
      int li, lj, ri, rj;
      bool needed;
      double sensor = 1.0;
      if constexpr (AX == 0) {
        const int fi = f % XFX, j = f / XFX;
        li = fi - 1; lj = j; ri = fi; rj = j;
        needed = fi <= lim_x && j < lim_y;
        if constexpr (HP) {
          double m = rx(fi, j);
          m = mn(m, ry(fi - 1, j));
          m = mn(m, ry(fi - 1, j + 1));
          m = mn(m, ry(fi, j));
          m = mn(m, ry(fi, j + 1));
          const double c = W ? m : clamp01(m);
          sensor = c * c * c;
        }
      } else {
        const int i = f % TX, fj = f / TX;
        li = i; lj = fj - 1; ri = i; rj = fj;
        needed = i < lim_x && fj <= lim_y;
        if constexpr (HP) {
          double m = ry(i, fj);
          m = mn(m, rx(i, fj - 1));
          m = mn(m, rx(i + 1, fj - 1));
          m = mn(m, rx(i, fj));
          m = mn(m, rx(i + 1, fj));
          const double c = W ? m : clamp01(m);
          sensor = c * c * c;
        }
      }
      dflag = 0;
      out = Cons{0.0, 0.0, 0.0, 0.0};
      if (needed) {
        using HS = std::bool_constant<W>;
        const Prim L = face_state(HS{}, li, lj, AX, 0.5);
        const Prim R = face_state(HS{}, ri, rj, AX, -0.5);
        if constexpr (W) {
          bool degen;
          out = interface_flux_w<AX, RIEMANN>(L, R, gamma, rg, sensor, degen);
          dflag = degen ? 1 : 0;
        } else {
          out = interface_flux(L, R, AX, gamma, RIEMANN, sensor);
          if constexpr (RIEMANN != 0) {
            if (diag) dflag = face_degenerate(L, R, AX, gamma, RIEMANN) ? 1 : 0;
          }
        }
      }
    };
    Cons fxv[NPX], fyv[NPY];
    unsigned char dfx[NPX], dfy[NPY];
    // phase 2 of the split slopes: the y-slopes of rho, u and v, over the
    // x-slopes the x-faces no longer need
    auto slopes2 = [&](auto fast_c, auto half_c) {
      constexpr bool FAST = decltype(fast_c)::value;
      constexpr bool HALF = decltype(half_c)::value;
      constexpr double SC = HALF ? 0.5 : 1.0;
      for (int t = tid; t < NS; t += NT) {
        const int qc = (t / SX + 1) * RX + (t % SX + 1);
        double st[3][3];
  #pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double* q = lds + OFF_Q + c * NR;
          st[c][0] = q[qc];
          st[c][1] = q[qc - RX];
          st[c][2] = q[qc + RX];
        }
  #pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double vc = st[c][0];
          const double sy = FAST ? minmod_finite(vc - st[c][1], st[c][2] - vc)
                                 : limited_slope(vc - st[c][1], st[c][2] - vc, g.limiter);
          lds[OFF_S + c * NS + t] = SC * sy;
        }
      }
    };
    auto all_faces = [&](auto w_c) {
      constexpr bool W = decltype(w_c)::value;
      prio_base();
  #pragma unroll
      for (int p = 0; p < NPX; ++p) {
        const int f = tid + p * NT;
        if (f < NXF) face(std::integral_constant<int, 0>{}, w_c, f, fxv[p], dfx[p]);
      }
      prio_raised();
      if constexpr (SPLIT) {
        __syncthreads();  // every x-face has read the x-slopes
        if (EC_LIKELY(W && g.limiter == 0)) slopes2(std::true_type{}, std::bool_constant<W>{});
        else slopes2(std::false_type{}, std::bool_constant<W>{});
        __syncthreads();
      }
      prio_base();
  #pragma unroll
      for (int p = 0; p < NPY; ++p) {
        const int f = tid + p * NT - YOFF;
        if ((YOFF == 0 || f >= 0) && f < NYF) face(std::integral_constant<int, 1>{}, w_c, f, fyv[p], dfy[p]);
      }
      prio_raised();
    };
    if (EC_LIKELY(win)) all_faces(std::true_type{});
    else all_faces(std::false_type{});
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
      const int f = tid + p * NT - YOFF;
      if ((YOFF == 0 || f >= 0) && f < NYF) {
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
    if (dt_src) dt = *dt_src;
    const double inv_dx = g.inv_dx, inv_dy = g.inv_dy;
    for (int t = tid; t < TX * TY; t += NT) {
      const int ci = t % TX, cj = t / TX;
      if (ci < lim_x && cj < lim_y) {
        const long long k = (long long)(j0 + cj) * rs + (i0 + ci);
        const int xl = cj * XFX + ci, yb = cj * TX + ci;
        double nu[4];
  #pragma unroll
        for (int c = 0; c < 4; ++c) {
          const double fxl = lds[OFF_FX + c * NXF + xl], fxr = lds[OFF_FX + c * NXF + xl + 1];
          const double fyb = lds[OFF_FY + c * NYF + yb], fyt = lds[OFF_FY + c * NYF + yb + TX];
          const double old = OWN ? ld.own[c] : uin[c * cs + k];
          nu[c] = old - dt * ((fxr - fxl) * inv_dx + (fyt - fyb) * inv_dy);
          uout[c * cs + k] = nu[c];
        }
        if (diag) {
          mine.deg += degx[xl] + degx[xl + 1] + degy[yb] + degy[yb + TX];
          red_cell_w(mine, Cons{nu[0], nu[1], nu[2], nu[3]}, g);
        }
      }
    }
    if (!diag) return mine;
    return tile_reduce<NT>(mine, lds + OFF_RED);
  }
};

}  // namespace ec
