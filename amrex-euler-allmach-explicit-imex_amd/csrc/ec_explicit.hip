// ec_explicit.hip -- fused explicit O2 step and the CFL / diagnostics
// reduction for gfx950 (FP64 stencils, no MFMA).
//
// One workgroup advances one TX x TY tile (the stages are in ec_tile.hpp)
// and stores its epilogue reduction (max CFL rate, min rho, min p,
// non-finite and degenerate counts) as a per-workgroup partial.  A
// one-workgroup finalize kernel, launched next on the stream, folds the
// partials and advances the device-resident run control (dt, time, step,
// status), so consecutive steps need no host round trip and no device-scope
// fence.
//
// Workgroups are numbered so that the 8 XCDs each receive a contiguous band
// of tile rows: neighbouring tiles then share halo rows through one L2.

#include "ec_tile.hpp"

namespace ec {

struct StepArgs {
  const double* __restrict__ uin;
  double* __restrict__ uout;
  long long rs, cs;
  GridParams g;
  double dt;        // used when device_dt == 0
  int device_dt;    // 1: dt from ctrl (device-resident loop)
  int diag;         // 0: state only (no epilogue reductions, no finalize)
  DevCtrl* ctrl;
  Partial* partials;
  double* dt_log;   // optional: dt of every device-loop step, indexed by step
  int tiles_x, nblocks;
};

// First level of the partials reduction of a large grid: workgroup b
// reduces partials [b * chunk, min((b + 1) * chunk, nparts)) into out[b], one
// load per thread in flight, so that the one-workgroup finalize below reads a
// few dozen partials instead of thousands.  Order independent, like every
// combine here.
template <int NT>
__global__ void __launch_bounds__(NT) partials_reduce_kernel(const Partial* __restrict__ partials,
                                                             int nparts, int chunk, Partial* __restrict__ out)
{
  __shared__ Red scratch[NT / 64];
  Red r = red_identity();
  const int lo = blockIdx.x * chunk, hi = min(lo + chunk, nparts);
  for (int p = lo + (int)threadIdx.x; p < hi; p += NT) {
    const Partial q = partials[p];
    red_combine(r, Red{q.rate, q.rho_min, q.p_min, q.bad, q.degenerate});
  }
  r = block_reduce<NT>(r, scratch);
  store_partial(out, blockIdx.x, r);
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
  ctrl_fold(*ctrl, r, mode, dt_log);
}


// Waves per SIMD the step kernel is compiled for: the 256-thread tiles (LDS
// below 32 KiB with the split slopes) run five workgroups per CU when their
// registers fit 96 VGPRs; the 512-thread tile is held to two workgroups by its
// LDS anyway.  The strip tile (TY = 4) serves grids too small to fill the chip,
// where the length of one tile's chain matters more than occupancy: it keeps
// every slope plane, so that its y-faces run on wavefronts of their own
// (TileStep::YOFF), with the registers of four workgroups per CU.
template <int TY>
constexpr bool step_split_slopes()
{
  return TY > 4;
}
template <int TY, int NT>
constexpr int step_waves_per_eu()
{
  return NT == 256 ? (step_split_slopes<TY>() ? 5 : 4) : 1;
}

template <int TX, int TY, int NT, int RIEMANN, int ORDER>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(step_waves_per_eu<TY, NT>())))
explicit_step_kernel(StepArgs A)
{
  using T = TileStep<TX, TY, NT, RIEMANN, ORDER, step_split_slopes<TY>()>;
  __shared__ typename T::Smem sm;
  const int tile = xcd_banded(blockIdx.x, A.nblocks);
  const int i0 = (tile % A.tiles_x) * TX;
  const int j0 = (tile / A.tiles_x) * TY;

  // Run control: every workgroup derives the same dt and activity.  In the
  // device loop the activity test is cheap and every wavefront makes it; dt
  // (a division) is formed by wavefront 0 alone into LDS and read by the
  // others in stage E, and a bad dt (which ends the run) is read after
  // stage A's barrier, where the tile is then carried through.
  double dt = A.dt;
  bool active = true;
  __shared__ double s_dt;
  __shared__ int s_skip;
  const double* dt_src = nullptr;
  const int* skip_src = nullptr;
  if (A.device_dt) {
    const DevCtrl* c = A.ctrl;
    active = loop_active(c);
    if (active) {
      if (threadIdx.x < 64) {
        const double d = euler_dt(c->max_rate, c->cfl, c->time, c->stop_time);
        if (threadIdx.x == 0) {
          s_dt = d;
          s_skip = dt_valid(d) ? 0 : 1;
        }
      }
      dt_src = &s_dt;
      skip_src = &s_skip;
    }
  }
  if (!active) {
    T::carry(A.uin, A.uout, A.rs, A.cs, A.g, i0, j0);
    store_partial(A.partials, tile, red_identity());
    return;
  }
  const Red r = T::run(sm, A.uin, A.uout, A.rs, A.cs, A.g, i0, j0, dt, A.diag != 0, nullptr, dt_src,
                       skip_src);
  if (A.diag) store_partial(A.partials, tile, r);
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

// Launches the step kernel and returns its workgroup count (the number of
// partials it writes), or 0 on a HIP error.
using StepFn = int (*)(const StepArgs&, hipStream_t);

template <int TX, int TY, int NT, int R, int O>
int launch_step(const StepArgs& a, hipStream_t s)
{
  hipLaunchKernelGGL((explicit_step_kernel<TX, TY, NT, R, O>), dim3(a.nblocks), dim3(NT), 0, s, a);
  return a.nblocks;
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

static const TileShape kShapes[] = {{16, 15, 256}, {32, 15, 512}, {32, 7, 256}, {24, 4, 256}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);
static_assert(TileStep<24, 4, 256, 3, 2, step_split_slopes<4>()>::YOFF == 128, "24x4: y-faces on wavefronts 2-3");

static StepFn step_fn(int tile_id, int riemann, int order)
{
  switch (tile_id) {
    case 1: return TileFns<32, 15, 512>::get(riemann, order);
    case 2: return TileFns<32, 7, 256>::get(riemann, order);
    case 3: return TileFns<24, 4, 256>::get(riemann, order);
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

int tiles_needed(int nx, int ny, int tile_id)
{
  const TileShape s = kShapes[tile_id];
  return ((nx + s.tx - 1) / s.tx) * ((ny + s.ty - 1) / s.ty);
}

// The tile a new context starts with: the 16x15 tile, or the 24x4 strip tile
// (id 3) for a grid whose 16x15 tiles number fewer than 1.5 per CU.  On such a
// grid the step waits on one tile's chain of dependent work, not on the issue
// rate, and the strip tile's chain is shorter (its y-faces run beside its
// x-faces, and stage A has one cell per thread).  Measured with the device
// loop: 0.76-0.84 of the 16x15 tile's step time on grids up to 256^2 (288
// tiles of 16x15) and on 4-row strips; 1.04-1.21 from 102,400 cells (440
// tiles) up.
int default_tile_id(int nx, int ny, int device)
{
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) {
    return 0;
  }
  return 2 * tiles_needed(nx, ny, 0) < 3 * cus ? 3 : 0;
}

// The step kernel and its finalize on a caller-supplied workspace (run
// control block and max_partials tile partials).  ev0/ev1: optional events
// recorded around the step kernel.  device_dt as below.
ec_status launch_step_raw(const GridParams& grid, int tile_id, hipStream_t stream, const double* uin,
                          double* uout, long long rs, long long cs, double dt, int device_dt,
                          DevCtrl* ctrl, Partial* partials, int max_partials, double* dt_log,
                          hipEvent_t ev0, hipEvent_t ev1)
{
  const TileShape s = kShapes[tile_id];
  StepArgs a;
  a.uin = uin;
  a.uout = uout;
  a.rs = rs;
  a.cs = cs;
  a.g = grid;
  // Exact ghost states at the step's start time: the run control's time in
  // the device loop (read by the kernel), the boundary time else.
  a.g.time_ptr = device_dt == 1 ? &ctrl->time : nullptr;
  a.dt = dt;
  a.device_dt = device_dt == 1 ? 1 : 0;
  a.diag = device_dt == 2 ? 0 : 1;
  a.ctrl = ctrl;
  a.partials = partials;
  a.dt_log = device_dt ? dt_log : nullptr;
  a.tiles_x = (grid.nx + s.tx - 1) / s.tx;
  const int tiles_y = (grid.ny + s.ty - 1) / s.ty;
  a.nblocks = a.tiles_x * tiles_y;
  if (a.nblocks > max_partials) return EC_ERR_INVALID_ARGUMENT;
  if (ev0 && hipEventRecord(ev0, stream) != hipSuccess) return EC_ERR_HIP;
  const int nparts = step_fn(tile_id, grid.riemann, grid.order)(a, stream);
  if (nparts <= 0) return EC_ERR_HIP;
  if (ev1 && hipEventRecord(ev1, stream) != hipSuccess) return EC_ERR_HIP;
  if (a.diag) {
    // beyond kTwoLevelMin partials they are first reduced by up to
    // kPartialGroupsMax workgroups into the slots after them
    const Partial* fin = partials;
    int nfin = nparts;
    if (nparts >= kTwoLevelMin && nparts + kPartialGroupsMax <= max_partials) {
      const int chunk = (nparts + kPartialGroupsMax - 1) / kPartialGroupsMax;
      const int groups = (nparts + chunk - 1) / chunk;
      hipLaunchKernelGGL((partials_reduce_kernel<256>), dim3(groups), dim3(256), 0, stream, partials,
                         nparts, chunk, partials + nparts);
      fin = partials + nparts;
      nfin = groups;
    }
    if (nfin <= 256) {
      hipLaunchKernelGGL((finalize_kernel<256>), dim3(1), dim3(256), 0, stream, fin, nfin, ctrl,
                         device_dt ? 1 : 0, a.dt_log);
    } else {
      hipLaunchKernelGGL((finalize_kernel<kFinalizeThreads>), dim3(1), dim3(kFinalizeThreads), 0, stream,
                         fin, nfin, ctrl, device_dt ? 1 : 0, a.dt_log);
    }
  }
  return hipGetLastError() == hipSuccess ? EC_OK : EC_ERR_HIP;
}

// device_dt: 0 host dt, 1 device-loop dt, 2 host dt without the diagnostics
// epilogue (state only; the run control is left untouched).
ec_status launch_explicit_step_on(ec_ctx* ctx, hipStream_t stream, const double* uin, double* uout,
                                  long long rs, long long cs, double dt, int device_dt)
{
  if (tiles_needed(ctx->grid.nx, ctx->grid.ny, ctx->tile_id) > ctx->max_partials) {
    ctx->last_error = "partials buffer too small for the tile shape";
    return EC_ERR_INVALID_ARGUMENT;
  }
  const bool timed =
      ctx->timing && stream == ctx->stream && ctx->tused + 2 <= (int)ctx->tev.size();
  hipEvent_t e0 = timed ? ctx->tev[ctx->tused] : nullptr;
  hipEvent_t e1 = timed ? ctx->tev[ctx->tused + 1] : nullptr;
  const ec_status st = launch_step_raw(ctx->grid, ctx->tile_id, stream, uin, uout, rs, cs, dt, device_dt,
                                       ctx->ctrl, ctx->partials, ctx->max_partials, ctx->dt_log, e0, e1);
  if (st == EC_ERR_HIP) {
    ctx->last_error = std::string("step launch: ") + hipGetErrorString(hipGetLastError());
    return st;
  }
  if (st == EC_OK && timed) ctx->tused += 2;
  return st;
}

ec_status launch_explicit_step(ec_ctx* ctx, const double* uin, double* uout, long long rs,
                               long long cs, double dt, int device_dt)
{
  return launch_explicit_step_on(ctx, ctx->stream, uin, uout, rs, cs, dt, device_dt);
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

extern "C" ec_status ec_step_kernel_timing(ec_ctx* ctx, int enable, int max_launches)
{
  if (!ctx || max_launches < 0) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  EC_HIP_TRY(ctx, ctx->stream ? hipStreamSynchronize(ctx->stream) : hipDeviceSynchronize());
  for (hipEvent_t e : ctx->tev) (void)hipEventDestroy(e);
  ctx->tev.clear();
  ctx->tused = 0;
  ctx->timing = enable != 0;
  if (!ctx->timing) return EC_OK;
  ctx->tev.resize(2 * (size_t)max_launches, nullptr);
  for (hipEvent_t& e : ctx->tev) EC_HIP_TRY(ctx, hipEventCreate(&e));
  return EC_OK;
}

extern "C" ec_status ec_step_kernel_time(ec_ctx* ctx, double* total_ms, int* launches)
{
  if (!ctx || !total_ms || !launches) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  double sum = 0.0;
  for (int k = 0; k + 1 < ctx->tused; k += 2) {
    EC_HIP_TRY(ctx, hipEventSynchronize(ctx->tev[k + 1]));
    float ms = 0.f;
    EC_HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->tev[k], ctx->tev[k + 1]));
    sum += ms;
  }
  *total_ms = sum;
  *launches = ctx->tused / 2;
  ctx->tused = 0;
  return EC_OK;
}
