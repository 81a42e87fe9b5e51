// ec_batch.hip -- ensembles of independent explicit runs of uneven
// resolution (Riemann strips, Gresho boxes, ...) stepped together on one GPU.
//
// Every member follows the explicit branch of the driver loop
// (app/driver_main.hpp:216-333) exactly as ec_explicit_advance does on the
// member alone: the same tile arithmetic (ec_tile.hpp), the same dt sequence
// and statuses, so each member's result is bit-identical to its single-grid
// run.  The batch is a list of tiles (member-major); a "round" advances every
// member that is still running by one step:
//   1. fold: the member's previous-round partials (one per tile) are reduced
//      and folded into the member's run control, redundantly by every
//      workgroup of the member (max/min/integer sums: order independent);
//   2. step: the workgroup's tiles of the member are advanced by the dt the
//      fold yields, writing one partial per tile;
//   3. publish: the workgroup owning the member's first tile writes the run
//      control for the next round.
// Partials and run controls are double-buffered by round parity, and each
// member ping-pongs its state between buffers A and B by step parity, so a
// round reads only what the previous round wrote.  Two ways of separating
// rounds:
//   graph:      one kernel launch per round (kernel boundaries order the
//               rounds); S rounds are captured in a hipGraph and replayed;
//   persistent: one cooperative launch in which each workgroup loops over
//               rounds, with a per-member arrival counter between rounds (the
//               workgroups of one member wait only for each other).  Every
//               wait is bounded: a timed-out wait raises an error flag that
//               makes every workgroup leave, so the kernel always drains.
// A project-written autotuner times every (mode, tile shape) pair on a copy
// of the initial states and keeps the fastest.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ec_exact_api.hpp"
#include "ec_tile.hpp"

namespace ec {

// A round of an ensemble is one tile per workgroup and waits on latency, not
// on issue: the tile keeps all eight slope planes (the split form adds two
// barriers) and loads the cells of its stage E with those of stage A.
template <int TX, int TY, int NT, int RIEMANN, int ORDER>
using BatchTile = TileStep<TX, TY, NT, RIEMANN, ORDER, /*SPLIT_SLOPES=*/false, /*PREFETCH_OWN=*/true>;

struct MemberDev {
  GridParams g;
  long long off;   // first double of the member's state in each buffer
  int tiles_x;     // tiles per row of the member for the active shape
  int ntiles;
  int tile0;       // global index of the member's first tile
  int pad;
};

// Everything a round needs about one tile, in one record (one dependent load).
struct TileInfo {
  MemberDev md;
  int m;   // member
  int lt;  // member-local tile index
};

struct BatchArgs {
  const MemberDev* members;
  const TileInfo* tiles;   // per global tile
  const TileInfo* round_tiles;  // graph mode: the tile of each workgroup of one group's launch
  const int* tile_member;  // global tile -> member
  double* ua;
  double* ub;
  DevCtrl* ctrl;           // [2][nm]: counter holds the pending fold mode (0, 1 or 3)
  Partial* part;           // [2][ntiles]
  int nm, ntiles;
  int* error;              // persistent mode: set by a timed-out wait
  unsigned int* arrive;    // persistent mode: per-member arrival counters
  unsigned int spin_limit;
  long long max_rounds;    // persistent mode: rounds no run can exceed (largest max_step + 1)
};

struct RoundShared {
  DevCtrl c;
  double dt;
  int active;
  int abort;
};

// Epilogue reduction of the tile's current state (initial diagnostics).
template <int TX, int TY, int NT>
__device__ Red tile_state_reduce(Red* scratch, const double* __restrict__ u, long long rs,
                                 long long cs, const GridParams& g, int i0, int j0)
{
  Red r = red_identity();
  for (int t = threadIdx.x; t < TX * TY; t += NT) {
    const int gi = i0 + t % TX, gj = j0 + t / TX;
    if (gi < g.nx && gj < g.ny) {
      const long long k = (long long)gj * rs + gi;
      red_cell(r, Cons{u[k], u[cs + k], u[2 * cs + k], u[3 * cs + k]}, g);
    }
  }
  return block_reduce<NT>(r, scratch);
}

// One round of member m for the tiles [ta, tb) (member-local indices) owned
// by this workgroup.  Returns whether the member stepped (so another round
// of it follows).
template <int TX, int TY, int NT, int RIEMANN, int ORDER>
__device__ bool member_round(typename BatchTile<TX, TY, NT, RIEMANN, ORDER>::Smem& sm,
                             RoundShared& rsh, const BatchArgs& A, const MemberDev& md, int m, int ta,
                             int tb, int p)
{
  using T = BatchTile<TX, TY, NT, RIEMANN, ORDER>;
  const int tid = threadIdx.x;
  const DevCtrl* cin = A.ctrl + (long long)p * A.nm + m;
  const long long rs = md.g.nx, cs = (long long)md.g.nx * md.g.ny;

  // 0. stage A's loads of the first tile, issued before the fold so that the
  // two wait as one: every member steps once per round from step 0 while it
  // runs, so a running member's state at the round of parity p is in buffer p
  // (step & 1 == p).  The step is checked after the fold, and a tile whose
  // state is elsewhere is loaded again.
  typename T::Loaded ld;
  const double* guess = (p ? A.ub : A.ua) + md.off;
  if (ta < tb) T::load(ld, guess, rs, cs, md.g, (ta % md.tiles_x) * TX, (ta / md.tiles_x) * TY);

  // 1. fold the previous round's partials of every tile of the member.  The
  // partials are read whether or not a fold is pending (the buffer always
  // holds finite data), so their loads overlap the control block's.
  Red r = red_identity();
  {
    const Partial* prev = A.part + (long long)(p ^ 1) * A.ntiles + md.tile0;
    for (int q = tid; q < md.ntiles; q += NT) {
      const Partial v = prev[q];
      red_combine(r, Red{v.rate, v.rho_min, v.p_min, v.bad, v.degenerate});
    }
  }
  r = block_reduce<NT>(r, sm.red_scratch);
  if (tid == 0) {
    DevCtrl c = *cin;
    const int pending = c.counter;
    if (pending != 0) ctrl_fold(c, r, pending, nullptr);
    double dt;
    bool active, dt_bad;
    loop_control(&c, dt, active, dt_bad);
    // A bad dt ends the run with invalid_time_step, as the finalize of the
    // single-grid loop records it.
    if (!active && dt_bad) c.status = 2;
    c.counter = active ? 1 : 0;
    rsh.c = c;
    rsh.dt = dt;
    rsh.active = active ? 1 : 0;
  }
  __syncthreads();
  const bool active = rsh.active != 0;
  const double dt = rsh.dt;
  const int step = rsh.c.step;

  // 2. step the owned tiles
  if (active) {
    const double* uin = ((step & 1) ? A.ub : A.ua) + md.off;
    double* uout = ((step & 1) ? A.ua : A.ub) + md.off;
    Partial* out = A.part + (long long)p * A.ntiles + md.tile0;
    for (int t = ta; t < tb; ++t) {
      const int i0 = (t % md.tiles_x) * TX, j0 = (t / md.tiles_x) * TY;
      if (t != ta || uin != guess) T::load(ld, uin, rs, cs, md.g, i0, j0);
      // exact-Dirichlet ghosts take the step's start time, held in LDS
      const Red rr = T::run_loaded(sm, ld, uin, uout, rs, cs, md.g, i0, j0, dt, true, &rsh.c.time);
      store_partial(out, t, rr);
      __syncthreads();
    }
  }

  // 3. publish the run control for the next round
  if (tid == 0 && ta == 0) A.ctrl[(long long)(p ^ 1) * A.nm + m] = rsh.c;
  __syncthreads();
  return active;
}

// Initial diagnostics of every tile (the reference's compute_diagnostics at
// app/driver_main.hpp:216-222), as round -1: partials into parity 1.
template <int TX, int TY, int NT>
__global__ void __launch_bounds__(NT) batch_init_kernel(BatchArgs A)
{
  __shared__ Red scratch[NT / 64];
  const int t = blockIdx.x;
  const int m = A.tile_member[t];
  const MemberDev& md = A.members[m];
  const int lt = t - md.tile0;
  const int i0 = (lt % md.tiles_x) * TX, j0 = (lt / md.tiles_x) * TY;
  const Red r = tile_state_reduce<TX, TY, NT>(scratch, A.ua + md.off, md.g.nx,
                                              (long long)md.g.nx * md.g.ny, md.g, i0, j0);
  store_partial(A.part + A.ntiles, t, r);
}

// Graph mode: one launch = one round of one member group; workgroup b owns
// the tile round_tiles[b].  Between replays the host moves the tiles of
// finished members to the end of the table, marked m = -1.
template <int TX, int TY, int NT, int RIEMANN, int ORDER>
__global__ void __launch_bounds__(NT) batch_round_kernel(BatchArgs A, int p)
{
  using T = BatchTile<TX, TY, NT, RIEMANN, ORDER>;
  __shared__ typename T::Smem sm;
  __shared__ RoundShared rsh;
  const TileInfo& ti = A.round_tiles[blockIdx.x];
  if (ti.m < 0) return;  // a member that had finished when this replay began
  (void)member_round<TX, TY, NT, RIEMANN, ORDER>(sm, rsh, A, ti.md, ti.m, ti.lt, ti.lt + 1, p);
}

// Waits until the member's arrival counter reaches target.  Bounded: after
// spin_limit polls, or when another workgroup has raised the error flag, it
// raises the flag and reports failure.  The workgroup barrier orders the
// other threads' stores before thread 0's agent-scope release fence and
// arrival, and thread 0's acquire fence after the wait before every
// thread's loads (one fence pair per workgroup: each agent-scope release
// writes the XCD's L2 back, and a fence in every wavefront cost 1.5x the
// round time).
__device__ bool member_barrier(const BatchArgs& A, RoundShared& rsh, int m, unsigned add,
                               unsigned target)
{
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    rsh.abort = 0;
    atomicAdd(A.arrive + m, add);
    unsigned spins = 0;
    while (__hip_atomic_load(A.arrive + m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(A.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
          ++spins > A.spin_limit) {
        atomicOr(A.error, 1);
        rsh.abort = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __threadfence();
  }
  __syncthreads();
  return rsh.abort == 0;
}

// Persistent mode: workgroup b owns the global tiles [range.x, range.y),
// visited member by member in increasing member order every round, so the
// per-member waits can never form a cycle.
template <int TX, int TY, int NT, int RIEMANN, int ORDER>
__global__ void __launch_bounds__(NT) batch_persistent_kernel(BatchArgs A, const int2* ranges)
{
  using T = BatchTile<TX, TY, NT, RIEMANN, ORDER>;
  __shared__ typename T::Smem sm;
  __shared__ RoundShared rsh;
  constexpr int kMaxOwned = 64;   // members one workgroup may hold (host-checked)
  __shared__ unsigned char done[kMaxOwned];
  const int2 range = ranges[blockIdx.x];
  if (range.x >= range.y) return;
  const int first_member = A.tile_member[range.x];
  for (int k = threadIdx.x; k < kMaxOwned; k += NT) done[k] = 0;
  __syncthreads();
  for (long long round = 0;; ++round) {
    // every member has stopped by then: a run control that still reads
    // "running" is an error, and the workgroup leaves (so the launch drains)
    if (round > A.max_rounds) {
      if (threadIdx.x == 0) atomicOr(A.error, 1);
      return;
    }
    const int p = (int)(round & 1);
    bool any = false;
    int t = range.x;
    while (t < range.y) {
      const int m = A.tile_member[t];
      const MemberDev& md = A.members[m];
      const int tb = min(range.y, md.tile0 + md.ntiles);
      const int slot = m - first_member;
      if (!done[slot]) {
        const bool stepped = member_round<TX, TY, NT, RIEMANN, ORDER>(sm, rsh, A, md, m, t - md.tile0,
                                                                      tb - md.tile0, p);
        if (stepped) {
          any = true;
          // Steps taken so far by the member, counting this one.
          const unsigned target = (unsigned)md.ntiles * (unsigned)(rsh.c.step + 1);
          if (!member_barrier(A, rsh, m, (unsigned)(tb - t), target)) return;
        } else if (threadIdx.x == 0) {
          done[slot] = 1;
        }
        __syncthreads();
      }
      t = tb;
    }
    if (!any) return;
  }
}

// Copies the members whose final state sits in buffer B back to buffer A.
__global__ void batch_gather_kernel(const MemberDev* members, const DevCtrl* final_ctrl,
                                    const double* __restrict__ ub, double* __restrict__ ua)
{
  const int m = blockIdx.y;
  if ((final_ctrl[m].step & 1) == 0) return;
  const MemberDev& md = members[m];
  const long long n = 4LL * md.g.nx * md.g.ny;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (long long)gridDim.x * blockDim.x) {
    ua[md.off + k] = ub[md.off + k];
  }
}

// Picks, per member, the run-control slot of its last round (the one whose
// pending mode is 0) into out[m].
__global__ void batch_final_ctrl_kernel(const DevCtrl* ctrl, int nm, DevCtrl* out)
{
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const DevCtrl& a = ctrl[m];
  const DevCtrl& b = ctrl[nm + m];
  out[m] = (a.counter == 0 && (b.counter != 0 || a.step >= b.step)) ? a : b;
}

// ---------------------------------------------------------------- dispatch
struct BatchShape {
  int tx, ty, nt;
};
// 48x4: both face sets of a 4-row strip tile in one pass of 256 threads
// (196 x-faces, 240 y-faces); 60x4 needs a second y pass but has fewer halo
// columns per cell; 16x15 suits 2D members.  24x4: the y-faces get wavefronts
// of their own (TileStep::YOFF; 100 x-faces on wavefronts 0-1, 120 y-faces on
// 2-3), so each wavefront computes one face, and stage A has one cell per
// thread: the shortest chain through a round that waits on its slowest tile.
static const BatchShape kBatchShapes[] = {{60, 4, 256}, {124, 4, 512}, {16, 15, 256}, {48, 4, 256}, {24, 4, 256}};
constexpr int kNumBatchShapes = sizeof(kBatchShapes) / sizeof(kBatchShapes[0]);
static_assert(BatchTile<24, 4, 256, 3, 2>::YOFF == 128, "24x4/256: y-faces on wavefronts 2-3");

struct BatchFns {
  const void* round;       // batch_round_kernel
  const void* persistent;  // batch_persistent_kernel
  const void* init;        // batch_init_kernel
};

template <int TX, int TY, int NT, int R, int O>
static BatchFns fns_of()
{
  return BatchFns{(const void*)batch_round_kernel<TX, TY, NT, R, O>,
                  (const void*)batch_persistent_kernel<TX, TY, NT, R, O>,
                  (const void*)batch_init_kernel<TX, TY, NT>};
}

template <int TX, int TY, int NT>
static BatchFns fns_for_shape(int riemann, int order)
{
  const int k = riemann * 2 + (order == 2 ? 1 : 0);
  switch (k) {
    case 0: return fns_of<TX, TY, NT, 0, 1>();
    case 1: return fns_of<TX, TY, NT, 0, 2>();
    case 2: return fns_of<TX, TY, NT, 1, 1>();
    case 3: return fns_of<TX, TY, NT, 1, 2>();
    case 4: return fns_of<TX, TY, NT, 2, 1>();
    case 5: return fns_of<TX, TY, NT, 2, 2>();
    case 6: return fns_of<TX, TY, NT, 3, 1>();
    default: return fns_of<TX, TY, NT, 3, 2>();
  }
}

static BatchFns batch_fns(int shape, int riemann, int order)
{
  switch (shape) {
    case 1: return fns_for_shape<124, 4, 512>(riemann, order);
    case 2: return fns_for_shape<16, 15, 256>(riemann, order);
    case 3: return fns_for_shape<48, 4, 256>(riemann, order);
    case 4: return fns_for_shape<24, 4, 256>(riemann, order);
    default: return fns_for_shape<60, 4, 256>(riemann, order);
  }
}

}  // namespace ec

using namespace ec;

constexpr int kMaxGroups = 4;

struct ec_batch {
  int device = 0;
  hipStream_t user_stream = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev_user = nullptr, ev_a = nullptr, ev_b = nullptr;
  int nm = 0;
  std::vector<ec_member_desc> desc;
  double gamma = 1.4;
  int riemann = 1, limiter = 0, order = 2;
  std::vector<long long> off;
  long long total = 0;       // doubles per buffer
  double* ua = nullptr;
  double* ub = nullptr;
  double* backup = nullptr;  // autotuner copy of the initial states
  DevCtrl* ctrl = nullptr;   // [2][nm]
  DevCtrl* final_ctrl = nullptr;  // [nm]
  DevCtrl* ctrl_host = nullptr;   // pinned [2][nm]
  Partial* part = nullptr;   // [2][max tiles]
  int part_cap = 0;
  int* error = nullptr;
  unsigned* arrive = nullptr;
  MemberDev* members = nullptr;
  int* tile_member = nullptr;
  TileInfo* tiles = nullptr;
  int2* ranges = nullptr;
  int ranges_cap = 0;
  int mode = EC_BATCH_AUTO;
  int shape = -1;
  int rounds_per_graph = 32;
  int cfg_shape = -1;        // shape the device tables hold
  int ntiles = 0;
  int graph_shape = -1;
  int graph_rounds = 0;      // rounds the instantiated graph holds
  // Graph mode splits the members into groups (balanced by tile count), each
  // group's rounds captured in its own graph and replayed on its own stream:
  // the groups' rounds overlap (one group's loads with another's arithmetic,
  // one group's launch gap with another's round).  gstream[0] is `stream`.
  int groups = 2;
  int cfg_groups = 0;        // groups the device tables hold
  int graph_groups = 0;      // groups the instantiated graphs hold
  hipStream_t gstream[kMaxGroups] = {};
  hipEvent_t gev[kMaxGroups] = {};
  TileInfo* gtiles[kMaxGroups] = {};
  int gblocks[kMaxGroups] = {};
  // host copy of each group's full tile table (member order), and the table
  // last uploaded (kept until the copy has completed: every replay ends with
  // a stream synchronisation)
  std::vector<TileInfo> gtiles_host[kMaxGroups];
  std::vector<TileInfo> gtiles_up[kMaxGroups];
  hipGraphExec_t ggraph[kMaxGroups] = {};
  int last_mode = 0, last_shape = 0, last_rounds = 0;
  double last_ms = 0.0;
  unsigned spin_limit = 1u << 20;
  std::vector<ec::ExactDevice> exact;  // [nm]: exact-Dirichlet sides of a member
  std::string last_error;
};

namespace {

ec_status fail(ec_batch* b, hipError_t e, const char* what)
{
  b->last_error = std::string(what) + ": " + hipGetErrorString(e);
  return EC_ERR_HIP;
}

#define BT_TRY(b, expr)                                  \
  do {                                                   \
    hipError_t _e = (expr);                              \
    if (_e != hipSuccess) return fail((b), _e, #expr);   \
  } while (0)

bool member_ok(const ec_member_desc& d)
{
  if (d.nx < 1 || d.ny < 1 || !(d.dx > 0.0) || !(d.dy > 0.0)) return false;
  if (!std::isfinite(d.dx) || !std::isfinite(d.dy)) return false;
  for (int s = 0; s < 4; ++s)
    if (d.bc[s] < 0 || d.bc[s] > 3) return false;
  if ((d.bc[0] == BC_PERIODIC) != (d.bc[1] == BC_PERIODIC)) return false;
  if ((d.bc[2] == BC_PERIODIC) != (d.bc[3] == BC_PERIODIC)) return false;
  if (!(d.cfl > 0.0) || !std::isfinite(d.stop_time) || d.max_step < 0) return false;
  return true;
}

void drop_graphs(ec_batch* b)
{
  for (auto& g : b->ggraph) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  b->graph_shape = -1;
}

int tiles_of(const ec_member_desc& d, int shape, int* tx_out)
{
  const BatchShape s = kBatchShapes[shape];
  const int tx = (d.nx + s.tx - 1) / s.tx, ty = (d.ny + s.ty - 1) / s.ty;
  if (tx_out) *tx_out = tx;
  return tx * ty;
}

// Device tables (member descriptors, tile -> member) for one tile shape.
ec_status configure_tables(ec_batch* b, int shape)
{
  if (b->cfg_shape == shape && b->cfg_groups == b->groups) return EC_OK;
  std::vector<MemberDev> md(b->nm);
  std::vector<int> tm;
  for (int m = 0; m < b->nm; ++m) {
    const ec_member_desc& d = b->desc[m];
    ec_grid_desc gd{};
    gd.nx = d.nx;
    gd.ny = d.ny;
    gd.dx = d.dx;
    gd.dy = d.dy;
    gd.gamma = b->gamma;
    gd.riemann = b->riemann;
    gd.limiter = b->limiter;
    gd.spatial_order = b->order;
    for (int s = 0; s < 4; ++s) {
      gd.bc[s] = d.bc[s];
      for (int c = 0; c < 4; ++c) gd.fixed_state[s][c] = d.fixed_state[s][c];
    }
    md[m].g = make_grid_params(gd);
    md[m].g.exact = m < (int)b->exact.size() ? b->exact[m].desc : nullptr;
    md[m].g.time_ptr = nullptr;
    md[m].off = b->off[m];
    md[m].ntiles = tiles_of(d, shape, &md[m].tiles_x);
    md[m].tile0 = (int)tm.size();
    md[m].pad = 0;
    for (int t = 0; t < md[m].ntiles; ++t) tm.push_back(m);
  }
  const int nt = (int)tm.size();
  if (nt > b->part_cap) {
    b->last_error = "tile count exceeds the partials buffer";
    return EC_ERR_INVALID_ARGUMENT;
  }
  if (b->tile_member) (void)hipFree(b->tile_member);
  b->tile_member = nullptr;
  if (b->tiles) (void)hipFree(b->tiles);
  b->tiles = nullptr;
  BT_TRY(b, hipMalloc(&b->tile_member, sizeof(int) * (size_t)nt));
  BT_TRY(b, hipMalloc(&b->tiles, sizeof(TileInfo) * (size_t)nt));
  std::vector<TileInfo> ti(nt);
  for (int t = 0; t < nt; ++t) {
    ti[t].md = md[tm[t]];
    ti[t].m = tm[t];
    ti[t].lt = t - md[tm[t]].tile0;
  }
  // Member groups of the graph mode: largest member first, to the group with
  // the fewest tiles; each group's workgroups take its tiles member by member.
  const int G = std::min(b->groups, b->nm);
  std::vector<int> order(b->nm);
  for (int m = 0; m < b->nm; ++m) order[m] = m;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return md[x].ntiles > md[y].ntiles; });
  std::vector<std::vector<int>> members_of(G);
  std::vector<int> load(G, 0);
  for (int m : order) {
    const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    members_of[g].push_back(m);
    load[g] += md[m].ntiles;
  }
  for (int g = 0; g < kMaxGroups; ++g) {
    if (b->gtiles[g]) (void)hipFree(b->gtiles[g]);
    b->gtiles[g] = nullptr;
    b->gblocks[g] = 0;
    b->gtiles_host[g].clear();
  }
  for (int g = 0; g < G; ++g) {
    std::sort(members_of[g].begin(), members_of[g].end());
    std::vector<TileInfo>& gt = b->gtiles_host[g];
    gt.clear();
    for (int m : members_of[g])
      for (int t = 0; t < md[m].ntiles; ++t) gt.push_back(ti[md[m].tile0 + t]);
    BT_TRY(b, hipMalloc(&b->gtiles[g], sizeof(TileInfo) * gt.size()));
    BT_TRY(b, hipMemcpyAsync(b->gtiles[g], gt.data(), sizeof(TileInfo) * gt.size(),
                             hipMemcpyHostToDevice, b->stream));
    b->gblocks[g] = (int)gt.size();
  }
  BT_TRY(b, hipMemcpyAsync(b->members, md.data(), sizeof(MemberDev) * (size_t)b->nm,
                           hipMemcpyHostToDevice, b->stream));
  BT_TRY(b, hipMemcpyAsync(b->tile_member, tm.data(), sizeof(int) * (size_t)nt,
                           hipMemcpyHostToDevice, b->stream));
  BT_TRY(b, hipMemcpyAsync(b->tiles, ti.data(), sizeof(TileInfo) * (size_t)nt, hipMemcpyHostToDevice,
                           b->stream));
  BT_TRY(b, hipStreamSynchronize(b->stream));
  b->ntiles = nt;
  b->cfg_shape = shape;
  b->cfg_groups = b->groups;
  drop_graphs(b);
  return EC_OK;
}

BatchArgs make_args(ec_batch* b)
{
  BatchArgs a;
  a.members = b->members;
  a.tiles = b->tiles;
  a.round_tiles = b->gtiles[0];
  a.tile_member = b->tile_member;
  a.ua = b->ua;
  a.ub = b->ub;
  a.ctrl = b->ctrl;
  a.part = b->part;
  a.nm = b->nm;
  a.ntiles = b->ntiles;
  a.error = b->error;
  a.arrive = b->arrive;
  a.spin_limit = b->spin_limit;
  long long mx = 0;
  for (const auto& d : b->desc) mx = std::max<long long>(mx, d.max_step);
  a.max_rounds = mx + 1;
  return a;
}

// Run controls of a fresh run (app/driver_main.hpp:216-222 start values),
// both parities, pending the initial diagnostics.
ec_status reset_run(ec_batch* b)
{
  for (int q = 0; q < 2; ++q) {
    for (int m = 0; m < b->nm; ++m) {
      DevCtrl& c = b->ctrl_host[q * b->nm + m];
      std::memset(&c, 0, sizeof(DevCtrl));
      c.stop_time = b->desc[m].stop_time;
      c.cfl = b->desc[m].cfl;
      c.max_step = b->desc[m].max_step;
      c.dt_min = INFINITY;
      c.counter = q == 0 ? 3 : 0;
    }
  }
  BT_TRY(b, hipMemcpyAsync(b->ctrl, b->ctrl_host, sizeof(DevCtrl) * 2 * (size_t)b->nm,
                           hipMemcpyHostToDevice, b->stream));
  BatchArgs a = make_args(b);
  void* args[] = {&a};
  const BatchFns f = batch_fns(b->cfg_shape, b->riemann, b->order);
  BT_TRY(b, hipLaunchKernel(f.init, dim3(b->ntiles), dim3(kBatchShapes[b->cfg_shape].nt), args, 0,
                            b->stream));
  return EC_OK;
}

long long round_bound(const ec_batch* b)
{
  long long mx = 0;
  for (const auto& d : b->desc) mx = std::max<long long>(mx, d.max_step);
  return mx + 2;
}

// Uploads each group's tile table with the tiles of members marked in
// `finished` moved to the end as m = -1 (their workgroups exit at once); the
// order of the other tiles is kept.  The copies are ordered before the next
// replay on `stream`, which every group's stream waits for.
static ec_status upload_round_tables(ec_batch* b, const std::vector<char>& finished)
{
  for (int g = 0; g < kMaxGroups; ++g) {
    const std::vector<TileInfo>& full = b->gtiles_host[g];
    if (full.empty() || !b->gtiles[g]) continue;
    std::vector<TileInfo>& t = b->gtiles_up[g];
    t.clear();
    t.reserve(full.size());
    for (const TileInfo& x : full)
      if (!finished[x.m]) t.push_back(x);
    while (t.size() < full.size()) {
      TileInfo x = full[0];
      x.m = -1;
      t.push_back(x);
    }
    BT_TRY(b, hipMemcpyAsync(b->gtiles[g], t.data(), sizeof(TileInfo) * t.size(), hipMemcpyHostToDevice,
                             b->stream));
  }
  return EC_OK;
}

ec_status run_graph(ec_batch* b, int* rounds_out)
{
  const BatchShape s = kBatchShapes[b->cfg_shape];
  const BatchFns f = batch_fns(b->cfg_shape, b->riemann, b->order);
  const int R = b->rounds_per_graph;
  if (b->graph_shape != b->cfg_shape || b->graph_rounds != R || b->graph_groups != b->cfg_groups) {
    drop_graphs(b);
    for (int gi = 0; gi < kMaxGroups; ++gi) {
      if (b->gblocks[gi] == 0) continue;
      hipStream_t st = b->gstream[gi];
      hipGraph_t g = nullptr;
      BatchArgs a = make_args(b);
      a.round_tiles = b->gtiles[gi];
      BT_TRY(b, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < R; ++k) {
        int p = k & 1;
        void* args[] = {&a, &p};
        const hipError_t e = hipLaunchKernel(f.round, dim3(b->gblocks[gi]), dim3(s.nt), args, 0, st);
        if (e != hipSuccess) {
          (void)hipStreamEndCapture(st, &g);
          if (g) (void)hipGraphDestroy(g);
          return fail(b, e, "hipLaunchKernel(batch_round_kernel) under capture");
        }
      }
      BT_TRY(b, hipStreamEndCapture(st, &g));
      const hipError_t e = hipGraphInstantiate(&b->ggraph[gi], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(b, e, "hipGraphInstantiate");
    }
    b->graph_shape = b->cfg_shape;
    b->graph_rounds = R;
    b->graph_groups = b->cfg_groups;
  }
  // After an even number of rounds the latest run controls are in parity 0;
  // a member is finished when its latest round did not step it.  Finished
  // members leave the round tables at the next replay (their final control
  // stays in parity 0, which no later round writes).
  const long long bound = round_bound(b);
  long long rounds = 0;
  std::vector<char> finished(b->nm, 0);
  ec_status st = upload_round_tables(b, finished);
  if (st != EC_OK) return st;
  for (;;) {
    // the groups' streams start after `stream` and `stream` waits for them
    BT_TRY(b, hipEventRecord(b->gev[0], b->stream));
    for (int gi = 1; gi < kMaxGroups; ++gi)
      if (b->ggraph[gi]) BT_TRY(b, hipStreamWaitEvent(b->gstream[gi], b->gev[0], 0));
    for (int gi = 0; gi < kMaxGroups; ++gi)
      if (b->ggraph[gi]) BT_TRY(b, hipGraphLaunch(b->ggraph[gi], b->gstream[gi]));
    for (int gi = 1; gi < kMaxGroups; ++gi) {
      if (!b->ggraph[gi]) continue;
      BT_TRY(b, hipEventRecord(b->gev[gi], b->gstream[gi]));
      BT_TRY(b, hipStreamWaitEvent(b->stream, b->gev[gi], 0));
    }
    rounds += R;
    BT_TRY(b, hipMemcpyAsync(b->ctrl_host, b->ctrl, sizeof(DevCtrl) * (size_t)b->nm,
                             hipMemcpyDeviceToHost, b->stream));
    BT_TRY(b, hipStreamSynchronize(b->stream));
    bool done = true, changed = false;
    for (int m = 0; m < b->nm; ++m) {
      const bool fin = b->ctrl_host[m].counter == 0;
      done = done && fin;
      if (fin && !finished[m]) {
        finished[m] = 1;
        changed = true;
      }
    }
    if (done) break;
    if (changed && (st = upload_round_tables(b, finished)) != EC_OK) return st;
    if (rounds > bound + R) {
      b->last_error = "ensemble members still running past their max_step";
      return EC_ERR_HIP;
    }
  }
  *rounds_out = (int)rounds;
  return EC_OK;
}

ec_status run_persistent(ec_batch* b, int* rounds_out)
{
  const BatchShape s = kBatchShapes[b->cfg_shape];
  const BatchFns f = batch_fns(b->cfg_shape, b->riemann, b->order);
  int per_cu = 0, cus = 0;
  BT_TRY(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f.persistent, s.nt, 0));
  BT_TRY(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  const int cap = per_cu * cus;
  if (cap < 1) {
    b->last_error = "persistent ensemble kernel cannot be resident";
    return EC_ERR_UNSUPPORTED;
  }
  const int G = std::min(b->ntiles, cap);
  // Contiguous, member-major tile ranges of near-equal size.
  std::vector<int2> ranges(G);
  std::vector<int> tm(b->ntiles);
  {
    int t = 0;
    for (int m = 0; m < b->nm; ++m) {
      const int n = tiles_of(b->desc[m], b->cfg_shape, nullptr);
      for (int k = 0; k < n; ++k) tm[t++] = m;
    }
  }
  for (int w = 0; w < G; ++w) {
    const int lo = (int)((long long)b->ntiles * w / G), hi = (int)((long long)b->ntiles * (w + 1) / G);
    ranges[w] = make_int2(lo, hi);
    if (hi > lo && tm[hi - 1] - tm[lo] >= 64) {
      b->last_error = "too many members per persistent workgroup";
      return EC_ERR_UNSUPPORTED;
    }
  }
  if (G > b->ranges_cap) {
    if (b->ranges) (void)hipFree(b->ranges);
    b->ranges = nullptr;
    BT_TRY(b, hipMalloc(&b->ranges, sizeof(int2) * (size_t)G));
    b->ranges_cap = G;
  }
  BT_TRY(b, hipMemcpyAsync(b->ranges, ranges.data(), sizeof(int2) * (size_t)G,
                           hipMemcpyHostToDevice, b->stream));
  BT_TRY(b, hipMemsetAsync(b->arrive, 0, sizeof(unsigned) * (size_t)b->nm, b->stream));
  BT_TRY(b, hipMemsetAsync(b->error, 0, sizeof(int), b->stream));
  BatchArgs a = make_args(b);
  int2* rp = b->ranges;
  void* args[] = {&a, &rp};
  BT_TRY(b, hipLaunchCooperativeKernel(f.persistent, dim3(G), dim3(s.nt), args, 0, b->stream));
  int err = 0;
  BT_TRY(b, hipMemcpyAsync(&b->ctrl_host[0].counter, b->error, sizeof(int), hipMemcpyDeviceToHost,
                           b->stream));
  BT_TRY(b, hipStreamSynchronize(b->stream));
  err = (int)b->ctrl_host[0].counter;
  if (err != 0) {
    b->last_error = "persistent ensemble: a member wait timed out, or a member ran past every max_step";
    return EC_ERR_HIP;
  }
  *rounds_out = 0;
  return EC_OK;
}

// One complete run of every member from the states in buffer A, leaving the
// final states in buffer A and the run controls in final_ctrl.
ec_status run_once(ec_batch* b, int mode, int shape, double* ms)
{
  ec_status st = configure_tables(b, shape);
  if (st != EC_OK) return st;
  BT_TRY(b, hipEventRecord(b->ev_a, b->stream));
  if ((st = reset_run(b)) != EC_OK) return st;
  int rounds = 0;
  st = mode == EC_BATCH_PERSISTENT ? run_persistent(b, &rounds) : run_graph(b, &rounds);
  if (st != EC_OK) return st;
  hipLaunchKernelGGL(batch_final_ctrl_kernel, dim3((b->nm + 63) / 64), dim3(64), 0, b->stream,
                     b->ctrl, b->nm, b->final_ctrl);
  hipLaunchKernelGGL(batch_gather_kernel, dim3(64, b->nm), dim3(256), 0, b->stream, b->members,
                     b->final_ctrl, b->ub, b->ua);
  BT_TRY(b, hipGetLastError());
  BT_TRY(b, hipEventRecord(b->ev_b, b->stream));
  BT_TRY(b, hipEventSynchronize(b->ev_b));
  float t = 0.f;
  BT_TRY(b, hipEventElapsedTime(&t, b->ev_a, b->ev_b));
  if (ms) *ms = t;
  b->last_mode = mode;
  b->last_shape = shape;
  b->last_rounds = rounds;
  b->last_ms = t;
  return EC_OK;
}

void to_run_state(const DevCtrl& c, ec_run_state* r)
{
  r->time = c.time;
  r->stop_time = c.stop_time;
  r->cfl = c.cfl;
  r->dt_last = c.dt_last;
  r->dt_min = c.dt_min;
  r->dt_max = c.dt_max;
  r->dt_sum = c.dt_sum;
  r->rho_min = c.rho_min;
  r->pressure_min = c.p_min;
  r->max_rate = c.max_rate;
  r->degenerate_total = c.degenerate_total;
  r->step = c.step;
  r->max_step = c.max_step;
  r->nonfinite_count = c.nonfinite;
  r->status = c.status;
}

ec_status enter(ec_batch* b)
{
  BT_TRY(b, hipSetDevice(b->device));
  // Order the batch's stream after the caller's work.
  BT_TRY(b, hipEventRecord(b->ev_user, b->user_stream));
  BT_TRY(b, hipStreamWaitEvent(b->stream, b->ev_user, 0));
  return EC_OK;
}

}  // namespace

extern "C" ec_status ec_batch_create(const ec_member_desc* members, int n_members, double gamma,
                                     int riemann, int limiter, int spatial_order, void* stream,
                                     ec_batch** out)
{
  if (!out) return EC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!members || n_members < 1 || !(gamma > 1.0) || riemann < 0 || riemann > 3 || limiter < 0 ||
      limiter > 2 || (spatial_order != 1 && spatial_order != 2)) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  for (int m = 0; m < n_members; ++m)
    if (!member_ok(members[m])) return EC_ERR_INVALID_ARGUMENT;
  ec_batch* b = new ec_batch();
  b->nm = n_members;
  b->desc.assign(members, members + n_members);
  b->gamma = gamma;
  b->riemann = riemann;
  b->limiter = limiter;
  b->order = spatial_order;
  b->user_stream = static_cast<hipStream_t>(stream);
  b->off.resize(n_members);
  long long tot = 0;
  for (int m = 0; m < n_members; ++m) {
    b->off[m] = tot;
    const long long n = 4LL * members[m].nx * members[m].ny;
    tot += (n + 31) / 32 * 32;  // 256-byte aligned member blocks
  }
  b->total = tot;
  int cap = 0;
  for (int sh = 0; sh < kNumBatchShapes; ++sh) {
    int n = 0;
    for (int m = 0; m < n_members; ++m) n += tiles_of(members[m], sh, nullptr);
    cap = std::max(cap, n);
  }
  b->part_cap = cap;
  auto bad = [&](ec_status st) {
    ec_batch_destroy(b);
    return st;
  };
  if (hipGetDevice(&b->device) != hipSuccess) return bad(EC_ERR_HIP);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&b->ev_user) != hipSuccess || hipEventCreate(&b->ev_a) != hipSuccess ||
      hipEventCreate(&b->ev_b) != hipSuccess) {
    return bad(EC_ERR_HIP);
  }
  b->gstream[0] = b->stream;
  for (int g = 0; g < kMaxGroups; ++g) {
    if (g > 0 && hipStreamCreateWithFlags(&b->gstream[g], hipStreamNonBlocking) != hipSuccess)
      return bad(EC_ERR_HIP);
    if (hipEventCreateWithFlags(&b->gev[g], hipEventDisableTiming) != hipSuccess) return bad(EC_ERR_HIP);
  }
  const size_t sb = sizeof(double) * (size_t)tot;
  if (hipMalloc(&b->ua, sb) != hipSuccess || hipMalloc(&b->ub, sb) != hipSuccess ||
      hipMalloc(&b->ctrl, sizeof(DevCtrl) * 2 * (size_t)n_members) != hipSuccess ||
      hipMalloc(&b->final_ctrl, sizeof(DevCtrl) * (size_t)n_members) != hipSuccess ||
      hipHostMalloc(&b->ctrl_host, sizeof(DevCtrl) * 2 * (size_t)n_members) != hipSuccess ||
      hipMalloc(&b->part, sizeof(Partial) * 2 * (size_t)cap) != hipSuccess ||
      hipMalloc(&b->error, sizeof(int)) != hipSuccess ||
      hipMalloc(&b->arrive, sizeof(unsigned) * (size_t)n_members) != hipSuccess ||
      hipMalloc(&b->members, sizeof(MemberDev) * (size_t)n_members) != hipSuccess) {
    return bad(EC_ERR_OUT_OF_MEMORY);
  }
  if (hipMemsetAsync(b->ua, 0, sb, b->stream) != hipSuccess ||
      hipMemsetAsync(b->ub, 0, sb, b->stream) != hipSuccess ||
      hipMemsetAsync(b->part, 0, sizeof(Partial) * 2 * (size_t)cap, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess) {
    return bad(EC_ERR_HIP);
  }
  *out = b;
  return EC_OK;
}

extern "C" ec_status ec_batch_destroy(ec_batch* b)
{
  if (!b) return EC_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (int g = 0; g < kMaxGroups; ++g) {
    if (b->gtiles[g]) (void)hipFree(b->gtiles[g]);
    if (b->gev[g]) (void)hipEventDestroy(b->gev[g]);
    if (g > 0 && b->gstream[g]) (void)hipStreamDestroy(b->gstream[g]);
  }
  double* bufs[] = {b->ua, b->ub, b->backup};
  for (double* p : bufs)
    if (p) (void)hipFree(p);
  if (b->ctrl) (void)hipFree(b->ctrl);
  if (b->final_ctrl) (void)hipFree(b->final_ctrl);
  if (b->ctrl_host) (void)hipHostFree(b->ctrl_host);
  if (b->part) (void)hipFree(b->part);
  if (b->error) (void)hipFree(b->error);
  if (b->arrive) (void)hipFree(b->arrive);
  if (b->members) (void)hipFree(b->members);
  if (b->tile_member) (void)hipFree(b->tile_member);
  if (b->tiles) (void)hipFree(b->tiles);
  if (b->ranges) (void)hipFree(b->ranges);
  for (auto& e : b->exact) ec::exact_free(e);
  if (b->ev_user) (void)hipEventDestroy(b->ev_user);
  if (b->ev_a) (void)hipEventDestroy(b->ev_a);
  if (b->ev_b) (void)hipEventDestroy(b->ev_b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return EC_OK;
}

extern "C" const char* ec_batch_last_error(const ec_batch* b)
{
  return b ? b->last_error.c_str() : "";
}

static ec_status member_copy(ec_batch* b, int m, double* dst, const double* src,
                             const ec_layout* user, bool to_batch)
{
  if (!b || m < 0 || m >= b->nm || !user) return EC_ERR_INVALID_ARGUMENT;
  const ec_member_desc& d = b->desc[m];
  if (user->row_stride < d.nx || user->comp_stride < user->row_stride * d.ny)
    return EC_ERR_INVALID_ARGUMENT;
  ec_status st = enter(b);
  if (st != EC_OK) return st;
  const size_t w = sizeof(double) * (size_t)d.nx;
  const size_t mine_pitch = w, user_pitch = sizeof(double) * (size_t)user->row_stride;
  for (int c = 0; c < 4; ++c) {
    const long long mine = b->off[m] + (long long)c * d.nx * d.ny;
    const long long theirs = (long long)c * user->comp_stride;
    if (to_batch) {
      BT_TRY(b, hipMemcpy2DAsync(b->ua + mine, mine_pitch, src + theirs, user_pitch, w, d.ny,
                                 hipMemcpyDefault, b->stream));
    } else {
      BT_TRY(b, hipMemcpy2DAsync(dst + theirs, user_pitch, b->ua + mine, mine_pitch, w, d.ny,
                                 hipMemcpyDefault, b->stream));
    }
  }
  BT_TRY(b, hipStreamSynchronize(b->stream));
  return EC_OK;
}

extern "C" ec_status ec_batch_set_exact_boundary(ec_batch* b, int member, const ec_problem_desc* pd,
                                                 const int* periodic)
{
  if (!b || !pd || !periodic || member < 0 || member >= b->nm) return EC_ERR_INVALID_ARGUMENT;
  ec_member_desc& d = b->desc[member];
  if (pd->nx != d.nx || pd->ny != d.ny || pd->gamma != b->gamma) {
    b->last_error = "ec_batch_set_exact_boundary: problem grid or gamma differs from the member's";
    return EC_ERR_INVALID_ARGUMENT;
  }
  ec_status st = enter(b);
  if (st != EC_OK) return st;
  // Work in flight may read the member's current descriptor.
  BT_TRY(b, hipStreamSynchronize(b->stream));
  int bc[4], plain[4];
  ec::ExactDevice e;
  std::string err;
  st = ec::exact_build(pd, periodic[0], periodic[1], bc, plain, &e, &err);
  if (st != EC_OK) {
    if (!err.empty()) b->last_error = "ec_batch_set_exact_boundary: " + err;
    return st;
  }
  if (b->exact.empty()) b->exact.resize(b->nm);
  ec::exact_free(b->exact[member]);
  b->exact[member] = e;
  for (int s = 0; s < 4; ++s) d.bc[s] = bc[s];
  b->cfg_shape = -1;  // the device tables hold the old sides
  return EC_OK;
}

extern "C" ec_status ec_batch_set_state(ec_batch* b, int member, const double* src,
                                        const ec_layout* layout)
{
  if (!src) return EC_ERR_INVALID_ARGUMENT;
  return member_copy(b, member, nullptr, src, layout, true);
}

extern "C" ec_status ec_batch_get_state(ec_batch* b, int member, double* dst,
                                        const ec_layout* layout)
{
  if (!dst) return EC_ERR_INVALID_ARGUMENT;
  return member_copy(b, member, dst, nullptr, layout, false);
}

extern "C" int ec_batch_tile_count(void) { return kNumBatchShapes; }

extern "C" ec_status ec_batch_tile_shape(int tile_id, int* tx, int* ty, int* threads)
{
  if (tile_id < 0 || tile_id >= kNumBatchShapes) return EC_ERR_INVALID_ARGUMENT;
  if (tx) *tx = kBatchShapes[tile_id].tx;
  if (ty) *ty = kBatchShapes[tile_id].ty;
  if (threads) *threads = kBatchShapes[tile_id].nt;
  return EC_OK;
}

extern "C" ec_status ec_batch_configure(ec_batch* b, int mode, int tile_id, int rounds_per_graph)
{
  if (!b || mode < EC_BATCH_AUTO || mode > EC_BATCH_PERSISTENT || tile_id < -1 ||
      tile_id >= kNumBatchShapes || rounds_per_graph < 0 || (rounds_per_graph & 1)) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  b->mode = mode;
  b->shape = tile_id;
  if (rounds_per_graph > 0 && rounds_per_graph != b->rounds_per_graph) {
    b->rounds_per_graph = rounds_per_graph;
    drop_graphs(b);
  }
  return EC_OK;
}

extern "C" ec_status ec_batch_set_groups(ec_batch* b, int groups)
{
  if (!b || groups < 1 || groups > kMaxGroups) return EC_ERR_INVALID_ARGUMENT;
  b->groups = groups;   // the tables and graphs follow on the next run
  return EC_OK;
}

extern "C" int ec_batch_groups(const ec_batch* b) { return b ? b->groups : 0; }

extern "C" ec_status ec_batch_autotune(ec_batch* b, ec_batch_timing* table, int table_cap,
                                       int* n_entries)
{
  if (!b || (table_cap > 0 && !table)) return EC_ERR_INVALID_ARGUMENT;
  ec_status st = enter(b);
  if (st != EC_OK) return st;
  const size_t sb = sizeof(double) * (size_t)b->total;
  if (!b->backup && hipMalloc(&b->backup, sb) != hipSuccess) return EC_ERR_OUT_OF_MEMORY;
  BT_TRY(b, hipMemcpyAsync(b->backup, b->ua, sb, hipMemcpyDeviceToDevice, b->stream));
  double best = INFINITY;
  int best_mode = EC_BATCH_GRAPH, best_shape = 0, n = 0;
  // Graph mode only: the persistent mode's per-member waits need agent-scope
  // release/acquire fences, which write back and invalidate the XCD's L2 at
  // every wait, and it measured about twice graph mode's time on every tile
  // shape (64 Riemann strips, HLLC-P: 88-182 ms against 47-60 ms).  It stays
  // selectable through ec_batch_configure.
  for (int mode = EC_BATCH_GRAPH; mode <= EC_BATCH_GRAPH; ++mode) {
    for (int sh = 0; sh < kNumBatchShapes; ++sh) {
      BT_TRY(b, hipMemcpyAsync(b->ua, b->backup, sb, hipMemcpyDeviceToDevice, b->stream));
      double ms = 0.0;
      st = run_once(b, mode, sh, &ms);
      const bool ok = st == EC_OK;
      if (!ok && st != EC_ERR_UNSUPPORTED) return st;
      if (n < table_cap) table[n] = ec_batch_timing{mode, sh, ok ? ms : -1.0};
      ++n;
      if (ok && ms < best) {
        best = ms;
        best_mode = mode;
        best_shape = sh;
      }
    }
  }
  // Rounds per graph replay for the best shape: each replay ends in one host
  // synchronisation, and a run ends up to R - 1 idle rounds late.
  int best_r = b->rounds_per_graph;
  for (int R : {32, 64, 128}) {
    if (R == b->rounds_per_graph) continue;
    const int keep = b->rounds_per_graph;
    b->rounds_per_graph = R;
    BT_TRY(b, hipMemcpyAsync(b->ua, b->backup, sb, hipMemcpyDeviceToDevice, b->stream));
    double ms = 0.0;
    st = run_once(b, best_mode, best_shape, &ms);
    if (st != EC_OK) return st;
    if (ms < best) {
      best = ms;
      best_r = R;
    }
    b->rounds_per_graph = keep;
  }
  if (best_r != b->rounds_per_graph) {
    b->rounds_per_graph = best_r;
    drop_graphs(b);
  }
  // Member groups (streams) for the best shape and rounds per replay.
  int best_g = b->groups;
  for (int G = 1; G <= kMaxGroups; ++G) {
    if (G == b->groups) continue;
    const int keep = b->groups;
    b->groups = G;
    BT_TRY(b, hipMemcpyAsync(b->ua, b->backup, sb, hipMemcpyDeviceToDevice, b->stream));
    double ms = 0.0;
    st = run_once(b, best_mode, best_shape, &ms);
    if (st != EC_OK) return st;
    if (ms < best) {
      best = ms;
      best_g = G;
    }
    b->groups = keep;
  }
  b->groups = best_g;
  BT_TRY(b, hipMemcpyAsync(b->ua, b->backup, sb, hipMemcpyDeviceToDevice, b->stream));
  BT_TRY(b, hipStreamSynchronize(b->stream));
  b->mode = best_mode;
  b->shape = best_shape;
  if (n_entries) *n_entries = n;
  return EC_OK;
}

extern "C" ec_status ec_batch_run(ec_batch* b, ec_run_state* results, ec_batch_timing* used)
{
  if (!b) return EC_ERR_INVALID_ARGUMENT;
  ec_status st = enter(b);
  if (st != EC_OK) return st;
  if (b->mode == EC_BATCH_AUTO || b->shape < 0) {
    if (b->mode == EC_BATCH_AUTO) {
      st = ec_batch_autotune(b, nullptr, 0, nullptr);
      if (st != EC_OK) return st;
    } else {
      b->shape = 0;
    }
  }
  double ms = 0.0;
  st = run_once(b, b->mode, b->shape, &ms);
  if (st != EC_OK) return st;
  if (used) *used = ec_batch_timing{b->last_mode, b->last_shape, ms};
  if (results) {
    std::vector<DevCtrl> fc(b->nm);
    BT_TRY(b, hipMemcpyAsync(fc.data(), b->final_ctrl, sizeof(DevCtrl) * (size_t)b->nm,
                             hipMemcpyDeviceToHost, b->stream));
    BT_TRY(b, hipStreamSynchronize(b->stream));
    for (int m = 0; m < b->nm; ++m) to_run_state(fc[m], results + m);
  }
  return EC_OK;
}
