// ec_context.hpp -- internal definition of ec_ctx and the device-side run
// control block shared by the explicit, IMEX and batch kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ec_amd.h"

namespace ec {

enum BcKind { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECT = 2, BC_FIXED = 3, BC_EXACT = 4 };

// Exact-Dirichlet ghost states of one test problem
// (euler.field_boundary=exact_dirichlet, problems/boundary_conditions.hpp:
// 133-205).  RING: time-independent states (Gresho, Riemann quadrant)
// evaluated once on the host into a table of the 2-cell ghost ring; TORO and
// BLOB: time-dependent states evaluated on read from the constants below.
enum ExactKind { EXACT_NONE = 0, EXACT_RING = 1, EXACT_TORO = 2, EXACT_BLOB = 3 };

struct ExactBc {
  int kind;
  int nx, ny;
  int periodic_x, periodic_y;
  double xlo, ylo, dx, dy, gamma;  // cell centres xlo + (i + 1/2) dx
  const double* ring;              // RING: 4 conserved values per ghost cell
  // TORO: left/right primitive states and the host-evaluated constants of
  // the exact solution (problems/exact_riemann.hpp:114-205).
  double ql[4], qr[4];
  double x0, p_star, u_star, a_l, a_r;
  int shock_l, shock_r;
  double speed_l, head_l, tail_l, rho_star_l;
  double speed_r, head_r, tail_r, rho_star_r;
  // BLOB: problems/problem_initialization.hpp:209-263.
  double blob_cx, blob_cy, blob_u, blob_v, blob_p, blob_r, rho_in, rho_out;
};

// Device-resident state of the explicit driver loop.  The last workgroup of
// every step kernel updates it; the next step kernel reads it at start.
struct DevCtrl {
  double time, stop_time, cfl;
  double dt_last, dt_min, dt_max, dt_sum;
  double rho_min, p_min, max_rate;
  long long degenerate_total;
  int step, max_step, nonfinite, status;
  int last_degenerate;  // degenerate count of the most recent step kernel
  unsigned int counter; // arrival counter of the last-workgroup reduction
};

// Per-workgroup partial of the fused epilogue reductions.
// Partials of a step above kTwoLevelMin tiles are reduced in two levels, by
// at most kPartialGroupsMax workgroups first; workspaces hold that many more
// slots after the tile partials.
constexpr int kTwoLevelMin = 1024;
constexpr int kPartialGroupsMax = 128;

struct Partial {
  double rate, rho_min, p_min;
  int bad, degenerate;
};

// Grid description in the form the kernels take by value.
struct GridParams {
  int nx, ny;
  double dx, dy, inv_dx, inv_dy, gamma;
  int riemann, limiter, order;
  int dx_pow2, dy_pow2;  // dx (dy) is a power of two: x / dx == x * inv_dx exactly
  int bc[4];
  double fixed[4][4];
  const ExactBc* exact;     // device copy; null unless a side is BC_EXACT
  const double* time_ptr;   // device-loop steps: the run control's time
  double time;              // otherwise: time of the exact ghost states
};

}  // namespace ec

struct ec_ctx {
  ec_grid_desc desc;
  ec::GridParams grid;
  hipStream_t stream = nullptr;
  int device = 0;
  int tile_id = 0;
  ec::DevCtrl* ctrl = nullptr;        // device
  ec::DevCtrl* ctrl_host = nullptr;   // pinned host mirror
  ec::Partial* partials = nullptr;    // device, capacity max_partials
  int max_partials = 0;
  double* scratch_state = nullptr;    // in-place step buffer (4*nx*ny)
  std::string last_error;
  double* dt_log = nullptr;           // device-loop dt record (checkpointed runs)
  // IMEX, adjoint and batch workspaces are attached lazily by their modules.
  void* imex = nullptr;
  void* adjoint = nullptr;
  void* direct = nullptr;
  // Whole-field metric reductions: workgroup partials (device) and their
  // pinned host copy, allocated on first use.
  double* metrics_part = nullptr;
  double* metrics_part_host = nullptr;
  // Optional step-kernel timing: event pairs recorded around each explicit
  // step kernel launch on the context's stream.
  bool timing = false;
  std::vector<hipEvent_t> tev;
  int tused = 0;
  // Exact-Dirichlet boundary: device descriptor and ghost-ring table.
  ec::ExactBc* exact_dev = nullptr;
  double* exact_ring = nullptr;
  int exact_kind = 0;
  // The rule of each side where exact states are not read (the IMEX step):
  // periodic on a periodic axis, else outflow.
  int exact_plain_bc[4] = {0, 0, 0, 0};
};

namespace ec {

inline ec_status hip_fail(ec_ctx* ctx, hipError_t e, const char* what)
{
  if (ctx) ctx->last_error = std::string(what) + ": " + hipGetErrorString(e);
  return EC_ERR_HIP;
}

#define EC_HIP_TRY(ctx, expr)                         \
  do {                                                \
    hipError_t _e = (expr);                           \
    if (_e != hipSuccess) return ec::hip_fail((ctx), _e, #expr); \
  } while (0)

GridParams make_grid_params(const ec_grid_desc& d);

}  // namespace ec
