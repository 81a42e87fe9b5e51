// ec_context.hpp -- internal definition of ec_ctx and the device-side run
// control block shared by the explicit, IMEX and batch kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ec_amd.h"

namespace ec {

enum BcKind { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_REFLECT = 2, BC_FIXED = 3 };

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
  // Optional step-kernel timing: event pairs recorded around each explicit
  // step kernel launch on the context's stream.
  bool timing = false;
  std::vector<hipEvent_t> tev;
  int tused = 0;
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
