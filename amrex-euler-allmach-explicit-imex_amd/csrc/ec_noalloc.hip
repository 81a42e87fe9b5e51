// ec_noalloc.hip -- the explicit step for callers that cannot allocate.
//
// ec_explicit_step_noalloc runs the same step kernel and finalize kernel as
// ec_explicit_step, with the same results, but needs no context: its
// workspace (the run control block the finalize kernel writes and one
// epilogue partial per tile) is the file-scope device array `scratch` below,
// which every GPU the library runs on holds until the process ends.  The
// host-side name of that array is not a device address; its address on each
// device comes from hipGetSymbolAddress, taken once per device with that
// device current, and only that address is handed to kernels and copies.
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU race.  Every call overwrites all of the workspace it reads (each tile
// writes its partial, the finalize kernel writes the run control fields the
// diagnostics come from), so nothing needs clearing between calls.
#include <mutex>
#include <string>

#include "ec_context.hpp"

namespace ec {

ec_status launch_step_raw(const GridParams& grid, int tile_id, hipStream_t stream, const double* uin,
                          double* uout, long long rs, long long cs, double dt, int device_dt,
                          DevCtrl* ctrl, Partial* partials, int max_partials, double* dt_log,
                          hipEvent_t ev0, hipEvent_t ev1);
int tiles_needed(int nx, int ny, int tile_id);
bool grid_desc_valid(const ec_grid_desc* d);

struct NoallocWorkspace {
  DevCtrl ctrl;
  Partial partials[EC_NOALLOC_MAX_TILES];
};

}  // namespace ec

static __device__ __attribute__((aligned(256))) ec::NoallocWorkspace scratch;

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
ec::NoallocWorkspace* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (taken once).
ec_status workspace_on(int device, ec::NoallocWorkspace** out)
{
  if (device < 0 || device >= kMaxDevices) return EC_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ptr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return EC_ERR_HIP;
    g_ptr[device] = static_cast<ec::NoallocWorkspace*>(p);
  }
  *out = g_ptr[device];
  return EC_OK;
}

}  // namespace

using namespace ec;

extern "C" ec_status ec_explicit_step_noalloc(const ec_grid_desc* desc, const double* U_in, double* U_out,
                                              const ec_layout* layout, double dt, ec_step_diag* diag,
                                              void* stream)
{
  if (!grid_desc_valid(desc) || !U_in || !U_out || U_in == U_out || !layout ||
      layout->row_stride < desc->nx || layout->comp_stride < layout->row_stride * desc->ny) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  if (tiles_needed(desc->nx, desc->ny, 0) > EC_NOALLOC_MAX_TILES) return EC_ERR_TOO_LARGE;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return EC_ERR_HIP;
  NoallocWorkspace* w = nullptr;
  const ec_status ws = workspace_on(device, &w);
  if (ws != EC_OK) return ws;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const GridParams g = make_grid_params(*desc);
  const ec_status st = launch_step_raw(g, 0, s, U_in, U_out, layout->row_stride, layout->comp_stride,
                                       dt, 0, &w->ctrl, w->partials, EC_NOALLOC_MAX_TILES, nullptr,
                                       nullptr, nullptr);
  if (st != EC_OK || !diag) return st;
  DevCtrl h;
  if (hipMemcpyAsync(&h, &w->ctrl, sizeof(DevCtrl), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    return EC_ERR_HIP;
  }
  diag->rho_min = h.rho_min;
  diag->pressure_min = h.p_min;
  diag->nonfinite_count = h.nonfinite;
  diag->degenerate_count = h.last_degenerate;
  diag->max_rate = h.max_rate;
  return EC_OK;
}
