// ec_capi.hip -- context lifetime and generic C-ABI entry points.
#include <cmath>
#include <cstring>

#include "ec_context.hpp"

namespace ec {

int num_tile_shapes();
int step_partials_needed(const ec_ctx* ctx, int tile_id);
int default_tile_id(int nx, int ny, int device);
int reduce_blocks(const ec_ctx* ctx);
void imex_release(ec_ctx* ctx);
void adjoint_release(ec_ctx* ctx);
void exact_release(ec_ctx* ctx);
void direct_release(ec_ctx* ctx);

GridParams make_grid_params(const ec_grid_desc& d)
{
  GridParams g{};
  g.nx = d.nx;
  g.ny = d.ny;
  g.dx = d.dx;
  g.dy = d.dy;
  // numerics/explicit_update.hpp:69-70
  g.inv_dx = 1.0 / d.dx;
  g.inv_dy = 1.0 / d.dy;
  int e;
  g.dx_pow2 = std::frexp(d.dx, &e) == 0.5 ? 1 : 0;
  g.dy_pow2 = std::frexp(d.dy, &e) == 0.5 ? 1 : 0;
  g.gamma = d.gamma;
  g.riemann = d.riemann;
  g.limiter = d.limiter;
  g.order = d.spatial_order;
  for (int s = 0; s < 4; ++s) {
    g.bc[s] = d.bc[s];
    for (int c = 0; c < 4; ++c) g.fixed[s][c] = d.fixed_state[s][c];
  }
  return g;
}

static bool desc_ok(const ec_grid_desc* d)
{
  if (!d || d->nx < 1 || d->ny < 1) return false;
  if (!(d->dx > 0.0) || !(d->dy > 0.0) || !std::isfinite(d->dx) || !std::isfinite(d->dy)) return false;
  if (!(d->gamma > 1.0)) return false;
  if (d->riemann < 0 || d->riemann > 3 || d->limiter < 0 || d->limiter > 2) return false;
  if (d->spatial_order != 1 && d->spatial_order != 2) return false;
  for (int s = 0; s < 4; ++s) {
    if (d->bc[s] < 0 || d->bc[s] > 3) return false;
  }
  // A periodic axis is periodic on both sides.
  if ((d->bc[0] == BC_PERIODIC) != (d->bc[1] == BC_PERIODIC)) return false;
  if ((d->bc[2] == BC_PERIODIC) != (d->bc[3] == BC_PERIODIC)) return false;
  return true;
}

bool grid_desc_valid(const ec_grid_desc* d) { return desc_ok(d); }

}  // namespace ec

using namespace ec;

extern "C" const char* ec_status_string(ec_status s)
{
  switch (s) {
    case EC_OK: return "ok";
    case EC_ERR_INVALID_ARGUMENT: return "invalid argument";
    case EC_ERR_HIP: return "HIP runtime error";
    case EC_ERR_OUT_OF_MEMORY: return "out of device memory";
    case EC_ERR_UNSUPPORTED: return "unsupported";
    case EC_ERR_TOO_LARGE: return "grid exceeds the fixed workspace of the allocation-free entry point";
  }
  return "unknown status";
}

extern "C" const char* ec_last_error(const ec_ctx* ctx)
{
  return ctx ? ctx->last_error.c_str() : "";
}

extern "C" ec_status ec_create(const ec_grid_desc* desc, void* stream, ec_ctx** out)
{
  if (!out) return EC_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!desc_ok(desc)) return EC_ERR_INVALID_ARGUMENT;
  ec_ctx* ctx = new ec_ctx();
  ctx->desc = *desc;
  ctx->grid = make_grid_params(*desc);
  ctx->stream = static_cast<hipStream_t>(stream);
  if (hipGetDevice(&ctx->device) != hipSuccess) {
    delete ctx;
    return EC_ERR_HIP;
  }
  int need = reduce_blocks(ctx);
  for (int t = 0; t < num_tile_shapes(); ++t) {
    const int n = step_partials_needed(ctx, t);
    if (n > need) need = n;
  }
  // room for the second-level partials of a large grid's reduction
  need += kPartialGroupsMax;
  ctx->max_partials = need;
  ctx->tile_id = default_tile_id(desc->nx, desc->ny, ctx->device);
  const size_t state_bytes = sizeof(double) * 4 * (size_t)desc->nx * desc->ny;
  if (hipMalloc(&ctx->ctrl, sizeof(DevCtrl)) != hipSuccess ||
      hipHostMalloc(&ctx->ctrl_host, sizeof(DevCtrl)) != hipSuccess ||
      hipMalloc(&ctx->partials, sizeof(Partial) * (size_t)need) != hipSuccess ||
      hipMalloc(&ctx->scratch_state, state_bytes) != hipSuccess) {
    ec_destroy(ctx);
    return EC_ERR_OUT_OF_MEMORY;
  }
  std::memset(ctx->ctrl_host, 0, sizeof(DevCtrl));
  ctx->ctrl_host->dt_min = INFINITY;
  if (hipMemcpy(ctx->ctrl, ctx->ctrl_host, sizeof(DevCtrl), hipMemcpyHostToDevice) != hipSuccess) {
    ec_destroy(ctx);
    return EC_ERR_HIP;
  }
  *out = ctx;
  return EC_OK;
}

extern "C" ec_status ec_destroy(ec_ctx* ctx)
{
  if (!ctx) return EC_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  imex_release(ctx);
  adjoint_release(ctx);
  exact_release(ctx);
  direct_release(ctx);
  for (hipEvent_t e : ctx->tev) (void)hipEventDestroy(e);
  if (ctx->ctrl) (void)hipFree(ctx->ctrl);
  if (ctx->ctrl_host) (void)hipHostFree(ctx->ctrl_host);
  if (ctx->partials) (void)hipFree(ctx->partials);
  if (ctx->scratch_state) (void)hipFree(ctx->scratch_state);
  if (ctx->metrics_part) (void)hipFree(ctx->metrics_part);
  if (ctx->metrics_part_host) (void)hipHostFree(ctx->metrics_part_host);
  delete ctx;
  return EC_OK;
}

extern "C" ec_status ec_alloc(ec_ctx* ctx, size_t bytes, void** dev_ptr)
{
  if (!ctx || !dev_ptr) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (hipMalloc(dev_ptr, bytes) != hipSuccess) return EC_ERR_OUT_OF_MEMORY;
  return EC_OK;
}

extern "C" ec_status ec_free(ec_ctx* ctx, void* dev_ptr)
{
  if (!ctx) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  EC_HIP_TRY(ctx, hipFree(dev_ptr));
  return EC_OK;
}

extern "C" ec_status ec_synchronize(ec_ctx* ctx)
{
  if (!ctx) return EC_ERR_INVALID_ARGUMENT;
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  EC_HIP_TRY(ctx, ctx->stream ? hipStreamSynchronize(ctx->stream) : hipDeviceSynchronize());
  return EC_OK;
}
