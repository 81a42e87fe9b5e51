// ec_exact.hip -- exact-Dirichlet ghost states (euler.field_boundary=
// exact_dirichlet): set-up of the device descriptor the ghost rule reads.
//
// The reference fills the ghost cells of a non-periodic axis with the
// problem's exact state before every step (fill_problem_ghosts,
// amrex/apps/euler_compare/src/euler_compare/problems/boundary_conditions.hpp:
// 133-205).  Here the sides become BC_EXACT and the step evaluates those
// states on read: time-independent ones (Gresho, Riemann quadrant) come from
// a table of the 2-cell ghost ring evaluated once on the host with the
// reference's formulas; the Toro exact solution and the advected blob are
// evaluated in the kernel at the step's time, with every constant of the
// exact Riemann solution (the Newton-iterated star state, wave speeds and
// star densities) computed here.
#include <cmath>
#include <cstring>
#include <vector>

#include "ec_context.hpp"
#include "ec_exact_api.hpp"
#include "ec_ghost.hpp"
#include "ec_problems.hpp"

namespace ec {

// Toro constants of exact_riemann_sample (problems/exact_riemann.hpp:114-205),
// evaluated once in the reference's operation order.
void toro_constants(ExactBc& e, const prob::P4& L, const prob::P4& R, double gamma)
{
  prob::exact_riemann_star(L, R, gamma, e.p_star, e.u_star);
  e.a_l = std::sqrt(gamma * L.p / L.r);
  e.a_r = std::sqrt(gamma * R.p / R.r);
  const double gm1 = gamma - 1.0, gp1 = gamma + 1.0;
  auto shock_density = [&](const prob::P4& q) {
    const double ratio = e.p_star / q.p;
    return q.r * ((ratio + gm1 / gp1) / ((gm1 / gp1) * ratio + 1.0));
  };
  auto rarefaction_density = [&](const prob::P4& q) { return q.r * std::pow(e.p_star / q.p, 1.0 / gamma); };
  e.shock_l = e.p_star > L.p ? 1 : 0;
  if (e.shock_l) {
    e.speed_l = L.u - e.a_l * std::sqrt(gp1 / (2.0 * gamma) * e.p_star / L.p + gm1 / (2.0 * gamma));
    e.rho_star_l = shock_density(L);
  } else {
    e.head_l = L.u - e.a_l;
    const double a_star = e.a_l * std::pow(e.p_star / L.p, gm1 / (2.0 * gamma));
    e.tail_l = e.u_star - a_star;
    e.rho_star_l = rarefaction_density(L);
  }
  e.shock_r = e.p_star > R.p ? 1 : 0;
  if (e.shock_r) {
    e.speed_r = R.u + e.a_r * std::sqrt(gp1 / (2.0 * gamma) * e.p_star / R.p + gm1 / (2.0 * gamma));
    e.rho_star_r = shock_density(R);
  } else {
    e.head_r = R.u + e.a_r;
    const double a_star = e.a_r * std::pow(e.p_star / R.p, gm1 / (2.0 * gamma));
    e.tail_r = e.u_star + a_star;
    e.rho_star_r = rarefaction_density(R);
  }
}

void exact_free(ExactDevice& d)
{
  if (d.desc) (void)hipFree(d.desc);
  if (d.ring) (void)hipFree(d.ring);
  d = ExactDevice{};
}

void exact_release(ec_ctx* ctx)
{
  if (ctx->exact_dev) (void)hipFree(ctx->exact_dev);
  if (ctx->exact_ring) (void)hipFree(ctx->exact_ring);
  ctx->exact_dev = nullptr;
  ctx->exact_ring = nullptr;
  ctx->exact_kind = EXACT_NONE;
}

bool has_exact_side(const GridParams& g)
{
  for (int s = 0; s < 4; ++s)
    if (g.bc[s] == BC_EXACT) return true;
  return false;
}

ec_status exact_build(const ec_problem_desc* pd, int px, int py, int bc[4], int plain_bc[4], ExactDevice* out,
                      std::string* err)
{
  *out = ExactDevice{};
  const int kinds[] = {EC_PROBLEM_GRESHO, EC_PROBLEM_TORO, EC_PROBLEM_RIEMANN_QUADRANT,
                       EC_PROBLEM_ADVECTION_BLOB};
  bool known = false;
  for (int k : kinds) known = known || pd->kind == k;
  if (!known) {
    *err = "the problem has no exact-Dirichlet ghost state";
    return EC_ERR_INVALID_ARGUMENT;
  }
  prob::P4 tl{}, tr{};
  if (pd->kind == EC_PROBLEM_TORO && !prob::toro_state(pd->toro_test, tl, tr)) {
    *err = "unknown Toro test id";
    return EC_ERR_INVALID_ARGUMENT;
  }
  px = px ? 1 : 0;
  py = py ? 1 : 0;
  // The exact fill applies when an axis is not periodic
  // (use_*_exact_dirichlet_boundary, problems/boundary_conditions.hpp:165-191);
  // otherwise every side is periodic.
  for (int s = 0; s < 4; ++s) bc[s] = BC_PERIODIC;
  if (!px || !py) {
    // Toro's fill writes the x ghosts whatever the x periodicity
    // (problems/problem_initialization.hpp:329-330).
    const int x_exact = (pd->kind == EC_PROBLEM_TORO || !px) ? BC_EXACT : BC_PERIODIC;
    bc[0] = bc[1] = x_exact;
    bc[2] = bc[3] = py ? BC_PERIODIC : BC_EXACT;
  }
  for (int s = 0; s < 4; ++s) plain_bc[s] = (s < 2 ? px : py) ? BC_PERIODIC : BC_OUTFLOW;
  if (px && py) return EC_OK;

  ExactBc e;
  std::memset(&e, 0, sizeof(e));
  e.nx = pd->nx;
  e.ny = pd->ny;
  e.periodic_x = px;
  e.periodic_y = py;
  e.xlo = pd->prob_lo[0];
  e.ylo = pd->prob_lo[1];
  // AMReX Geometry: dx = (prob_hi - prob_lo) / n_cell.
  e.dx = (pd->prob_hi[0] - pd->prob_lo[0]) / pd->nx;
  e.dy = (pd->prob_hi[1] - pd->prob_lo[1]) / pd->ny;
  e.gamma = pd->gamma;
  if (pd->kind == EC_PROBLEM_TORO) {
    e.kind = EXACT_TORO;
    const double l[4] = {tl.r, tl.u, tl.v, tl.p}, r[4] = {tr.r, tr.u, tr.v, tr.p};
    for (int c = 0; c < 4; ++c) {
      e.ql[c] = l[c];
      e.qr[c] = r[c];
    }
    e.x0 = pd->riemann_interface_x;
    toro_constants(e, tl, tr, pd->gamma);
  } else if (pd->kind == EC_PROBLEM_ADVECTION_BLOB) {
    e.kind = EXACT_BLOB;
    e.blob_cx = pd->blob_center[0];
    e.blob_cy = pd->blob_center[1];
    e.blob_u = pd->velocity[0];
    e.blob_v = pd->velocity[1];
    e.blob_p = pd->pressure;
    e.blob_r = pd->blob_radius;
    e.rho_in = pd->density_inner;
    e.rho_out = pd->density_outer;
  } else {
    // Gresho and the quadrant: states at the ghost cell's own centre, no
    // periodic image (problems/boundary_conditions.hpp:152-160,
    // problems/problem_initialization.hpp:365-372).
    e.kind = EXACT_RING;
    const long long n = ring_cells(pd->nx, pd->ny);
    std::vector<double> ring(4 * n);
    for (int j = -2; j < pd->ny + 2; ++j) {
      for (int i = -2; i < pd->nx + 2; ++i) {
        if (i >= 0 && i < pd->nx && j >= 0 && j < pd->ny) continue;
        const double x = e.xlo + ((double)i + 0.5) * e.dx;
        const double y = e.ylo + ((double)j + 0.5) * e.dy;
        const prob::P4 q = pd->kind == EC_PROBLEM_GRESHO ? prob::gresho_state(x, y, *pd)
                                                         : prob::quadrant_state(x, y, *pd);
        prob::to_cons(q, pd->gamma, &ring[4 * ring_index(i, j, pd->nx, pd->ny)]);
      }
    }
    if (hipMalloc(&out->ring, sizeof(double) * ring.size()) != hipSuccess) {
      exact_free(*out);
      return EC_ERR_OUT_OF_MEMORY;
    }
    if (hipMemcpy(out->ring, ring.data(), sizeof(double) * ring.size(), hipMemcpyHostToDevice) != hipSuccess) {
      exact_free(*out);
      *err = "copying the exact ghost ring";
      return EC_ERR_HIP;
    }
    e.ring = out->ring;
  }
  if (hipMalloc(&out->desc, sizeof(ExactBc)) != hipSuccess) {
    exact_free(*out);
    return EC_ERR_OUT_OF_MEMORY;
  }
  if (hipMemcpy(out->desc, &e, sizeof(ExactBc), hipMemcpyHostToDevice) != hipSuccess) {
    exact_free(*out);
    *err = "copying the exact descriptor";
    return EC_ERR_HIP;
  }
  out->kind = e.kind;
  return EC_OK;
}

}  // namespace ec

using namespace ec;

extern "C" ec_status ec_set_exact_boundary(ec_ctx* ctx, const ec_problem_desc* pd, const int* periodic)
{
  if (!ctx || !pd || !periodic) return EC_ERR_INVALID_ARGUMENT;
  if (pd->nx != ctx->grid.nx || pd->ny != ctx->grid.ny || pd->gamma != ctx->grid.gamma) {
    ctx->last_error = "ec_set_exact_boundary: problem grid or gamma differs from the context's";
    return EC_ERR_INVALID_ARGUMENT;
  }
  EC_HIP_TRY(ctx, hipSetDevice(ctx->device));
  // Kernels in flight may read the current descriptor.
  EC_HIP_TRY(ctx, ctx->stream ? hipStreamSynchronize(ctx->stream) : hipDeviceSynchronize());
  int bc[4], plain[4];
  ExactDevice d;
  std::string err;
  const ec_status st = exact_build(pd, periodic[0], periodic[1], bc, plain, &d, &err);
  if (st != EC_OK) {
    if (!err.empty()) ctx->last_error = "ec_set_exact_boundary: " + err;
    return st;
  }
  exact_release(ctx);
  for (int s = 0; s < 4; ++s) {
    ctx->grid.bc[s] = bc[s];
    ctx->desc.bc[s] = bc[s];
    ctx->exact_plain_bc[s] = plain[s];
  }
  ctx->exact_dev = d.desc;
  ctx->exact_ring = d.ring;
  ctx->exact_kind = d.kind;
  ctx->grid.exact = d.desc;
  return EC_OK;
}

extern "C" ec_status ec_set_boundary_time(ec_ctx* ctx, double time)
{
  if (!ctx || !std::isfinite(time)) return EC_ERR_INVALID_ARGUMENT;
  ctx->grid.time = time;
  return EC_OK;
}
