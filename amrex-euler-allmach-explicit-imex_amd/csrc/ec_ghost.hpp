// ec_ghost.hpp -- the explicit ghost rule, applied on read: the state of any
// cell index, inside or up to two cells outside the valid box, from the valid
// cells and the per-side boundary kind (or, on an exact-Dirichlet side, from
// the problem's exact solution).
#pragma once

#include "ec_context.hpp"
#include "ec_state.hpp"

namespace ec {

// Ghost rule of one axis; returns the valid index to read and whether the
// normal momentum flips (boundary_conditions.hpp:2-25, :27-73).
EC_HD int map_axis(int i, int n, int lo, int hi, bool& refl)
{
  refl = false;
  if (i < 0) {
    if (lo == BC_PERIODIC) {
      const int m = i % n;
      return m < 0 ? m + n : m;
    }
    if (lo == BC_REFLECT) {
      refl = true;
      i = -i - 1;
    }
  } else if (i >= n) {
    if (hi == BC_PERIODIC) return i % n;
    if (hi == BC_REFLECT) {
      refl = true;
      i = 2 * n - 1 - i;
    }
  } else {
    return i;
  }
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

// Position of ghost cell (i, j) in the ghost-ring table of an nx x ny grid:
// the two rows below (width nx + 4), the two rows above, then the four ghost
// cells of every valid row.
EC_HD long long ring_index(int i, int j, int nx, int ny)
{
  const long long w = nx + 4;
  if (j < 0) return (j + 2) * w + (i + 2);
  if (j >= ny) return (2 + j - ny) * w + (i + 2);
  return 4 * w + 4LL * j + (i < 0 ? i + 2 : 2 + i - nx);
}

EC_HD long long ring_cells(int nx, int ny) { return 4LL * (nx + 4) + 4LL * ny; }

// Whether the ghost value of (i, j) is an exact-Dirichlet state.
EC_HD bool exact_cell(const GridParams& g, int i, int j)
{
  return (i < 0 && g.bc[0] == BC_EXACT) || (i >= g.nx && g.bc[1] == BC_EXACT) ||
         (j < 0 && g.bc[2] == BC_EXACT) || (j >= g.ny && g.bc[3] == BC_EXACT);
}

// exact_riemann_sample (problems/exact_riemann.hpp:114-205) with the star
// state, wave speeds and star densities taken from the host; only the
// rarefaction fan is evaluated here.
EC_HD Prim toro_sample(const ExactBc& e, double x, double time)
{
  const Prim L{e.ql[0], e.ql[1], e.ql[2], e.ql[3]};
  const Prim R{e.qr[0], e.qr[1], e.qr[2], e.qr[3]};
  if (time <= 0.0) return x < e.x0 ? L : R;
  const double xi = (x - e.x0) / time;
  const double gm1 = e.gamma - 1.0;
  const double gp1 = e.gamma + 1.0;
  if (xi <= e.u_star) {
    if (e.shock_l) {
      if (xi <= e.speed_l) return L;
      return Prim{e.rho_star_l, e.u_star, L.v, e.p_star};
    }
    if (xi <= e.head_l) return L;
    if (xi >= e.tail_l) return Prim{e.rho_star_l, e.u_star, L.v, e.p_star};
    const double u = 2.0 / gp1 * (e.a_l + 0.5 * gm1 * L.u + xi);
    const double a = 2.0 / gp1 * (e.a_l + 0.5 * gm1 * (L.u - xi));
    const double rho = L.r * pow(a / e.a_l, 2.0 / gm1);
    const double p = L.p * pow(a / e.a_l, 2.0 * e.gamma / gm1);
    return Prim{rho, u, L.v, p};
  }
  if (e.shock_r) {
    if (xi >= e.speed_r) return R;
    return Prim{e.rho_star_r, e.u_star, R.v, e.p_star};
  }
  if (xi >= e.head_r) return R;
  if (xi <= e.tail_r) return Prim{e.rho_star_r, e.u_star, R.v, e.p_star};
  const double u = 2.0 / gp1 * (-e.a_r + 0.5 * gm1 * R.u + xi);
  const double a = 2.0 / gp1 * (e.a_r - 0.5 * gm1 * (R.u - xi));
  const double rho = R.r * pow(a / e.a_r, 2.0 / gm1);
  const double p = R.p * pow(a / e.a_r, 2.0 * e.gamma / gm1);
  return Prim{rho, u, R.v, p};
}

// The exact ghost state of (i, j) at the given time: the ring table
// (Gresho, problems/boundary_conditions.hpp:133-163; quadrant,
// problems/problem_initialization.hpp:344-375), the Toro exact solution at
// the ghost cell's x (:308-342) or the advected blob, evaluated at the
// periodic image of a periodic axis (:209-263).
EC_HD Cons exact_state(const ExactBc& e, double time, int i, int j)
{
  if (e.kind == EXACT_RING) {
    // Loads of a partial tile's unused lanes reach past the two ghost
    // layers; they read the nearest ring cell.
    const int ic = i < -2 ? -2 : (i > e.nx + 1 ? e.nx + 1 : i);
    const int jc = j < -2 ? -2 : (j > e.ny + 1 ? e.ny + 1 : j);
    const double* r = e.ring + 4 * ring_index(ic, jc, e.nx, e.ny);
    return Cons{r[0], r[1], r[2], r[3]};
  }
  if (e.kind == EXACT_TORO) {
    const double x = e.xlo + ((double)i + 0.5) * e.dx;
    return to_cons(toro_sample(e, x, time), e.gamma);
  }
  int ii = i, jj = j;
  if (e.periodic_x) {
    const int o = i % e.nx;
    ii = o < 0 ? o + e.nx : o;
  }
  if (e.periodic_y) {
    const int o = j % e.ny;
    jj = o < 0 ? o + e.ny : o;
  }
  const double x = e.xlo + ((double)ii + 0.5) * e.dx;
  const double y = e.ylo + ((double)jj + 0.5) * e.dy;
  const double cx = e.blob_cx + e.blob_u * time;
  const double cy = e.blob_cy + e.blob_v * time;
  const double r2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
  const double rho = r2 <= e.blob_r * e.blob_r ? e.rho_in : e.rho_out;
  return to_cons(Prim{rho, e.blob_u, e.blob_v, e.blob_p}, e.gamma);
}

EC_HD Cons fixed_state(const GridParams& g, int side)
{
  return Cons{g.fixed[side][0], g.fixed[side][1], g.fixed[side][2], g.fixed[side][3]};
}

// load_cell in two phases, so that a tile can issue all of its loads before
// it converts any cell: cell_src says which valid cell the ghost rule reads
// (with the momentum flips) or that the value is not read at all (an exact or
// fixed side); cell_value then forms the state from what was loaded.
struct CellSrc {
  long long k;   // offset of the valid cell read (kind 0)
  int kind;      // 0 read, 1 fixed side `side`, 2 exact
  int side;
  bool rx, ry;   // normal momentum flips of a reflecting side
};

EC_HD CellSrc cell_src(const GridParams& g, long long rs, int i, int j)
{
  CellSrc c{0, 0, 0, false, false};
  if (g.exact && exact_cell(g, i, j)) {
    c.kind = 2;
    return c;
  }
  const int side = (i < 0 && g.bc[0] == BC_FIXED)       ? 0
                   : (i >= g.nx && g.bc[1] == BC_FIXED) ? 1
                   : (j < 0 && g.bc[2] == BC_FIXED)     ? 2
                   : (j >= g.ny && g.bc[3] == BC_FIXED) ? 3
                                                        : -1;
  if (side >= 0) {
    c.kind = 1;
    c.side = side;
    return c;
  }
  const int ii = map_axis(i, g.nx, g.bc[0], g.bc[1], c.rx);
  const int jj = map_axis(j, g.ny, g.bc[2], g.bc[3], c.ry);
  c.k = (long long)jj * rs + ii;
  return c;
}

// The loaded values of a kind-0 source (unused otherwise).
EC_HD Cons cell_read(const double* __restrict__ u, long long cs, const CellSrc& c)
{
  if (c.kind != 0) return Cons{0.0, 0.0, 0.0, 0.0};
  return Cons{u[c.k], u[cs + c.k], u[2 * cs + c.k], u[3 * cs + c.k]};
}

EC_HD Cons cell_value(const CellSrc& c, Cons v, const GridParams& g, int i, int j,
                      const double* time_src)
{
  if (EC_UNLIKELY(c.kind == 2)) {
    const double t = time_src ? *time_src : (g.time_ptr ? *g.time_ptr : g.time);
    return exact_state(*g.exact, t, i, j);
  }
  if (EC_UNLIKELY(c.kind == 1)) return fixed_state(g, c.side);
  if (c.rx) v.mx = -v.mx;
  if (c.ry) v.my = -v.my;
  return v;
}

// time_src: where exact ghost states read the step's time, if not from g
// (the ensembles keep each member's step time in LDS).
EC_HD Cons load_cell(const double* __restrict__ u, long long rs, long long cs,
                                          const GridParams& g, int i, int j,
                                          const double* time_src = nullptr)
{
  if (g.exact && exact_cell(g, i, j)) {
    const double t = time_src ? *time_src : (g.time_ptr ? *g.time_ptr : g.time);
    return exact_state(*g.exact, t, i, j);
  }
  if (i < 0 && g.bc[0] == BC_FIXED) return fixed_state(g, 0);
  if (i >= g.nx && g.bc[1] == BC_FIXED) return fixed_state(g, 1);
  if (j < 0 && g.bc[2] == BC_FIXED) return fixed_state(g, 2);
  if (j >= g.ny && g.bc[3] == BC_FIXED) return fixed_state(g, 3);
  bool rx, ry;
  const int ii = map_axis(i, g.nx, g.bc[0], g.bc[1], rx);
  const int jj = map_axis(j, g.ny, g.bc[2], g.bc[3], ry);
  const long long k = (long long)jj * rs + ii;
  Cons c{u[k], u[cs + k], u[2 * cs + k], u[3 * cs + k]};
  if (rx) c.mx = -c.mx;
  if (ry) c.my = -c.my;
  return c;
}

}  // namespace ec
