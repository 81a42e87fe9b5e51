// ec_ghost.hpp -- the explicit ghost rule, applied on read: the state of any
// cell index, inside or up to two cells outside the valid box, from the valid
// cells and the per-side boundary kind.
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

EC_HD Cons fixed_state(const GridParams& g, int side)
{
  return Cons{g.fixed[side][0], g.fixed[side][1], g.fixed[side][2], g.fixed[side][3]};
}

EC_HD Cons load_cell(const double* __restrict__ u, long long rs, long long cs,
                                          const GridParams& g, int i, int j)
{
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
