// ec_exact_api.hpp -- building the device copy of an exact-Dirichlet
// boundary (euler.field_boundary=exact_dirichlet) for one grid: the side
// rules and the descriptor the ghost rule of ec_ghost.hpp reads.  Shared by
// the single-grid context and the batched ensembles.
#pragma once

#include <string>

#include "ec_context.hpp"

namespace ec {

struct ExactDevice {
  ExactBc* desc = nullptr;   // device descriptor (nullptr: every side periodic)
  double* ring = nullptr;    // ghost-ring table of the time-independent problems
  int kind = EXACT_NONE;
};

// Side rules for problem pd with geometry.is_periodic (px, py): bc[4] the
// rule of each side, plain_bc[4] the rule where exact states are not read
// (periodic on a periodic axis, else outflow), and the device copy (built on
// the current device; *out is left empty when both axes are periodic).
ec_status exact_build(const ec_problem_desc* pd, int px, int py, int bc[4], int plain_bc[4], ExactDevice* out,
                      std::string* err);
void exact_free(ExactDevice& d);

}  // namespace ec
