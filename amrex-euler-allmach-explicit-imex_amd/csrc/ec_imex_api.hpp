// ec_imex_api.hpp -- the IMEX module's pressure solver as used by the
// direct BDLTV20 step: the workspace buffers and one solve of
//   diag p - cx [x part] - cy [y part] = b
// (the matrix-free row of ec_imex.hip over face enthalpies hx, hy), FP64
// GMRES or the mixed FGMRES form, to the tolerance of ec_imex_opts.
#pragma once

#include "ec_context.hpp"

namespace ec {

struct ImexBuffers {
  int nx, ny;
  double* star;   // 4n
  double* p;      // n: initial guess of each solve, then the accepted pressure
  double* x;      // n: the solution of the last solve
  double* mom;    // 2n
  double* hx;     // (nx+1)*ny, x-face f of row j between cells f-1 and f
  double* hy;     // nx*(ny+1)
  float* hx32;    // FP32 copies for the mixed form
  float* hy32;
  double* b;      // n
};

struct PressureSolve {
  int iterations, refinements, bad, breakdown;
  double rinf, rel, rhs_norm;
};

ec_status imex_buffers(ec_ctx* ctx, int restart, ImexBuffers* out);

// Solves with the right-hand side in b from the initial guess in p; the
// solution is left in x.  bc: per side BC_PERIODIC (wrapped columns) or
// BC_FIXED (columns outside are Dirichlet values already in b).
ec_status imex_solve_pressure(ec_ctx* ctx, const int bc[4], double cx, double cy, double diag,
                              const ec_imex_opts& o, PressureSolve* out);

}  // namespace ec
