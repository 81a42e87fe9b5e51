// ec_problems.hpp -- host-side analytic states of the test problems shared by
// the initial data and the exact-Dirichlet boundary set-up.  Citations are
// relative to amrex/apps/euler_compare/src/euler_compare/.
#pragma once

#include "../../include/ec_amd.h"

namespace ec::prob {

struct P4 { double r, u, v, p; };

// core/state_primitives.hpp:113-117, written into u[0..3].
void to_cons(const P4& q, double gamma, double* u);
// Left/right states of a Toro test id (problems/exact_riemann.hpp:2-35).
bool toro_state(int id, P4& l, P4& r);
// problems/problem_state_functions.hpp:2-30.
P4 gresho_state(double x, double y, const ec_problem_desc& d);
// problems/problem_state_functions.hpp:73-98.
P4 quadrant_state(double x, double y, const ec_problem_desc& d);
// Star pressure and velocity of the exact Riemann solution
// (problems/exact_riemann.hpp:37-112).
void exact_riemann_star(const P4& L, const P4& R, double gamma, double& p_star, double& u_star);
// The sampled exact Riemann solution (problems/exact_riemann.hpp:114-205).
P4 exact_riemann_sample(double x, double t, const P4& L, const P4& R, double gamma, double x0);

}  // namespace ec::prob
