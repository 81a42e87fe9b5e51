// ec_problems.cpp -- host-side initial data of the euler_compare test
// problems, written into any ec_layout.  Host code on purpose: the reference
// keeps problem setup on the host and both sides then see bit-identical
// initial data.  Citations are relative to
// amrex/apps/euler_compare/src/euler_compare/.
#include <cmath>

#include "ec_problems.hpp"

namespace ec::prob {

void to_cons(const P4& q, double gamma, double* u)
{
  // core/state_primitives.hpp:113-117
  const double kinetic = 0.5 * q.r * (q.u * q.u + q.v * q.v);
  u[0] = q.r;
  u[1] = q.r * q.u;
  u[2] = q.r * q.v;
  u[3] = q.p / (gamma - 1.0) + kinetic;
}

// problems/exact_riemann.hpp:2-35
bool toro_state(int id, P4& l, P4& r)
{
  switch (id) {
    case 1: l = {1.0, 0.0, 0.0, 1.0}; r = {0.125, 0.0, 0.0, 0.1}; return true;
    case 2: l = {1.0, -2.0, 0.0, 0.4}; r = {1.0, 2.0, 0.0, 0.4}; return true;
    case 3: l = {1.0, 0.0, 0.0, 1000.0}; r = {1.0, 0.0, 0.0, 0.01}; return true;
    case 4: l = {1.0, 0.0, 0.0, 0.01}; r = {1.0, 0.0, 0.0, 100.0}; return true;
    case 5: l = {5.99924, 19.5975, 0.0, 460.894}; r = {5.99242, -6.19633, 0.0, 46.0950}; return true;
    case 93: l = {0.445, 1.698, 0.0, 3.528}; r = {0.5, 0.0, 0.0, 0.571}; return true;
    case 94: l = {1000.0, 1.0, 0.0, 1.0e5}; r = {0.01, 1.0, 0.0, 1.0e5}; return true;
    case 95: l = {1.0, 0.0, 0.0, 1.01}; r = {1.0, 0.0, 0.0, 1.0}; return true;
    case 96: l = {1.0, 0.0, 0.0, 1.0001}; r = {1.0, 0.0, 0.0, 1.0}; return true;
    case 97: l = {1.0, 2.0, 0.0, 0.1}; r = {1.0, -2.0, 0.0, 0.1}; return true;
    case 90: l = {1.0, 0.0, 0.0, 1.0}; r = {1.0, 0.0, 0.0, 1.0}; return true;
    case 91: l = {2.0, 0.0, 0.0, 1.0}; r = {1.0, 0.0, 0.0, 1.0}; return true;
    case 92: l = {2.0, 0.1, 0.0, 1.0}; r = {1.0, 0.1, 0.0, 1.0}; return true;
  }
  return false;
}

// problems/problem_state_functions.hpp:2-30 with the velocity scale
// mach * sqrt(gamma p0 / rho_outer) (core/state_primitives.hpp:207-215).
P4 gresho_state(double x, double y, const ec_problem_desc& d)
{
  const double dx = x - d.vortex_center[0];
  const double dy = y - d.vortex_center[1];
  const double r = std::sqrt(dx * dx + dy * dy);
  const double rho = d.density_outer;
  const double vs = d.mach * std::sqrt(d.gamma * d.pressure / d.density_outer);
  double tf = 0.0;
  double pf = -2.0 + 4.0 * std::log(2.0);
  if (r < 0.2) {
    tf = 5.0 * r;
    pf = 12.5 * r * r;
  } else if (r < 0.4) {
    tf = 2.0 - 5.0 * r;
    pf = 12.5 * r * r + 4.0 * (1.0 - 5.0 * r - std::log(0.2) + std::log(r));
  }
  const double vt = vs * tf;
  const double p = d.pressure + rho * vs * vs * pf;
  if (r < 1.0e-14) return P4{rho, 0.0, 0.0, p};
  return P4{rho, -vt * dy / r, vt * dx / r, p};
}

// problems/exact_riemann.hpp:37-112: Newton iteration on the pressure
// function f_L(p) + f_R(p) + u_R - u_L = 0 (Toro, ch. 4), from the
// two-rarefaction-style guess, halving on a non-positive or non-finite step.
static void pressure_function(double p, const P4& q, double gamma, double& value, double& derivative)
{
  const double a = std::sqrt(gamma * q.p / q.r);
  if (p > q.p) {
    const double acoef = 2.0 / ((gamma + 1.0) * q.r);
    const double bcoef = (gamma - 1.0) / (gamma + 1.0) * q.p;
    const double root = std::sqrt(acoef / (p + bcoef));
    value = (p - q.p) * root;
    derivative = root * (1.0 - 0.5 * (p - q.p) / (p + bcoef));
    return;
  }
  const double ratio = p / q.p;
  value = 2.0 * a / (gamma - 1.0) * (std::pow(ratio, (gamma - 1.0) / (2.0 * gamma)) - 1.0);
  derivative = 1.0 / (q.r * a) * std::pow(ratio, -(gamma + 1.0) / (2.0 * gamma));
}

void exact_riemann_star(const P4& L, const P4& R, double gamma, double& p_star, double& u_star)
{
  const double a_l = std::sqrt(gamma * L.p / L.r);
  const double a_r = std::sqrt(gamma * R.p / R.r);
  const double guess = 0.5 * (L.p + R.p) - 0.125 * (R.u - L.u) * (L.r + R.r) * (a_l + a_r);
  double p = 1.0e-12 < guess ? guess : 1.0e-12;
  for (int iter = 0; iter < 80; ++iter) {
    double fl, dl, fr, dr;
    pressure_function(p, L, gamma, fl, dl);
    pressure_function(p, R, gamma, fr, dr);
    double next = p - (fl + fr + R.u - L.u) / (dl + dr);
    if (!(next > 0.0) || !std::isfinite(next)) next = 0.5 * p;
    const bool done = std::fabs(next - p) <= 1.0e-12 * (1.0 < next ? next : 1.0);
    p = next;
    if (done) break;
  }
  double fl, dl, fr, dr;
  pressure_function(p, L, gamma, fl, dl);
  pressure_function(p, R, gamma, fr, dr);
  p_star = p;
  u_star = 0.5 * (L.u + R.u + fr - fl);
}

// problems/problem_state_functions.hpp:73-98 (Liska-Wendroff quadrant)
P4 quadrant_state(double x, double y, const ec_problem_desc& d)
{
  const bool right = x >= d.riemann_interface_x;
  const bool top = y >= d.riemann_interface_y;
  if (right && top) return P4{1.5, 0.0, 0.0, 1.5};
  if (!right && top) return P4{0.532258064516129, 1.206045378311055, 0.0, 0.3};
  if (!right && !top) {
    return P4{0.137992831541219, 1.206045378311055, 1.206045378311055, 0.029032258064516};
  }
  return P4{0.532258064516129, 0.0, 1.206045378311055, 0.3};
}

// exact_riemann_sample (problems/exact_riemann.hpp:114-205).
P4 exact_riemann_sample(double x, double t, const P4& L, const P4& R, double gamma, double x0)
{
  if (t <= 0.0) return x < x0 ? L : R;
  double ps, us;
  exact_riemann_star(L, R, gamma, ps, us);
  const double xi = (x - x0) / t;
  const double a_l = std::sqrt(gamma * L.p / L.r), a_r = std::sqrt(gamma * R.p / R.r);
  const double gm1 = gamma - 1.0, gp1 = gamma + 1.0;
  auto shock_rho = [&](const P4& q) {
    const double ratio = ps / q.p;
    return q.r * ((ratio + gm1 / gp1) / ((gm1 / gp1) * ratio + 1.0));
  };
  auto fan_rho = [&](const P4& q) { return q.r * std::pow(ps / q.p, 1.0 / gamma); };
  if (xi <= us) {
    if (ps > L.p) {
      const double s = L.u - a_l * std::sqrt(gp1 / (2.0 * gamma) * ps / L.p + gm1 / (2.0 * gamma));
      return xi <= s ? L : P4{shock_rho(L), us, L.v, ps};
    }
    const double head = L.u - a_l;
    const double tail = us - a_l * std::pow(ps / L.p, gm1 / (2.0 * gamma));
    if (xi <= head) return L;
    if (xi >= tail) return P4{fan_rho(L), us, L.v, ps};
    const double u = 2.0 / gp1 * (a_l + 0.5 * gm1 * L.u + xi);
    const double a = 2.0 / gp1 * (a_l + 0.5 * gm1 * (L.u - xi));
    return P4{L.r * std::pow(a / a_l, 2.0 / gm1), u, L.v, L.p * std::pow(a / a_l, 2.0 * gamma / gm1)};
  }
  if (ps > R.p) {
    const double s = R.u + a_r * std::sqrt(gp1 / (2.0 * gamma) * ps / R.p + gm1 / (2.0 * gamma));
    return xi >= s ? R : P4{shock_rho(R), us, R.v, ps};
  }
  const double head = R.u + a_r;
  const double tail = us + a_r * std::pow(ps / R.p, gm1 / (2.0 * gamma));
  if (xi >= head) return R;
  if (xi <= tail) return P4{fan_rho(R), us, R.v, ps};
  const double u = 2.0 / gp1 * (-a_r + 0.5 * gm1 * R.u + xi);
  const double a = 2.0 / gp1 * (a_r - 0.5 * gm1 * (R.u - xi));
  return P4{R.r * std::pow(a / a_r, 2.0 / gm1), u, R.v, R.p * std::pow(a / a_r, 2.0 * gamma / gm1)};
}

}  // namespace ec::prob

namespace {

using namespace ec::prob;

// problems/problem_state_functions.hpp:32-71
double periodic_nearest_delta(double x, double c, double lo, double hi)
{
  const double length = hi - lo;
  double delta = x - c;
  if (length > 0.0) delta -= length * std::floor(delta / length + 0.5);
  return delta;
}

P4 isentropic_vortex_state(double x, double y, const ec_problem_desc& d, double time = 0.0)
{
  const double pi = 3.141592653589793238462643383279502884;
  const double g = d.gamma;
  const double s = d.isentropic_vortex_strength;
  // the vortex advects with the mean flow (problem_state_functions.hpp:52-53)
  const double cx = d.vortex_center[0] + d.velocity[0] * time;
  const double cy = d.vortex_center[1] + d.velocity[1] * time;
  const double dx = periodic_nearest_delta(x, cx, d.prob_lo[0], d.prob_hi[0]);
  const double dy = periodic_nearest_delta(y, cy, d.prob_lo[1], d.prob_hi[1]);
  const double r2 = dx * dx + dy * dy;
  const double exp_half = std::exp(0.5 * (1.0 - r2));
  const double exp_full = std::exp(1.0 - r2);
  const double temp = 1.0 - (g - 1.0) * s * s / (8.0 * g * pi * pi) * exp_full;
  const double rho = d.density_outer * std::pow(temp, 1.0 / (g - 1.0));
  const double p = d.pressure * std::pow(temp, g / (g - 1.0));
  const double vf = s / (2.0 * pi) * exp_half;
  return P4{rho, d.velocity[0] - vf * dy, d.velocity[1] + vf * dx, p};
}


// problems/problem_initialization.hpp:2-102 (fixed geometry of the
// reference's shock-density bubble: shock x=0.2, bubble (0.5,0) r=0.2).
const P4 kBubblePostShock{2.8181818181818183, 1.6064386578049976, 0.0, 5.0};

double bubble_fraction(double xc, double yc, double dx, double dy, double shock_x, double bx,
                       double by, double br)
{
  double inside = 0.0;
  for (int sy = 0; sy < 4; ++sy) {
    for (int sx = 0; sx < 4; ++sx) {
      const double x = xc + (((double)sx + 0.5) / 4.0 - 0.5) * dx;
      const double y = yc + (((double)sy + 0.5) / 4.0 - 0.5) * dy;
      const double rx = x - bx, ry = y - by;
      if (x >= shock_x && rx * rx + ry * ry <= br * br) inside += 1.0;
    }
  }
  return inside / 16.0;
}

}  // namespace

extern "C" ec_status ec_post_shock_state(double gamma, double mach, double rho1, double p1,
                                         double* prim_out)
{
  if (!prim_out || !(gamma > 1.0) || !(mach >= 1.0) || !(rho1 > 0.0) || !(p1 > 0.0)) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  // Rankine-Hugoniot for a shock moving at mach * c1 into gas at rest.
  const double m2 = mach * mach;
  const double c1 = std::sqrt(gamma * p1 / rho1);
  const double rho2 = rho1 * (gamma + 1.0) * m2 / ((gamma - 1.0) * m2 + 2.0);
  const double p2 = p1 * (1.0 + 2.0 * gamma / (gamma + 1.0) * (m2 - 1.0));
  const double u2 = mach * c1 * (1.0 - rho1 / rho2);
  prim_out[0] = rho2;
  prim_out[1] = u2;
  prim_out[2] = 0.0;
  prim_out[3] = p2;
  return EC_OK;
}

extern "C" ec_status ec_init_state(const ec_problem_desc* d, double* U, const ec_layout* lay)
{
  if (!d || !U || !lay || d->nx < 1 || d->ny < 1 || lay->row_stride < d->nx ||
      lay->comp_stride < lay->row_stride * d->ny) {
    return EC_ERR_INVALID_ARGUMENT;
  }
  // AMReX Geometry: dx = (prob_hi - prob_lo) / n_cell.
  const double dx = (d->prob_hi[0] - d->prob_lo[0]) / d->nx;
  const double dy = (d->prob_hi[1] - d->prob_lo[1]) / d->ny;
  P4 tl{}, tr{};
  if (d->kind == EC_PROBLEM_TORO && !toro_state(d->toro_test, tl, tr)) return EC_ERR_INVALID_ARGUMENT;
  P4 post = kBubblePostShock;
  if (d->kind == EC_PROBLEM_SHOCK_BUBBLE_MACH) {
    double q[4];
    const ec_status st = ec_post_shock_state(d->gamma, d->shock_mach, d->density_outer, d->pressure, q);
    if (st != EC_OK) return st;
    post = P4{q[0], q[1], q[2], q[3]};
  }
  for (int j = 0; j < d->ny; ++j) {
    for (int i = 0; i < d->nx; ++i) {
      const double x = d->prob_lo[0] + ((double)i + 0.5) * dx;
      const double y = d->prob_lo[1] + ((double)j + 0.5) * dy;
      P4 q{};
      switch (d->kind) {
        case EC_PROBLEM_TORO:  // problem_initialization.hpp:265-284
          q = x < d->riemann_interface_x ? tl : tr;
          break;
        case EC_PROBLEM_GRESHO:
          q = gresho_state(x, y, *d);
          break;
        case EC_PROBLEM_ISENTROPIC_VORTEX:
          q = isentropic_vortex_state(x, y, *d);
          break;
        case EC_PROBLEM_RIEMANN_QUADRANT:
          q = quadrant_state(x, y, *d);
          break;
        case EC_PROBLEM_SHOCK_DENSITY_BUBBLE:  // problem_initialization.hpp:91-102
          if (x < 0.2) {
            q = kBubblePostShock;
          } else {
            const double f = bubble_fraction(x, y, dx, dy, 0.2, 0.5, 0.0, 0.2);
            q = P4{f * 0.1 + (1.0 - f) * 1.0, 0.0, 0.0, 1.0};
          }
          break;
        case EC_PROBLEM_SHOCK_BUBBLE_MACH:
          if (x < d->shock_x) {
            q = post;
          } else {
            const double f = bubble_fraction(x, y, dx, dy, d->shock_x, d->blob_center[0],
                                             d->blob_center[1], d->blob_radius);
            q = P4{f * d->density_inner + (1.0 - f) * d->density_outer, 0.0, 0.0, d->pressure};
          }
          break;
        case EC_PROBLEM_ADVECTION_BLOB: {  // problem_initialization.hpp:127-155
          // The reference places the blob's cell centres in y with the x
          // cell size (:146); on square cells this is the usual centre.
          const double yb = d->prob_lo[1] + ((double)j + 0.5) * dx;
          const double cx = d->blob_center[0], cy = d->blob_center[1];
          const double r2 = (x - cx) * (x - cx) + (yb - cy) * (yb - cy);
          const double rho = r2 <= d->blob_radius * d->blob_radius ? d->density_inner : d->density_outer;
          const double u = d->velocity[0], v = d->velocity[1];
          const long long k = (long long)j * lay->row_stride + i;
          U[k] = rho;
          U[lay->comp_stride + k] = rho * u;
          U[2 * lay->comp_stride + k] = rho * v;
          U[3 * lay->comp_stride + k] = d->pressure / (d->gamma - 1.0) + 0.5 * rho * (u * u + v * v);
          continue;
        }
        default:
          return EC_ERR_INVALID_ARGUMENT;
      }
      double u[4];
      to_cons(q, d->gamma, u);
      const long long k = (long long)j * lay->row_stride + i;
      for (int c = 0; c < 4; ++c) U[c * lay->comp_stride + k] = u[c];
    }
  }
  return EC_OK;
}

// problems/problem_initialization.hpp:157-184
static double periodic_signed_distance(double x, double c, double lo, double hi, bool periodic)
{
  double d = x - c;
  if (periodic) {
    const double length = hi - lo;
    if (length > 0.0) d -= std::round(d / length) * length;
  }
  return d;
}

static double wrap_periodic_coordinate(double x, double lo, double hi)
{
  const double length = hi - lo;
  if (length <= 0.0) return x;
  double w = lo + std::fmod(x - lo, length);
  if (w < lo) w += length;
  return w;
}

extern "C" ec_status ec_exact_state_at(const ec_problem_desc* d, const int* periodic, double x,
                                       double y, double t, double* prim)
{
  if (!d || !periodic || !prim) return EC_ERR_INVALID_ARGUMENT;
  P4 q{};
  switch (d->kind) {
    case EC_PROBLEM_TORO: {
      P4 l, r;
      if (!toro_state(d->toro_test, l, r)) return EC_ERR_INVALID_ARGUMENT;
      q = exact_riemann_sample(x, t, l, r, d->gamma, d->riemann_interface_x);
      break;
    }
    case EC_PROBLEM_GRESHO:
      q = gresho_state(x, y, *d);
      break;
    case EC_PROBLEM_ISENTROPIC_VORTEX:
      q = isentropic_vortex_state(x, y, *d, t);
      break;
    case EC_PROBLEM_RIEMANN_QUADRANT:
      q = quadrant_state(x, y, *d);
      break;
    case EC_PROBLEM_ADVECTION_BLOB: {  // problem_initialization.hpp:186-207
      double cx = d->blob_center[0] + d->velocity[0] * t;
      double cy = d->blob_center[1] + d->velocity[1] * t;
      if (periodic[0]) cx = wrap_periodic_coordinate(cx, d->prob_lo[0], d->prob_hi[0]);
      if (periodic[1]) cy = wrap_periodic_coordinate(cy, d->prob_lo[1], d->prob_hi[1]);
      const double ddx = periodic_signed_distance(x, cx, d->prob_lo[0], d->prob_hi[0], periodic[0] != 0);
      const double ddy = periodic_signed_distance(y, cy, d->prob_lo[1], d->prob_hi[1], periodic[1] != 0);
      const double r2 = ddx * ddx + ddy * ddy;
      q = P4{r2 <= d->blob_radius * d->blob_radius ? d->density_inner : d->density_outer,
             d->velocity[0], d->velocity[1], d->pressure};
      break;
    }
    default:
      return EC_ERR_INVALID_ARGUMENT;
  }
  prim[0] = q.r;
  prim[1] = q.u;
  prim[2] = q.v;
  prim[3] = q.p;
  return EC_OK;
}

extern "C" ec_status ec_problem_defaults(int kind, ec_problem_desc* d)
{
  if (!d) return EC_ERR_INVALID_ARGUMENT;
  // core/run_config.hpp:2-55 defaults.
  *d = ec_problem_desc{};
  d->kind = kind;
  d->nx = 64;
  d->ny = 64;
  d->prob_lo[0] = 0.0;
  d->prob_lo[1] = 0.0;
  d->prob_hi[0] = 1.0;
  d->prob_hi[1] = 1.0;
  d->gamma = 1.4;
  d->toro_test = 1;
  d->riemann_interface_x = 0.5;
  d->riemann_interface_y = 0.5;
  d->velocity[0] = 1.0;
  d->velocity[1] = 0.5;
  d->density_inner = 2.0;
  d->density_outer = 1.0;
  d->pressure = 1.0;
  d->mach = 0.1;
  d->blob_center[0] = 0.35;
  d->blob_center[1] = 0.35;
  d->blob_radius = 0.15;
  d->vortex_center[0] = 0.0;
  d->vortex_center[1] = 0.0;
  d->isentropic_vortex_strength = 5.0;
  d->shock_mach = 1.22;
  d->shock_x = 0.5;
  return EC_OK;
}
