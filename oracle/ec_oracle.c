/*
 * ORACLE -- test infrastructure only (see ec_oracle.h).
 *
 * Serial CPU restatement of the explicit branch of the euler_compare update
 * path.  Each function follows the reference's formula and floating-point
 * operation order so that an IEEE-754 build without FMA contraction
 * (-ffp-contract=off) reproduces the reference arithmetic.  Citations use
 * paths relative to amrex/apps/euler_compare/src/euler_compare/.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   to_primitive/to_conserved   round trip + closed form
 *   limited_slope               minmod/MC/van Leer closed forms
 *   hllc                        consistency F(q,q)=F(q); supersonic upwinding;
 *                               stationary contact is exact; Toro's textbook
 *                               HLLC star-state form (independent formula);
 *                               x/y rotation symmetry
 *   xie_hllcp                   consistency; reduces to HLLC for f=1 and a
 *                               locally supersonic state (theta=1); rotation
 *                               symmetry; f-affine structure; Roe's property
 *                               of the average recovered from beta_p (pins
 *                               u^, v^, H^, c^2, the (1+|M^|) factor and the
 *                               sign); the low-Mach pressure correction
 *                               against Toro's p*; the paper's arrangement
 *                               (HLLC + correction + dissipation)
 *   5-face sensor               f recovered from face fluxes = cube of the
 *                               minimum ratio over the face stencil, O1/O2
 *   low-Mach HLLC blend         z = 1 gives HLLC; z = 1/2 blend
 *   (tests/test_oracle_hllcp.py; tests/test_oracle_mutants.py plants a slip
 *   in each of these terms and checks that a pin rejects it)
 *   exact_riemann               Toro's Table 4.2 star values (tests 1-5)
 *   fv_step / run_explicit      Sod 400x4 HLLC O2 CFL 0.8 t=0.2 takes exactly
 *                               the 347 steps the reference run reports;
 *                               L1 error vs exact solution and its
 *                               convergence; periodic conservation; Gresho
 *                               steadiness and 90-degree rotation symmetry.
 */
#include "ec_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* std::min / std::max semantics, including the NaN behaviour
 * (AMReX_Algorithm.H:24-47 forwards amrex::min/max to std::min/max). */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double clamp01(double v) { return smin(smax(v, 0.0), 1.0); }

/* ------------------------------------------------------------------------ */
/* state_primitives.hpp:103-117 */
void ecor_to_primitive(const double* u, double gamma, double* q)
{
  const double inv_rho = 1.0 / u[0];
  const double ux = u[1] * inv_rho;
  const double uy = u[2] * inv_rho;
  const double kinetic = 0.5 * u[0] * (ux * ux + uy * uy);
  const double p = (gamma - 1.0) * (u[3] - kinetic);
  q[0] = u[0];
  q[1] = ux;
  q[2] = uy;
  q[3] = p;
}

void ecor_to_conserved(const double* q, double gamma, double* u)
{
  const double kinetic = 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]);
  u[0] = q[0];
  u[1] = q[0] * q[1];
  u[2] = q[0] * q[2];
  u[3] = q[3] / (gamma - 1.0) + kinetic;
}

/* state_primitives.hpp:57-63 */
static double minmod(double a, double b)
{
  if (a * b <= 0.0) return 0.0;
  return fabs(a) < fabs(b) ? a : b;
}

/* state_primitives.hpp:71-92 */
double ecor_limited_slope(double backward, double forward, int limiter)
{
  if (backward * forward <= 0.0) return 0.0;
  switch (limiter) {
    case ECOR_MINMOD:
      return minmod(backward, forward);
    case ECOR_MC:
      return minmod(minmod(2.0 * backward, 0.5 * (backward + forward)), 2.0 * forward);
    case ECOR_VANLEER: {
      const double denominator = backward + forward;
      if (fabs(denominator) <= 1.0e-14) return 0.0;
      return 2.0 * backward * forward / denominator;
    }
  }
  return minmod(backward, forward);
}

static double sound_speed(const double* q, double gamma)
{
  return sqrt(gamma * q[3] / q[0]); /* state_primitives.hpp:202-205 */
}

/* reconstruction_fluxes.hpp:66-74 */
void ecor_physical_flux(const double* q, int axis, double gamma, double* f)
{
  double u[4];
  ecor_to_conserved(q, gamma, u);
  if (axis == 0) {
    f[0] = u[1];
    f[1] = u[1] * q[1] + q[3];
    f[2] = u[2] * q[1];
    f[3] = (u[3] + q[3]) * q[1];
  } else {
    f[0] = u[2];
    f[1] = u[1] * q[2];
    f[2] = u[2] * q[2] + q[3];
    f[3] = (u[3] + q[3]) * q[2];
  }
}

static double normal_velocity(const double* q, int axis) { return axis == 0 ? q[1] : q[2]; }

/* reconstruction_fluxes.hpp:76-92 */
static void rusanov_flux(const double* L, const double* R, int axis, double gamma, double* f)
{
  double ul[4], ur[4], fl[4], fr[4];
  ecor_to_conserved(L, gamma, ul);
  ecor_to_conserved(R, gamma, ur);
  ecor_physical_flux(L, axis, gamma, fl);
  ecor_physical_flux(R, axis, gamma, fr);
  const double smax_ = smax(fabs(normal_velocity(L, axis)) + sound_speed(L, gamma),
                            fabs(normal_velocity(R, axis)) + sound_speed(R, gamma));
  for (int c = 0; c < 4; ++c) {
    f[c] = 0.5 * (fl[c] + fr[c]) - (0.5 * smax_) * (ur[c] - ul[c]);
  }
}

/* reconstruction_fluxes.hpp:106-115 */
static void hllc_star_state(const double* q, int axis, double s_m, double rho_s, double e_s,
                            double* us)
{
  const double tangential = axis == 0 ? q[2] : q[1];
  us[0] = rho_s;
  if (axis == 0) {
    us[1] = rho_s * s_m;
    us[2] = rho_s * tangential;
  } else {
    us[1] = rho_s * tangential;
    us[2] = rho_s * s_m;
  }
  us[3] = e_s;
}

/* reconstruction_fluxes.hpp:117-149 */
int ecor_hllc_star_degenerate(const double* L, const double* R, int axis, double gamma)
{
  const double rho_l = L[0], rho_r = R[0];
  const double un_l = normal_velocity(L, axis), un_r = normal_velocity(R, axis);
  const double p_l = L[3], p_r = R[3];
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double s_l = smin(un_l - c_l, un_r - c_r);
  const double s_r = smax(un_l + c_l, un_r + c_r);
  if (!(s_l < s_r)) return 1;
  if (s_l >= 0.0 || s_r <= 0.0) return 0;
  const double den = rho_l * (s_l - un_l) - rho_r * (s_r - un_r);
  if (fabs(den) <= 1.0e-12) return 1;
  const double s_m = (p_r - p_l + rho_l * un_l * (s_l - un_l) - rho_r * un_r * (s_r - un_r)) / den;
  return fabs(s_l - s_m) <= 1.0e-12 || fabs(s_r - s_m) <= 1.0e-12;
}

/* reconstruction_fluxes.hpp:151-210 */
static void hllc_flux(const double* L, const double* R, int axis, double gamma, double* f)
{
  double ul[4], ur[4], fl[4], fr[4];
  ecor_to_conserved(L, gamma, ul);
  ecor_to_conserved(R, gamma, ur);
  ecor_physical_flux(L, axis, gamma, fl);
  ecor_physical_flux(R, axis, gamma, fr);
  const double rho_l = L[0], rho_r = R[0];
  const double un_l = normal_velocity(L, axis), un_r = normal_velocity(R, axis);
  const double p_l = L[3], p_r = R[3];
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double s_l = smin(un_l - c_l, un_r - c_r);
  const double s_r = smax(un_l + c_l, un_r + c_r);
  if (!(s_l < s_r)) { rusanov_flux(L, R, axis, gamma, f); return; }
  if (s_l >= 0.0) { memcpy(f, fl, sizeof fl); return; }
  if (s_r <= 0.0) { memcpy(f, fr, sizeof fr); return; }
  const double den = rho_l * (s_l - un_l) - rho_r * (s_r - un_r);
  if (fabs(den) <= 1.0e-12) { rusanov_flux(L, R, axis, gamma, f); return; }
  const double s_m = (p_r - p_l + rho_l * un_l * (s_l - un_l) - rho_r * un_r * (s_r - un_r)) / den;
  const double lsd = s_l - s_m;
  const double rsd = s_r - s_m;
  if (fabs(lsd) <= 1.0e-12 || fabs(rsd) <= 1.0e-12) { rusanov_flux(L, R, axis, gamma, f); return; }
  const double p_star_l = p_l + rho_l * (s_l - un_l) * (s_m - un_l);
  const double p_star_r = p_r + rho_r * (s_r - un_r) * (s_m - un_r);
  const double p_star = 0.5 * (p_star_l + p_star_r);
  const double rho_star_l = rho_l * (s_l - un_l) / lsd;
  const double rho_star_r = rho_r * (s_r - un_r) / rsd;
  const double e_star_l = ((s_l - un_l) * ul[3] - p_l * un_l + p_star * s_m) / lsd;
  const double e_star_r = ((s_r - un_r) * ur[3] - p_r * un_r + p_star * s_m) / rsd;
  double us[4];
  if (s_m >= 0.0) {
    hllc_star_state(L, axis, s_m, rho_star_l, e_star_l, us);
    for (int c = 0; c < 4; ++c) f[c] = fl[c] + s_l * (us[c] - ul[c]);
  } else {
    hllc_star_state(R, axis, s_m, rho_star_r, e_star_r, us);
    for (int c = 0; c < 4; ++c) f[c] = fr[c] + s_r * (us[c] - ur[c]);
  }
}

/* reconstruction_fluxes.hpp:212-232 */
static void low_mach_correct(const double* L, const double* R, double gamma, double* cl,
                             double* cr)
{
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double speed_l = sqrt(L[1] * L[1] + L[2] * L[2]);
  const double speed_r = sqrt(R[1] * R[1] + R[2] * R[2]);
  const double m_l = speed_l / smax(c_l, 1.0e-12);
  const double m_r = speed_r / smax(c_r, 1.0e-12);
  const double z = smin(smax(smax(m_l, m_r), 0.0), 1.0);
  memcpy(cl, L, 4 * sizeof(double));
  memcpy(cr, R, 4 * sizeof(double));
  const double u_avg = 0.5 * (L[1] + R[1]);
  const double v_avg = 0.5 * (L[2] + R[2]);
  cl[1] = u_avg + 0.5 * z * (L[1] - R[1]);
  cr[1] = u_avg + 0.5 * z * (R[1] - L[1]);
  cl[2] = v_avg + 0.5 * z * (L[2] - R[2]);
  cr[2] = v_avg + 0.5 * z * (R[2] - L[2]);
}

/* reconstruction_fluxes.hpp:254-371 */
static void xie_hllcp_flux(const double* L, const double* R, int axis, double gamma,
                           double sensor_f, double* f)
{
  double ul[4], ur[4], fl[4], fr[4];
  ecor_to_conserved(L, gamma, ul);
  ecor_to_conserved(R, gamma, ur);
  ecor_physical_flux(L, axis, gamma, fl);
  ecor_physical_flux(R, axis, gamma, fr);
  const double rho_l = L[0], rho_r = R[0];
  const double q_l = normal_velocity(L, axis), q_r = normal_velocity(R, axis);
  const double p_l = L[3], p_r = R[3];
  const double c_l = sound_speed(L, gamma), c_r = sound_speed(R, gamma);
  const double s_l = smin(q_l - c_l, q_r - c_r);
  const double s_r = smax(q_l + c_l, q_r + c_r);
  if (!(s_l < s_r)) { hllc_flux(L, R, axis, gamma, f); return; }
  if (s_l >= 0.0) { memcpy(f, fl, sizeof fl); return; }
  if (s_r <= 0.0) { memcpy(f, fr, sizeof fr); return; }
  const double alpha_l = rho_l * (s_l - q_l);
  const double alpha_r = rho_r * (s_r - q_r);
  const double alpha_denom = alpha_r - alpha_l;
  if (fabs(alpha_denom) <= 1.0e-12) { hllc_flux(L, R, axis, gamma, f); return; }
  const double s_star = (alpha_r * q_r - alpha_l * q_l + p_l - p_r) / alpha_denom;
  const double rld = s_l - s_star;
  const double rrd = s_r - s_star;
  if (fabs(rld) <= 1.0e-12 || fabs(rrd) <= 1.0e-12 || fabs(alpha_l) <= 1.0e-12 ||
      fabs(alpha_r) <= 1.0e-12) {
    hllc_flux(L, R, axis, gamma, f);
    return;
  }
  const double rho_star_l = alpha_l / rld;
  const double rho_star_r = alpha_r / rrd;
  const double e_l = ul[3] / rho_l;
  const double e_r = ur[3] / rho_r;
  const double e_star_l = e_l + (s_star - q_l) * (s_star + p_l / alpha_l);
  const double e_star_r = e_r + (s_star - q_r) * (s_star + p_r / alpha_r);
  const double p_star = (alpha_r * p_l - alpha_l * p_r - alpha_l * alpha_r * (q_l - q_r)) / alpha_denom;
  const double sqrt_rho_l = sqrt(rho_l);
  const double sqrt_rho_r = sqrt(rho_r);
  const double roe_denom = sqrt_rho_l + sqrt_rho_r;
  if (!(roe_denom > 0.0)) { hllc_flux(L, R, axis, gamma, f); return; }
  const double u_hat = (sqrt_rho_l * L[1] + sqrt_rho_r * R[1]) / roe_denom;
  const double v_hat = (sqrt_rho_l * L[2] + sqrt_rho_r * R[2]) / roe_denom;
  const double q_hat = axis == 0 ? u_hat : v_hat;
  const double h_l = (ul[3] + p_l) / rho_l;
  const double h_r = (ur[3] + p_r) / rho_r;
  const double h_hat = (sqrt_rho_l * h_l + sqrt_rho_r * h_r) / roe_denom;
  const double c_hat_sq = (gamma - 1.0) * (h_hat - 0.5 * (u_hat * u_hat + v_hat * v_hat));
  if (!(c_hat_sq > 0.0) || !isfinite(c_hat_sq)) { hllc_flux(L, R, axis, gamma, f); return; }
  const double c_hat = sqrt(c_hat_sq);
  const double m_hat = q_hat / smax(c_hat, 1.0e-12);
  const double speed_l = sqrt(L[1] * L[1] + L[2] * L[2]);
  const double speed_r = sqrt(R[1] * R[1] + R[2] * R[2]);
  const double theta = clamp01(smax(speed_l / smax(c_l, 1.0e-12), speed_r / smax(c_r, 1.0e-12)));
  const double ff = clamp01(sensor_f);
  const double pss = theta * p_star + (1.0 - theta) * 0.5 * (p_l + p_r);
  const double psss = ff * pss + (1.0 - ff) * p_star;
  const double beta_p = (ff - 1.0) * s_l * s_r / (s_r - s_l) * (p_r - p_l) /
                        (c_hat_sq * (1.0 + fabs(m_hat)));
  const double phi[4] = {beta_p, beta_p * u_hat, beta_p * v_hat,
                         0.5 * beta_p * (u_hat * u_hat + v_hat * v_hat)};
  const double nx = axis == 0 ? 1.0 : 0.0;
  const double ny = axis == 1 ? 1.0 : 0.0;
  const double* q = s_star >= 0.0 ? L : R;
  const double qn = s_star >= 0.0 ? q_l : q_r;
  const double rs = s_star >= 0.0 ? rho_star_l : rho_star_r;
  const double es = s_star >= 0.0 ? e_star_l : e_star_r;
  double out[4];
  out[0] = rs * s_star;
  out[1] = rs * s_star * (q[1] + nx * (s_star - qn)) + nx * psss;
  out[2] = rs * s_star * (q[2] + ny * (s_star - qn)) + ny * psss;
  out[3] = s_star * (rs * es + p_star);
  for (int c = 0; c < 4; ++c) out[c] = out[c] + phi[c];
  if (!isfinite(out[0]) || !isfinite(out[1]) || !isfinite(out[2]) || !isfinite(out[3])) {
    hllc_flux(L, R, axis, gamma, f);
    return;
  }
  memcpy(f, out, sizeof out);
}

/* reconstruction_fluxes.hpp:373-388 */
void ecor_interface_flux(const double* L, const double* R, int axis, double gamma, int riemann,
                         double sensor_f, double* f)
{
  switch (riemann) {
    case ECOR_RUSANOV: rusanov_flux(L, R, axis, gamma, f); return;
    case ECOR_HLLC: hllc_flux(L, R, axis, gamma, f); return;
    case ECOR_LOWMACH_HLLC: {
      double cl[4], cr[4];
      low_mach_correct(L, R, gamma, cl, cr);
      hllc_flux(cl, cr, axis, gamma, f);
      return;
    }
    case ECOR_XIE_HLLCP: xie_hllcp_flux(L, R, axis, gamma, sensor_f, f); return;
  }
  hllc_flux(L, R, axis, gamma, f);
}

/* reconstruction_fluxes.hpp:390-397 */
double ecor_pressure_ratio(double pL, double pR)
{
  const double floor_ = 1.0e-14;
  const double p_l = isfinite(pL) ? smax(pL, floor_) : floor_;
  const double p_r = isfinite(pR) ? smax(pR, floor_) : floor_;
  return clamp01(smin(p_l / p_r, p_r / p_l));
}

/* ------------------------------------------------------------------------ */
/* Ghost rule.  boundary_conditions.hpp:2-25 (FillBoundary periodic copy then
 * clamp for non-periodic dims) and :27-73 (fixed inflow state at x-low with
 * priority, mirror with negated normal momentum at a reflecting side). */
static int map_index(int i, int n, int lo_bc, int hi_bc, int* reflected)
{
  *reflected = 0;
  if (i < 0) {
    if (lo_bc == ECOR_BC_PERIODIC) return ((i % n) + n) % n;
    if (lo_bc == ECOR_BC_REFLECT) { *reflected = 1; i = -i - 1; }
    return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
  }
  if (i > n - 1) {
    if (hi_bc == ECOR_BC_PERIODIC) return i % n;
    if (hi_bc == ECOR_BC_REFLECT) { *reflected = 1; i = 2 * n - 1 - i; }
    return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
  }
  return i;
}

static void gresho_prim(double x, double y, double gamma, double cx, double cy, double rho,
                        double mach, double p0, double* q);

/* The exact branches of fill_problem_ghosts (boundary_conditions.hpp:165-219):
 * active when an axis is not periodic; the physical ghosts are those of a
 * non-periodic axis (corners included), and every x ghost for Toro
 * (problem_initialization.hpp:329-330). */
int ecor_exact_ghost(const ecor_grid* g, int i, int j)
{
  const ecor_exact* e = g->exact;
  if (!e || e->kind == ECOR_EXACT_NONE || (e->periodic[0] && e->periodic[1])) return 0;
  const int x_out = i < 0 || i >= g->nx, y_out = j < 0 || j >= g->ny;
  if (e->kind == ECOR_EXACT_TORO) return x_out || (!e->periodic[1] && y_out);
  return (!e->periodic[0] && x_out) || (!e->periodic[1] && y_out);
}

/* Gresho (boundary_conditions.hpp:133-163) and quadrant
 * (problem_initialization.hpp:344-375) at the ghost's own centre; Toro at the
 * ghost's x and time (:308-342); the blob at the periodic image of a periodic
 * axis, its centre advected to the time (:209-263). */
void ecor_exact_state(const ecor_exact* e, int nx, int ny, double time, int i, int j, double* u)
{
  const double dx = (e->prob_hi[0] - e->prob_lo[0]) / nx;
  const double dy = (e->prob_hi[1] - e->prob_lo[1]) / ny;
  double q[4];
  if (e->kind == ECOR_EXACT_BLOB) {
    int ii = i, jj = j;
    if (e->periodic[0]) { const int o = i % nx; ii = o < 0 ? o + nx : o; }
    if (e->periodic[1]) { const int o = j % ny; jj = o < 0 ? o + ny : o; }
    const double x = e->prob_lo[0] + ((double)ii + 0.5) * dx;
    const double y = e->prob_lo[1] + ((double)jj + 0.5) * dy;
    const double cx = e->blob_center[0] + e->velocity[0] * time;
    const double cy = e->blob_center[1] + e->velocity[1] * time;
    const double r2 = (x - cx) * (x - cx) + (y - cy) * (y - cy);
    const double rho = r2 <= e->blob_radius * e->blob_radius ? e->density_inner : e->density_outer;
    q[0] = rho; q[1] = e->velocity[0]; q[2] = e->velocity[1]; q[3] = e->pressure;
    ecor_to_conserved(q, e->gamma, u);
    return;
  }
  const double x = e->prob_lo[0] + ((double)i + 0.5) * dx;
  const double y = e->prob_lo[1] + ((double)j + 0.5) * dy;
  ecor_exact_prim_at(e, x, y, time, q);
  ecor_to_conserved(q, e->gamma, u);
}

void ecor_exact_prim_at(const ecor_exact* e, double x, double y, double time, double* q)
{
  if (e->kind == ECOR_EXACT_GRESHO) {
    gresho_prim(x, y, e->gamma, e->vortex_center[0], e->vortex_center[1], e->density_outer,
                e->mach, e->pressure, q);
  } else if (e->kind == ECOR_EXACT_TORO) {
    ecor_exact_riemann_sample(x, time, e->ql, e->qr, e->gamma, e->interface_x, q);
  } else {
    /* problem_state_functions.hpp:73-98 */
    const int right = x >= e->interface_x, top = y >= e->interface_y;
    if (right && top) { q[0] = 1.5; q[1] = 0.0; q[2] = 0.0; q[3] = 1.5; }
    else if (!right && top) { q[0] = 0.532258064516129; q[1] = 1.206045378311055; q[2] = 0.0; q[3] = 0.3; }
    else if (!right && !top) {
      q[0] = 0.137992831541219; q[1] = 1.206045378311055; q[2] = 1.206045378311055;
      q[3] = 0.029032258064516;
    } else { q[0] = 0.532258064516129; q[1] = 0.0; q[2] = 1.206045378311055; q[3] = 0.3; }
  }
}

void ecor_load(const ecor_grid* g, const double* U, int i, int j, double* u)
{
  const int nx = g->nx, ny = g->ny;
  if (ecor_exact_ghost(g, i, j)) { ecor_exact_state(g->exact, nx, ny, g->time, i, j, u); return; }
  if (i < 0 && g->bc[0] == ECOR_BC_FIXED) { memcpy(u, g->fixed[0], 4 * sizeof(double)); return; }
  if (i >= nx && g->bc[1] == ECOR_BC_FIXED) { memcpy(u, g->fixed[1], 4 * sizeof(double)); return; }
  if (j < 0 && g->bc[2] == ECOR_BC_FIXED) { memcpy(u, g->fixed[2], 4 * sizeof(double)); return; }
  if (j >= ny && g->bc[3] == ECOR_BC_FIXED) { memcpy(u, g->fixed[3], 4 * sizeof(double)); return; }
  int rx, ry;
  const int ii = map_index(i, nx, g->bc[0], g->bc[1], &rx);
  const int jj = map_index(j, ny, g->bc[2], g->bc[3], &ry);
  const long plane = (long)nx * ny;
  const long k = (long)jj * nx + ii;
  u[0] = U[k];
  u[1] = U[plane + k];
  u[2] = U[2 * plane + k];
  u[3] = U[3 * plane + k];
  if (rx) u[1] = -u[1];
  if (ry) u[2] = -u[2];
}

static void load_prim(const ecor_grid* g, const double* U, int i, int j, double* q)
{
  double u[4];
  ecor_load(g, U, i, j, u);
  ecor_to_primitive(u, g->gamma, q);
}

/* reconstruction_fluxes.hpp:2-24 */
void ecor_reconstruct(const ecor_grid* g, const double* U, int i, int j, int axis, int side,
                      double* q)
{
  load_prim(g, U, i, j, q);
  if (g->order <= 1) return;
  double qm[4], qc[4], qp[4];
  if (axis == 0) {
    load_prim(g, U, i - 1, j, qm);
    load_prim(g, U, i, j, qc);
    load_prim(g, U, i + 1, j, qp);
  } else {
    load_prim(g, U, i, j - 1, qm);
    load_prim(g, U, i, j, qc);
    load_prim(g, U, i, j + 1, qp);
  }
  const double s = 0.5 * side;
  for (int c = 0; c < 4; ++c) {
    const double slope = ecor_limited_slope(qc[c] - qm[c], qp[c] - qc[c], g->limiter);
    q[c] = q[c] + s * slope;
  }
}

static double x_face_ratio(const ecor_grid* g, const double* U, int fi, int j)
{
  double l[4], r[4];
  ecor_reconstruct(g, U, fi - 1, j, 0, +1, l);
  ecor_reconstruct(g, U, fi, j, 0, -1, r);
  return ecor_pressure_ratio(l[3], r[3]);
}

static double y_face_ratio(const ecor_grid* g, const double* U, int i, int fj)
{
  double l[4], r[4];
  ecor_reconstruct(g, U, i, fj - 1, 1, +1, l);
  ecor_reconstruct(g, U, i, fj, 1, -1, r);
  return ecor_pressure_ratio(l[3], r[3]);
}

/* reconstruction_fluxes.hpp:421-440 */
static double multiface_sensor(const ecor_grid* g, const double* U, int fi, int fj, int axis)
{
  double f;
  if (axis == 0) {
    f = x_face_ratio(g, U, fi, fj);
    f = smin(f, y_face_ratio(g, U, fi - 1, fj));
    f = smin(f, y_face_ratio(g, U, fi - 1, fj + 1));
    f = smin(f, y_face_ratio(g, U, fi, fj));
    f = smin(f, y_face_ratio(g, U, fi, fj + 1));
  } else {
    f = y_face_ratio(g, U, fi, fj);
    f = smin(f, x_face_ratio(g, U, fi, fj - 1));
    f = smin(f, x_face_ratio(g, U, fi + 1, fj - 1));
    f = smin(f, x_face_ratio(g, U, fi, fj));
    f = smin(f, x_face_ratio(g, U, fi + 1, fj));
  }
  const double c = clamp01(f);
  return c * c * c;
}

/* reconstruction_fluxes.hpp:442-460 */
void ecor_face_flux(const ecor_grid* g, const double* U, int fi, int fj, int axis, double* f)
{
  double l[4], r[4];
  if (axis == 0) {
    ecor_reconstruct(g, U, fi - 1, fj, 0, +1, l);
    ecor_reconstruct(g, U, fi, fj, 0, -1, r);
  } else {
    ecor_reconstruct(g, U, fi, fj - 1, 1, +1, l);
    ecor_reconstruct(g, U, fi, fj, 1, -1, r);
  }
  const double sensor = g->riemann == ECOR_XIE_HLLCP ? multiface_sensor(g, U, fi, fj, axis) : 1.0;
  ecor_interface_flux(l, r, axis, g->gamma, g->riemann, sensor, f);
}

static int face_degenerate(const ecor_grid* g, const double* U, int fi, int fj, int axis)
{
  double l[4], r[4];
  if (axis == 0) {
    ecor_reconstruct(g, U, fi - 1, fj, 0, +1, l);
    ecor_reconstruct(g, U, fi, fj, 0, -1, r);
  } else {
    ecor_reconstruct(g, U, fi, fj - 1, 1, +1, l);
    ecor_reconstruct(g, U, fi, fj, 1, -1, r);
  }
  if (g->riemann == ECOR_LOWMACH_HLLC) {
    double cl[4], cr[4];
    low_mach_correct(l, r, g->gamma, cl, cr);
    return ecor_hllc_star_degenerate(cl, cr, axis, g->gamma);
  }
  return ecor_hllc_star_degenerate(l, r, axis, g->gamma);
}

/* explicit_update.hpp:2-56 (degenerate count) and :58-102 (the step). */
int ecor_fv_step(const ecor_grid* g, const double* U, double* Unew, double dt)
{
  const int nx = g->nx, ny = g->ny;
  const long plane = (long)nx * ny;
  int degenerate = 0;
  if (g->riemann != ECOR_RUSANOV) {
    for (int j = 0; j < ny; ++j) {
      for (int i = 0; i < nx; ++i) {
        degenerate += face_degenerate(g, U, i, j, 0);
        degenerate += face_degenerate(g, U, i + 1, j, 0);
        degenerate += face_degenerate(g, U, i, j, 1);
        degenerate += face_degenerate(g, U, i, j + 1, 1);
      }
    }
  }
  const double inv_dx = 1.0 / g->dx;
  const double inv_dy = 1.0 / g->dy;
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      double fxl[4], fxr[4], fyb[4], fyt[4];
      ecor_face_flux(g, U, i, j, 0, fxl);
      ecor_face_flux(g, U, i + 1, j, 0, fxr);
      ecor_face_flux(g, U, i, j, 1, fyb);
      ecor_face_flux(g, U, i, j + 1, 1, fyt);
      const long k = (long)j * nx + i;
      for (int c = 0; c < 4; ++c) {
        const double old = U[c * plane + k];
        Unew[c * plane + k] =
            old - dt * ((fxr[c] - fxl[c]) * inv_dx + (fyt[c] - fyb[c]) * inv_dy);
      }
    }
  }
  return degenerate;
}

/* diagnostic_metrics_timestep.hpp:227-246 (the max-rate reduction). */
double ecor_max_rate(const ecor_grid* g, const double* U)
{
  const long plane = (long)g->nx * g->ny;
  double max_rate = -INFINITY; /* ReduceOpMax identity */
  for (long k = 0; k < plane; ++k) {
    const double u[4] = {U[k], U[plane + k], U[2 * plane + k], U[3 * plane + k]};
    double q[4];
    ecor_to_primitive(u, g->gamma, q);
    const double c = sound_speed(q, g->gamma);
    const double rate = (fabs(q[1]) + c) / g->dx + (fabs(q[2]) + c) / g->dy;
    max_rate = smax(max_rate, rate);
  }
  return max_rate;
}

/* diagnostic_metrics_timestep.hpp:252-259 */
double ecor_euler_dt(double max_rate, double cfl, double time, double stop_time)
{
  double dt = stop_time - time;
  if (max_rate > 0.0) dt = cfl / max_rate;
  if (!isfinite(dt) || dt <= 0.0) dt = stop_time - time;
  return smin(dt, stop_time - time);
}

/* diagnostic_metrics_timestep.hpp:18-47 */
void ecor_diagnostics(const ecor_grid* g, const double* U, double* rho_min, double* p_min,
                      int* nonfinite)
{
  const long plane = (long)g->nx * g->ny;
  double rmin = INFINITY, pmin = INFINITY;
  int bad_count = 0;
  for (long k = 0; k < plane; ++k) {
    const double u[4] = {U[k], U[plane + k], U[2 * plane + k], U[3 * plane + k]};
    double q[4];
    ecor_to_primitive(u, g->gamma, q);
    const int bad = !isfinite(q[0]) || !isfinite(q[1]) || !isfinite(q[2]) || !isfinite(q[3]) ||
                    !isfinite(u[3]);
    rmin = smin(rmin, bad ? -INFINITY : q[0]);
    pmin = smin(pmin, bad ? -INFINITY : q[3]);
    bad_count += bad;
  }
  *rho_min = rmin;
  *p_min = pmin;
  *nonfinite = bad_count;
}

/* driver_main.hpp:216-340, explicit method, no snapshot clipping. */
void ecor_run_explicit(const ecor_grid* g, double* U, double cfl, double stop_time,
                       int max_step, double* dts, ecor_run_result* out)
{
  const long n = 4L * g->nx * g->ny;
  double* next = (double*)malloc((size_t)n * sizeof(double));
  memset(out, 0, sizeof *out);
  double time = 0.0;
  int step = 0;
  ecor_diagnostics(g, U, &out->rho_min, &out->p_min, &out->nonfinite);
  if (out->nonfinite > 0 || out->rho_min <= 0.0 || out->p_min <= 0.0) out->status = 1;
  ecor_grid gt = *g; /* the step's time for exact ghost states (explicit_update.hpp:65) */
  while (out->status == 0 && step < max_step && time < stop_time - 1.0e-14) {
    const double dt = ecor_euler_dt(ecor_max_rate(g, U), cfl, time, stop_time);
    if (!isfinite(dt) || dt <= 0.0) { out->status = 2; break; }
    gt.time = time;
    out->degenerate_total += ecor_fv_step(&gt, U, next, dt);
    memcpy(U, next, (size_t)n * sizeof(double));
    if (dts) dts[step] = dt;
    out->dt_last = dt;
    time += dt;
    ++step;
    ecor_diagnostics(g, U, &out->rho_min, &out->p_min, &out->nonfinite);
    if (out->nonfinite > 0) out->status = 3;
    else if (out->rho_min <= 0.0) out->status = 4;
    else if (out->p_min <= 0.0) out->status = 5;
  }
  out->steps = step;
  out->time = time;
  free(next);
}

/* ------------------------------------------------------------------------ */
/* problem_state_functions.hpp:2-30 with velocity_scale = mach * sqrt(gamma p0 / rho)
 * (state_primitives.hpp:207-215), and problem_initialization.hpp:377-392. */
static void gresho_prim(double x, double y, double gamma, double cx, double cy, double rho,
                        double mach, double p0, double* q)
{
  const double vscale = mach * sqrt(gamma * p0 / rho);
  const double ddx = x - cx, ddy = y - cy;
  const double r = sqrt(ddx * ddx + ddy * ddy);
  double tf = 0.0;
  double pf = -2.0 + 4.0 * log(2.0);
  if (r < 0.2) {
    tf = 5.0 * r;
    pf = 12.5 * r * r;
  } else if (r < 0.4) {
    tf = 2.0 - 5.0 * r;
    pf = 12.5 * r * r + 4.0 * (1.0 - 5.0 * r - log(0.2) + log(r));
  }
  const double vt = vscale * tf;
  const double p = p0 + rho * vscale * vscale * pf;
  q[0] = rho; q[1] = 0.0; q[2] = 0.0; q[3] = p;
  if (!(r < 1.0e-14)) {
    q[1] = -vt * ddy / r;
    q[2] = vt * ddx / r;
  }
}

void ecor_init_gresho(int nx, int ny, double xlo, double ylo, double dx, double dy,
                      double gamma, double cx, double cy, double rho, double mach, double p0,
                      double* U)
{
  const long plane = (long)nx * ny;
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      const double x = xlo + ((double)i + 0.5) * dx;
      const double y = ylo + ((double)j + 0.5) * dy;
      double q[4];
      gresho_prim(x, y, gamma, cx, cy, rho, mach, p0, q);
      double u[4];
      ecor_to_conserved(q, gamma, u);
      const long k = (long)j * nx + i;
      for (int c = 0; c < 4; ++c) U[c * plane + k] = u[c];
    }
  }
}

/* problem_initialization.hpp:265-284 */
void ecor_init_toro(int nx, int ny, double xlo, double dx, double gamma, const double* qL,
                    const double* qR, double x0, double* U)
{
  const long plane = (long)nx * ny;
  double uL[4], uR[4];
  ecor_to_conserved(qL, gamma, uL);
  ecor_to_conserved(qR, gamma, uR);
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      const double x = xlo + ((double)i + 0.5) * dx;
      const double* u = x < x0 ? uL : uR;
      const long k = (long)j * nx + i;
      for (int c = 0; c < 4; ++c) U[c * plane + k] = u[c];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Exact Riemann solver (Toro ch. 4), exact_riemann.hpp:37-205. */
static void pressure_function(double p, const double* q, double gamma, double* value,
                              double* derivative)
{
  const double a = sqrt(gamma * q[3] / q[0]);
  if (p > q[3]) {
    const double acoef = 2.0 / ((gamma + 1.0) * q[0]);
    const double bcoef = (gamma - 1.0) / (gamma + 1.0) * q[3];
    const double root = sqrt(acoef / (p + bcoef));
    *value = (p - q[3]) * root;
    *derivative = root * (1.0 - 0.5 * (p - q[3]) / (p + bcoef));
    return;
  }
  const double ratio = p / q[3];
  *value = 2.0 * a / (gamma - 1.0) * (pow(ratio, (gamma - 1.0) / (2.0 * gamma)) - 1.0);
  *derivative = 1.0 / (q[0] * a) * pow(ratio, -(gamma + 1.0) / (2.0 * gamma));
}

void ecor_exact_riemann_star(const double* L, const double* R, double gamma, double* p_star,
                             double* u_star)
{
  const double a_l = sqrt(gamma * L[3] / L[0]);
  const double a_r = sqrt(gamma * R[3] / R[0]);
  double p = smax(1.0e-12, 0.5 * (L[3] + R[3]) - 0.125 * (R[1] - L[1]) * (L[0] + R[0]) * (a_l + a_r));
  for (int iter = 0; iter < 80; ++iter) {
    double f_l, d_l, f_r, d_r;
    pressure_function(p, L, gamma, &f_l, &d_l);
    pressure_function(p, R, gamma, &f_r, &d_r);
    const double residual = f_l + f_r + R[1] - L[1];
    double p_next = p - residual / (d_l + d_r);
    if (!(p_next > 0.0) || !isfinite(p_next)) p_next = 0.5 * p;
    if (fabs(p_next - p) <= 1.0e-12 * smax(1.0, p_next)) { p = p_next; break; }
    p = p_next;
  }
  double f_l, d_l, f_r, d_r;
  pressure_function(p, L, gamma, &f_l, &d_l);
  pressure_function(p, R, gamma, &f_r, &d_r);
  *p_star = p;
  *u_star = 0.5 * (L[1] + R[1] + f_r - f_l);
}

static void set4(double* q, double a, double b, double c, double d)
{
  q[0] = a; q[1] = b; q[2] = c; q[3] = d;
}

void ecor_exact_riemann_sample(double x, double t, const double* L, const double* R, double gamma,
                               double x0, double* q)
{
  if (t <= 0.0) { memcpy(q, x < x0 ? L : R, 4 * sizeof(double)); return; }
  double p_star, u_star;
  ecor_exact_riemann_star(L, R, gamma, &p_star, &u_star);
  const double xi = (x - x0) / t;
  const double a_l = sqrt(gamma * L[3] / L[0]);
  const double a_r = sqrt(gamma * R[3] / R[0]);
  const double gm1 = gamma - 1.0, gp1 = gamma + 1.0;
  if (xi <= u_star) {
    if (p_star > L[3]) {
      const double speed = L[1] - a_l * sqrt(gp1 / (2.0 * gamma) * p_star / L[3] + gm1 / (2.0 * gamma));
      if (xi <= speed) { memcpy(q, L, 4 * sizeof(double)); return; }
      const double ratio = p_star / L[3];
      set4(q, L[0] * ((ratio + gm1 / gp1) / ((gm1 / gp1) * ratio + 1.0)), u_star, L[2], p_star);
      return;
    }
    const double head = L[1] - a_l;
    const double a_star = a_l * pow(p_star / L[3], gm1 / (2.0 * gamma));
    const double tail = u_star - a_star;
    if (xi <= head) { memcpy(q, L, 4 * sizeof(double)); return; }
    if (xi >= tail) { set4(q, L[0] * pow(p_star / L[3], 1.0 / gamma), u_star, L[2], p_star); return; }
    const double u = 2.0 / gp1 * (a_l + 0.5 * gm1 * L[1] + xi);
    const double a = 2.0 / gp1 * (a_l + 0.5 * gm1 * (L[1] - xi));
    set4(q, L[0] * pow(a / a_l, 2.0 / gm1), u, L[2], L[3] * pow(a / a_l, 2.0 * gamma / gm1));
    return;
  }
  if (p_star > R[3]) {
    const double speed = R[1] + a_r * sqrt(gp1 / (2.0 * gamma) * p_star / R[3] + gm1 / (2.0 * gamma));
    if (xi >= speed) { memcpy(q, R, 4 * sizeof(double)); return; }
    const double ratio = p_star / R[3];
    set4(q, R[0] * ((ratio + gm1 / gp1) / ((gm1 / gp1) * ratio + 1.0)), u_star, R[2], p_star);
    return;
  }
  const double head = R[1] + a_r;
  const double a_star = a_r * pow(p_star / R[3], gm1 / (2.0 * gamma));
  const double tail = u_star + a_star;
  if (xi >= head) { memcpy(q, R, 4 * sizeof(double)); return; }
  if (xi <= tail) { set4(q, R[0] * pow(p_star / R[3], 1.0 / gamma), u_star, R[2], p_star); return; }
  const double u = 2.0 / gp1 * (-a_r + 0.5 * gm1 * R[1] + xi);
  const double a = 2.0 / gp1 * (a_r - 0.5 * gm1 * (R[1] - xi));
  set4(q, R[0] * pow(a / a_r, 2.0 / gm1), u, R[2], R[3] * pow(a / a_r, 2.0 * gamma / gm1));
}

/* ------------------------------------------------------------------------ */
/* compute_admissibility_diagnostics (diagnostic_metrics_timestep.hpp:49-102),
 * serially in cell order.  out: rho_min, p_min, ie_min; counts: nonfinite,
 * density, pressure, internal energy failures. */
void ecor_admissibility(const ecor_grid* g, const double* U, double* out, int* counts)
{
  const long n = (long)g->nx * g->ny;
  double mins[3] = {INFINITY, INFINITY, INFINITY};
  int c[4] = {0, 0, 0, 0};
  for (long k = 0; k < n; ++k) {
    const double r = U[k], mx = U[n + k], my = U[2 * n + k], E = U[3 * n + k];
    const int cons_bad = !isfinite(r) || !isfinite(mx) || !isfinite(my) || !isfinite(E);
    const int density_bad = !cons_bad && r <= 0.0;
    double ie = -INFINITY, p = -INFINITY;
    int thermo_bad = 0;
    if (!cons_bad && !density_bad) {
      const double kinetic = 0.5 * (mx * mx + my * my) / r;
      ie = E - kinetic;
      p = (g->gamma - 1.0) * ie;
      thermo_bad = !isfinite(kinetic) || !isfinite(ie) || !isfinite(p);
    }
    const int bad = cons_bad || thermo_bad;
    const double v[3] = {bad ? -INFINITY : r, (bad || density_bad) ? -INFINITY : p,
                         (bad || density_bad) ? -INFINITY : ie};
    for (int q = 0; q < 3; ++q) mins[q] = smin(mins[q], v[q]);
    c[0] += bad;
    c[1] += density_bad;
    c[2] += !bad && !density_bad && p <= 0.0;
    c[3] += !bad && !density_bad && ie <= 0.0;
  }
  memcpy(out, mins, sizeof mins);
  memcpy(counts, c, sizeof c);
}

/* compute_gresho_metrics (diagnostic_metrics_timestep.hpp:147-225).  out[10]:
 * density, velocity, pressure L1 errors, pressure-perturbation L1 error and
 * its relative form, kinetic energy initial / final / ratio, reference sound
 * speed, velocity scale. */
void ecor_gresho_metrics(const ecor_grid* g, const double* U, double xlo, double ylo, double cx,
                         double cy, double rho, double mach, double p0, double* out)
{
  const int nx = g->nx, ny = g->ny;
  const long n = (long)nx * ny;
  const double vol = g->dx * g->dy;
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      const long k = (long)j * nx + i;
      const double x = xlo + ((double)i + 0.5) * g->dx, y = ylo + ((double)j + 0.5) * g->dy;
      const double u[4] = {U[k], U[n + k], U[2 * n + k], U[3 * n + k]};
      double q[4], e[4];
      ecor_to_primitive(u, g->gamma, q);
      gresho_prim(x, y, g->gamma, cx, cy, rho, mach, p0, e);
      const double du = q[1] - e[1], dv = q[2] - e[2];
      s[0] += fabs(q[0] - e[0]);
      s[1] += sqrt(du * du + dv * dv);
      s[2] += fabs(q[3] - e[3]);
      s[3] += fabs((q[3] - p0) - (e[3] - p0));
      s[4] += fabs(e[3] - p0);
      s[5] += 0.5 * e[0] * (e[1] * e[1] + e[2] * e[2]) * vol;
      s[6] += 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]) * vol;
    }
  }
  const double inv = 1.0 / (double)(nx * ny);
  const double c0 = sqrt(g->gamma * p0 / rho);
  out[0] = s[0] * inv;
  out[1] = s[1] * inv;
  out[2] = s[2] * inv;
  out[3] = s[3] * inv;
  out[4] = s[4] > 1.0e-30 ? s[3] / s[4] : NAN;
  out[5] = s[5];
  out[6] = s[6];
  out[7] = s[5] > 0.0 ? s[6] / s[5] : NAN;
  out[8] = c0;
  out[9] = mach * c0;
}

/* compute_shock_density_bubble_metrics (io/output_cases.hpp:329-407), the
 * neighbour densities through the grid's ghost rule.  out[10]: rho min/max,
 * p min/max, internal energy min, area, centroid x/y, schlieren max,
 * nonfinite count. */
void ecor_shock_bubble_metrics(const ecor_grid* g, const double* U, double xlo, double ylo,
                               double threshold, double* out)
{
  const int nx = g->nx, ny = g->ny;
  const long n = (long)nx * ny;
  const double vol = g->dx * g->dy;
  double rmin = INFINITY, rmax = -INFINITY, pmin = INFINITY, pmax = -INFINITY, emin = INFINITY;
  double area = 0.0, sx = 0.0, sy = 0.0, schl = 0.0;
  int bad_count = 0;
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      const long k = (long)j * nx + i;
      const double u[4] = {U[k], U[n + k], U[2 * n + k], U[3 * n + k]};
      double q[4];
      ecor_to_primitive(u, g->gamma, q);
      const double kinetic = u[0] > 0.0 ? 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0] : NAN;
      const double internal = u[3] - kinetic;
      if (!isfinite(u[0]) || !isfinite(u[1]) || !isfinite(u[2]) || !isfinite(u[3]) ||
          !isfinite(q[3]) || !isfinite(internal)) {
        ++bad_count;
        continue;
      }
      rmin = smin(rmin, q[0]);
      rmax = smax(rmax, q[0]);
      pmin = smin(pmin, q[3]);
      pmax = smax(pmax, q[3]);
      emin = smin(emin, internal);
      double a[4], b[4], c[4], d[4];
      ecor_load(g, U, i + 1, j, a);
      ecor_load(g, U, i - 1, j, b);
      ecor_load(g, U, i, j + 1, c);
      ecor_load(g, U, i, j - 1, d);
      const double rx = (a[0] - b[0]) / (2.0 * g->dx), ry = (c[0] - d[0]) / (2.0 * g->dy);
      const double grad = sqrt(rx * rx + ry * ry);
      if (isfinite(grad)) schl = smax(schl, grad);
      if (q[0] < threshold) {
        const double x = xlo + ((double)i + 0.5) * g->dx, y = ylo + ((double)j + 0.5) * g->dy;
        area += vol;
        sx += x * vol;
        sy += y * vol;
      }
    }
  }
  out[0] = rmin; out[1] = rmax; out[2] = pmin; out[3] = pmax; out[4] = emin;
  out[5] = area;
  out[6] = area > 0.0 ? sx / area : NAN;
  out[7] = area > 0.0 ? sy / area : NAN;
  out[8] = schl;
  out[9] = bad_count;
}

/* apply_shock_density_bubble_cylindrical_source
 * (problems/boundary_conditions.hpp:75-131): y is the radius; the source
 * -(rho v, rho u v, rho v^2, v (E + p)) / r is integrated over dt by the
 * midpoint rule, evaluated at q0 for the half step and at the half-step
 * state for the full step from q0.  Rows with r <= 0 are skipped. */
void ecor_cylindrical_source(const ecor_grid* g, double* U, double dt, double ylo)
{
  const int nx = g->nx, ny = g->ny;
  const long plane = (long)nx * ny;
  const double gm1 = g->gamma - 1.0;
  const double half_dt = 0.5 * dt;
  for (int j = 0; j < ny; ++j) {
    const double r = ylo + ((double)j + 0.5) * g->dy;
    if (!(r > 0.0)) continue;
    for (int i = 0; i < nx; ++i) {
      const long k = (long)j * nx + i;
      const double rho0 = U[k], mx0 = U[plane + k], my0 = U[2 * plane + k], E0 = U[3 * plane + k];
      const double ir0 = rho0 > 0.0 ? 1.0 / rho0 : 0.0;
      const double u0 = mx0 * ir0, v0 = my0 * ir0;
      const double p0 = gm1 * (E0 - 0.5 * (mx0 * mx0 + my0 * my0) * ir0);
      const double rhos = rho0 - half_dt * my0 / r;
      const double mxs = mx0 - half_dt * rho0 * u0 * v0 / r;
      const double mys = my0 - half_dt * rho0 * v0 * v0 / r;
      const double Es = E0 - half_dt * v0 * (E0 + p0) / r;
      const double irs = rhos > 0.0 ? 1.0 / rhos : 0.0;
      const double us = mxs * irs, vs = mys * irs;
      const double ps = gm1 * (Es - 0.5 * (mxs * mxs + mys * mys) * irs);
      U[k] = rho0 - dt * mys / r;
      U[plane + k] = mx0 - dt * rhos * us * vs / r;
      U[2 * plane + k] = my0 - dt * rhos * vs * vs / r;
      U[3 * plane + k] = E0 - dt * vs * (Es + ps) / r;
    }
  }
}
