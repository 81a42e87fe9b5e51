/*
 * ORACLE -- test infrastructure only.
 *
 * The direct BDLTV20 T1/S2 step (amrex/apps/euler_compare/src/euler_compare/
 * imex/imex_bdltv20_direct_driver.hpp:44-779), restated serially in C in the
 * reference's order of operations: the epsilon-scaled state
 * (core/state_primitives.hpp:119-139), the star update with conserved minmod
 * and the material Lax-Friedrichs flux (:176-253), face enthalpies from the
 * corrected cells (:267-383), the pressure rows with exact-Dirichlet columns
 * folded into the right-hand side (:424-494), the reference's full GMRES
 * (ecor_gmres_full), the pressure-gradient momentum correction and the
 * pressure closeout (:558-598), and the eq79 time step (:8-41).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ec_oracle.h"

static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

static int wrap(int i, int n) { return ((i % n) + n) % n; }

static double minmod(double a, double b)
{
  if (a * b <= 0.0) return 0.0;
  return fabs(a) < fabs(b) ? a : b;
}

/* core/state_primitives.hpp:119-139 */
void ecor_eps_to_primitive(const double* u, double gamma, double eps, double* q)
{
  const double inv_rho = 1.0 / u[0];
  const double ux = u[1] * inv_rho;
  const double uy = u[2] * inv_rho;
  const double kinetic = 0.5 * u[0] * (ux * ux + uy * uy);
  q[0] = u[0];
  q[1] = ux;
  q[2] = uy;
  q[3] = (gamma - 1.0) * (u[3] - eps * kinetic);
}

void ecor_eps_to_conserved(const double* q, double gamma, double eps, double* u)
{
  const double kinetic = 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]);
  u[0] = q[0];
  u[1] = q[0] * q[1];
  u[2] = q[0] * q[2];
  u[3] = q[3] / (gamma - 1.0) + eps * kinetic;
}

/* exact_boundary_state (:101-131): the exact state at the boundary point of
 * an outside index (the domain edge along an outside axis, the cell centre
 * along the other). */
void ecor_direct_boundary_state(const ecor_grid* g, const ecor_direct* d, int i, int j, double t,
                                double* u)
{
  const ecor_exact* e = &d->exact;
  const double dx = (e->prob_hi[0] - e->prob_lo[0]) / g->nx;
  const double dy = (e->prob_hi[1] - e->prob_lo[1]) / g->ny;
  const int ic = i < 0 ? 0 : (i > g->nx - 1 ? g->nx - 1 : i);
  const int jc = j < 0 ? 0 : (j > g->ny - 1 ? g->ny - 1 : j);
  double x = e->prob_lo[0] + ((double)ic + 0.5) * dx;
  double y = e->prob_lo[1] + ((double)jc + 0.5) * dy;
  if (i < 0) x = e->prob_lo[0];
  else if (i >= g->nx) x = e->prob_hi[0];
  if (j < 0) y = e->prob_lo[1];
  else if (j >= g->ny) y = e->prob_hi[1];
  double q[4];
  ecor_exact_prim_at(e, x, y, t, q);
  ecor_eps_to_conserved(q, e->gamma, d->epsilon, u);
}

static double boundary_pressure(const ecor_grid* g, const ecor_direct* d, int i, int j, double t)
{
  double u[4], q[4];
  ecor_direct_boundary_state(g, d, i, j, t, u);
  ecor_eps_to_primitive(u, g->gamma, d->epsilon, q);
  return q[3];
}

typedef struct {
  const ecor_grid* g;
  const ecor_direct* d;
  int px, py;
} Ctx;

/* cell_state_from / old_cell_from: wrap a periodic axis, else the boundary state */
static int cell_from(const Ctx* c, int i, int j, const double* cells, double t, double* u)
{
  const int nx = c->g->nx, ny = c->g->ny;
  if (c->px) i = wrap(i, nx);
  if (c->py) j = wrap(j, ny);
  if (i >= 0 && i < nx && j >= 0 && j < ny) {
    const long plane = (long)nx * ny, k = (long)j * nx + i;
    for (int q = 0; q < 4; ++q) u[q] = cells[q * plane + k];
    return 1;
  }
  ecor_direct_boundary_state(c->g, c->d, i, j, t, u);
  return 0;
}

static double pressure_from(const Ctx* c, int i, int j, const double* p, double t)
{
  const int nx = c->g->nx, ny = c->g->ny;
  if (c->px) i = wrap(i, nx);
  if (c->py) j = wrap(j, ny);
  if (i >= 0 && i < nx && j >= 0 && j < ny) return p[(long)j * nx + i];
  return boundary_pressure(c->g, c->d, i, j, t);
}

/* reconstructed_old (:188-203) */
static void recon(const Ctx* c, const double* U, int i, int j, int axis, int side, double t, double* u)
{
  double um[4], uc[4], up[4];
  cell_from(c, axis == 0 ? i - 1 : i, axis == 0 ? j : j - 1, U, t, um);
  cell_from(c, i, j, U, t, uc);
  cell_from(c, axis == 0 ? i + 1 : i, axis == 0 ? j : j + 1, U, t, up);
  const double s = 0.5 * side;
  for (int q = 0; q < 4; ++q) u[q] = uc[q] + s * minmod(uc[q] - um[q], up[q] - uc[q]);
}

/* explicit_flux (:204-233) */
static void lf_flux(const Ctx* c, const double* lu, const double* ru, int axis, double* f)
{
  const double gamma = c->g->gamma, eps = c->d->epsilon;
  double l[4], r[4];
  ecor_eps_to_primitive(lu, gamma, eps, l);
  ecor_eps_to_primitive(ru, gamma, eps, r);
  const double mx_l = l[0] * l[1], my_l = l[0] * l[2];
  const double mx_r = r[0] * r[1], my_r = r[0] * r[2];
  const double k_l = eps * 0.5 * l[0] * (l[1] * l[1] + l[2] * l[2]);
  const double k_r = eps * 0.5 * r[0] * (r[1] * r[1] + r[2] * r[2]);
  const double un_l = axis == 0 ? l[1] : l[2], un_r = axis == 0 ? r[1] : r[2];
  const double fl[4] = {axis == 0 ? mx_l : my_l, mx_l * un_l, my_l * un_l, k_l * un_l};
  const double fr[4] = {axis == 0 ? mx_r : my_r, mx_r * un_r, my_r * un_r, k_r * un_r};
  const double s = smax(fabs(un_l), fabs(un_r));
  for (int q = 0; q < 4; ++q) f[q] = 0.5 * (fl[q] + fr[q]) - 0.5 * s * (ru[q] - lu[q]);
}

/* bdltv20_eq46_face_enthalpy (problems/problem_state_functions.hpp) */
static double face_enthalpy(double rl, double pl, double ql, double rr, double pr, double qr,
                            double gamma, int* guards)
{
  const double hl = gamma * pl / ((gamma - 1.0) * rl);
  const double hr = gamma * pr / ((gamma - 1.0) * rr);
  const double denom = ql + qr;
  const double scale = smax(1.0, fabs(ql) + fabs(qr));
  if (fabs(denom) <= 1.0e-12 * scale) {
    ++*guards;
    return 0.5 * (hl + hr);
  }
  return (hl * ql + hr * qr) / denom;
}

typedef struct {
  long col;
  double val;
} DTerm;

static int cmp_dterm(const void* a, const void* b)
{
  const long x = ((const DTerm*)a)->col, y = ((const DTerm*)b)->col;
  return (x > y) - (x < y);
}

/* refresh_corrected_cells_and_faces (:301-383): corrected momentum (no
 * correction when p is NULL), then h on every face; returns the nonfinite
 * count.  hx: (nx+1)*ny, face f of row j between cells f-1 and f. */
static int refresh(const Ctx* c, const double* star, const double* p_old, const double* p, double dt,
                   double t, double* corr, double* hx, double* hy, int* guards)
{
  const ecor_grid* g = c->g;
  const int nx = g->nx, ny = g->ny;
  const long plane = (long)nx * ny;
  const double eps = c->d->epsilon;
  const double* pp = p ? p : p_old;
  int bad = 0;
  memcpy(corr, star, 4 * plane * sizeof(double));
  if (p) {
    for (int j = 0; j < ny; ++j) {
      for (int i = 0; i < nx; ++i) {
        const long k = (long)j * nx + i;
        const double pe = pressure_from(c, i + 1, j, p, t), pw = pressure_from(c, i - 1, j, p, t);
        const double pn = pressure_from(c, i, j + 1, p, t), ps = pressure_from(c, i, j - 1, p, t);
        corr[plane + k] -= dt * (pe - pw) / (2.0 * eps * g->dx);
        if (!c->d->x_only_row) corr[2 * plane + k] -= dt * (pn - ps) / (2.0 * eps * g->dy);
      }
    }
  }
  for (int j = 0; j < ny; ++j) {
    for (int f = 0; f <= nx; ++f) {
      double l[4], r[4];
      cell_from(c, f - 1, j, corr, t, l);
      cell_from(c, f, j, corr, t, r);
      const double pl = pressure_from(c, f - 1, j, pp, t), pr = pressure_from(c, f, j, pp, t);
      int dup = 0;
      const double h = face_enthalpy(l[0], pl, l[1], r[0], pr, r[1], g->gamma,
                                     (c->px && f == nx) ? &dup : guards);
      const double m = 0.5 * (l[1] + r[1]);
      /* a periodic axis has nx faces: face 0 and face nx are one face */
      if (!(c->px && f == nx) && (!isfinite(h) || !isfinite(m))) ++bad;
      hx[(long)j * (nx + 1) + f] = h;
    }
  }
  for (int f = 0; f <= ny; ++f) {
    for (int i = 0; i < nx; ++i) {
      double b[4], tt[4];
      cell_from(c, i, f - 1, corr, t, b);
      cell_from(c, i, f, corr, t, tt);
      const double pb = pressure_from(c, i, f - 1, pp, t), pt = pressure_from(c, i, f, pp, t);
      int dup = 0;
      const double h = face_enthalpy(b[0], pb, b[2], tt[0], pt, tt[2], g->gamma,
                                     (c->py && f == ny) ? &dup : guards);
      const double m = 0.5 * (b[2] + tt[2]);
      if (!(c->py && f == ny) && (!isfinite(h) || !isfinite(m))) ++bad;
      hy[(long)f * nx + i] = h;
    }
  }
  return bad;
}

/* Pressure rows (:424-494): CSR with duplicate columns summed in triplet
 * order; exact-Dirichlet columns folded into rhs.  Returns the nonfinite
 * count of the right-hand side and star density checks. */
static int rows(const Ctx* c, const double* star, const double* corr, const double* hx,
                const double* hy, double dt, double t, long* rp, long* cols, double* vals, double* rhs)
{
  const ecor_grid* g = c->g;
  const int nx = g->nx, ny = g->ny;
  const long plane = (long)nx * ny;
  const double eps = c->d->epsilon;
  const double eos = eps / (g->gamma - 1.0);
  const double cx = 0.25 * dt * dt / (g->dx * g->dx);
  const double cy = c->d->x_only_row ? 0.0 : 0.25 * dt * dt / (g->dy * g->dy);
  int bad = 0;
  long nnz = 0;
  rp[0] = 0;
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      const long k = (long)j * nx + i;
      const double he = hx[(long)j * (nx + 1) + i + 1], hw = hx[(long)j * (nx + 1) + i];
      const double hn = hy[(long)(j + 1) * nx + i], hs = hy[(long)j * nx + i];
      const double kinetic = eps * 0.5 * (corr[plane + k] * corr[plane + k] +
                                          corr[2 * plane + k] * corr[2 * plane + k]) / star[k];
      double se[4], sw[4], sn[4], ss[4];
      cell_from(c, i + 1, j, star, t, se);
      cell_from(c, i - 1, j, star, t, sw);
      cell_from(c, i, j + 1, star, t, sn);
      cell_from(c, i, j - 1, star, t, ss);
      const double div = (he * se[1] + (he - hw) * star[plane + k] - hw * sw[1]) / (2.0 * g->dx) +
                         (hn * sn[2] + (hn - hs) * star[2 * plane + k] - hs * ss[2]) / (2.0 * g->dy);
      double r = eps * (star[3 * plane + k] - kinetic - dt * div);
      const int ti[9] = {i + 2, i + 1, i, i - 1, i - 2, i, i, i, i};
      const int tj[9] = {j, j, j, j, j, j + 2, j + 1, j - 1, j - 2};
      const double tv[9] = {-cx * he, -cx * (he - hw), eos + cx * (he + hw) + cy * (hn + hs),
                            cx * (he - hw), -cx * hw, -cy * hn, -cy * (hn - hs), cy * (hn - hs),
                            -cy * hs};
      DTerm row[9];
      int nt = 0;
      for (int q = 0; q < 9; ++q) {
        if (tv[q] == 0.0) continue;
        int ci = ti[q], cj = tj[q];
        if (c->px) ci = wrap(ci, nx);
        if (c->py) cj = wrap(cj, ny);
        if (ci >= 0 && ci < nx && cj >= 0 && cj < ny) {
          const long col = (long)cj * nx + ci;
          int found = 0;
          for (int s = 0; s < nt; ++s) {
            if (row[s].col == col) { row[s].val += tv[q]; found = 1; break; }
          }
          if (!found) { row[nt].col = col; row[nt].val = tv[q]; ++nt; }
        } else {
          r -= tv[q] * boundary_pressure(g, c->d, ti[q], tj[q], t);
        }
      }
      qsort(row, (size_t)nt, sizeof(DTerm), cmp_dterm);
      for (int s = 0; s < nt; ++s) { cols[nnz] = row[s].col; vals[nnz] = row[s].val; ++nnz; }
      rp[k + 1] = nnz;
      rhs[k] = r;
      if (!isfinite(r) || !(star[k] > 0.0)) ++bad;
    }
  }
  return bad;
}

void ecor_direct_step(const ecor_grid* g, const ecor_direct* d, double* U, double time, double dt,
                      int picard, double tol, int max_iter, double* p_out, ecor_direct_result* out)
{
  const int nx = g->nx, ny = g->ny;
  const long n = (long)nx * ny, plane = n;
  const double gamma = g->gamma, eps = d->epsilon, gm1 = gamma - 1.0;
  Ctx c = {g, d, d->exact.periodic[0], d->exact.periodic[1]};
  memset(out, 0, sizeof *out);
  double* star = (double*)malloc(4 * n * sizeof(double));
  double* corr = (double*)malloc(4 * n * sizeof(double));
  double* p_old = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  double* x = (double*)malloc(n * sizeof(double));
  double* rhs = (double*)malloc(n * sizeof(double));
  double* hx = (double*)malloc((size_t)(nx + 1) * ny * sizeof(double));
  double* hy = (double*)malloc((size_t)nx * (ny + 1) * sizeof(double));
  long* rp = (long*)malloc((n + 1) * sizeof(long));
  long* cols = (long*)malloc(9 * n * sizeof(long));
  double* vals = (double*)malloc(9 * n * sizeof(double));

  for (long k = 0; k < n; ++k) {
    double u[4] = {U[k], U[plane + k], U[2 * plane + k], U[3 * plane + k]}, q[4];
    ecor_eps_to_primitive(u, gamma, eps, q);
    p_old[k] = q[3];
    p[k] = q[3];
  }
  /* star update (:234-265) */
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      double a[4], b[4], fxl[4], fxr[4], fyb[4], fyt[4];
      recon(&c, U, i - 1, j, 0, +1, time, a); recon(&c, U, i, j, 0, -1, time, b); lf_flux(&c, a, b, 0, fxl);
      recon(&c, U, i, j, 0, +1, time, a); recon(&c, U, i + 1, j, 0, -1, time, b); lf_flux(&c, a, b, 0, fxr);
      recon(&c, U, i, j - 1, 1, +1, time, a); recon(&c, U, i, j, 1, -1, time, b); lf_flux(&c, a, b, 1, fyb);
      recon(&c, U, i, j, 1, +1, time, a); recon(&c, U, i, j + 1, 1, -1, time, b); lf_flux(&c, a, b, 1, fyt);
      const long k = (long)j * nx + i;
      for (int q = 0; q < 4; ++q) {
        star[q * plane + k] = U[q * plane + k] - dt * ((fxr[q] - fxl[q]) / g->dx + (fyt[q] - fyb[q]) / g->dy);
      }
    }
  }
  const double t1 = time + dt;
  for (int it = 0; it < picard; ++it) {
    out->nonfinite_count += refresh(&c, star, p_old, it == 0 ? NULL : p, dt, t1, corr, hx, hy,
                                    &out->guard_count);
    out->nonfinite_count += rows(&c, star, corr, hx, hy, dt, t1, rp, cols, vals, rhs);
    ecor_gmres_result gr;
    ecor_gmres_full(n, rp, cols, vals, rhs, p, max_iter, tol, x, &gr);
    if (!gr.converged || gr.nonfinite || gr.breakdown) ++out->solver_failures;
    if (gr.iterations > out->iterations_max) out->iterations_max = gr.iterations;
    out->iterations += gr.iterations;
    out->residual_linf_max = smax(out->residual_linf_max, gr.residual_linf);
    out->relative_residual_linf_max = smax(out->relative_residual_linf_max, gr.relative_residual_linf);
    out->last_residual_linf = gr.residual_linf;
    memcpy(p, x, n * sizeof(double));
  }
  out->nonfinite_count += refresh(&c, star, p_old, p, dt, t1, corr, hx, hy, &out->guard_count);
  /* closeout (:568-598) */
  for (long k = 0; k < n; ++k) {
    double u[4] = {corr[k], corr[plane + k], corr[2 * plane + k], 0.0}, q[4];
    const double kinetic = eps * 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0];
    u[3] = p[k] / gm1 + kinetic;
    ecor_eps_to_primitive(u, gamma, eps, q);
    if (!isfinite(q[0]) || !isfinite(q[1]) || !isfinite(q[2]) || !isfinite(q[3])) ++out->nonfinite_count;
    if (!(q[0] > 0.0) || !(q[3] > 0.0)) ++out->nonpositive_count;
    for (int s = 0; s < 4; ++s) U[s * plane + k] = u[s];
  }
  if (p_out) memcpy(p_out, p, n * sizeof(double));
  free(star); free(corr); free(p_old); free(p); free(x); free(rhs);
  free(hx); free(hy); free(rp); free(cols); free(vals);
}

/* compute_bdltv20_paper_t1_s2_dt (:8-41) */
double ecor_direct_dt(const ecor_grid* g, const ecor_direct* d, const double* U, double imex_cfl,
                      double fixed_dt, double time, double stop_time)
{
  if (fixed_dt > 0.0) return smin(fixed_dt, stop_time - time);
  const long n = (long)g->nx * g->ny;
  const double min_dx = smin(g->dx, g->dy);
  double rate = 0.0;
  for (long k = 0; k < n; ++k) {
    double u[4] = {U[k], U[n + k], U[2 * n + k], U[3 * n + k]}, q[4];
    ecor_eps_to_primitive(u, g->gamma, d->epsilon, q);
    if (isfinite(q[1]) && isfinite(q[2])) {
      const double speed = sqrt(q[1] * q[1] + q[2] * q[2]);
      rate = smax(rate, 2.0 * speed / min_dx);
    }
  }
  const double remaining = stop_time - time;
  if (!(rate > 0.0)) return remaining;
  return smin(imex_cfl / rate, remaining);
}
