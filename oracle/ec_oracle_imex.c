/*
 * ORACLE -- test infrastructure only (see ec_oracle.h).
 *
 * Serial CPU restatement of the IMEX T1/S2 BDLTV20 source-map Picard step,
 * main-loop form (imex/imex_source_map_step.hpp:300-616), with the
 * reference's assembled pressure rows (imex/imex_source_map_helpers.hpp:
 * 473-571, duplicates summed in triplet order as setFromTriplets does,
 * imex/imex_source_map_step.hpp:8-38) and its unrestarted MGS-GMRES with
 * best-iterate tracking (imex/imex_pressure_helpers.hpp:92-217).  The small
 * least-squares problem is solved with Givens rotations instead of the
 * reference's per-iteration column-pivoting QR (the same minimiser up to
 * rounding); IMEX parity is defined at solver-tolerance level.
 *
 * Pins (tests/test_oracle_imex.py): the assembled operator equals the
 * differences-first matrix-free form and annihilates constants up to the
 * 1/(gamma-1) diagonal; with uniform enthalpy it is symmetric; GMRES on an
 * SPD system reaches the tolerance and agrees with a dense solve; an IMEX
 * step of a uniform flow is exact; the step conserves mass and energy on a
 * periodic grid; the star update reduces to the LF formula.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ec_oracle.h"

static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

/* ---------------------------------------------------------------- star */
/* reconstruction_fluxes.hpp:26-52: conserved minmod reconstruction, then
 * to_primitive of the reconstructed conserved state. */
static double minmod(double a, double b)
{
  if (a * b <= 0.0) return 0.0;
  return fabs(a) < fabs(b) ? a : b;
}

static void cons_minmod_state(const ecor_grid* g, const double* U, int i, int j, int axis,
                              int side, double* q)
{
  double u[4], um[4], up[4];
  ecor_load(g, U, i, j, u);
  if (g->order <= 1) { ecor_to_primitive(u, g->gamma, q); return; }
  if (axis == 0) { ecor_load(g, U, i - 1, j, um); ecor_load(g, U, i + 1, j, up); }
  else { ecor_load(g, U, i, j - 1, um); ecor_load(g, U, i, j + 1, up); }
  double r[4];
  const double s = 0.5 * side;
  for (int c = 0; c < 4; ++c) {
    const double slope = minmod(u[c] - um[c], up[c] - u[c]);
    r[c] = u[c] + s * slope;
  }
  ecor_to_primitive(r, g->gamma, q);
}

/* reconstruction_fluxes.hpp:462-484, 504-515 */
static void lf_flux(const double* L, const double* R, int axis, double gamma, int acoustic,
                    double* f)
{
  double ul[4], ur[4], fl[4], fr[4];
  ecor_to_conserved(L, gamma, ul);
  ecor_to_conserved(R, gamma, ur);
  const double* qs[2] = {L, R};
  double* fs[2] = {fl, fr};
  for (int s = 0; s < 2; ++s) {
    const double* q = qs[s];
    const double mx = q[0] * q[1], my = q[0] * q[2];
    const double kin = 0.5 * q[0] * (q[1] * q[1] + q[2] * q[2]);
    const double un = axis == 0 ? q[1] : q[2];
    fs[s][0] = axis == 0 ? mx : my;
    fs[s][1] = mx * un;
    fs[s][2] = my * un;
    fs[s][3] = kin * un;
  }
  const double unl = axis == 0 ? L[1] : L[2], unr = axis == 0 ? R[1] : R[2];
  double smax_ = smax(fabs(unl), fabs(unr));
  if (acoustic) {
    smax_ = smax(fabs(unl) + sqrt(gamma * L[3] / L[0]), fabs(unr) + sqrt(gamma * R[3] / R[0]));
  }
  for (int c = 0; c < 4; ++c) f[c] = 0.5 * (fl[c] + fr[c]) - (0.5 * smax_) * (ur[c] - ul[c]);
}

/* imex_source_map_helpers.hpp:2-68 */
void ecor_imex_star(const ecor_grid* g, const double* U, double dt, int acoustic, double* S)
{
  const int nx = g->nx, ny = g->ny;
  const long plane = (long)nx * ny;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      double l[4], r[4], fxl[4], fxr[4], fyb[4], fyt[4];
      cons_minmod_state(g, U, i - 1, j, 0, +1, l);
      cons_minmod_state(g, U, i, j, 0, -1, r);
      lf_flux(l, r, 0, g->gamma, acoustic, fxl);
      cons_minmod_state(g, U, i, j, 0, +1, l);
      cons_minmod_state(g, U, i + 1, j, 0, -1, r);
      lf_flux(l, r, 0, g->gamma, acoustic, fxr);
      cons_minmod_state(g, U, i, j - 1, 1, +1, l);
      cons_minmod_state(g, U, i, j, 1, -1, r);
      lf_flux(l, r, 1, g->gamma, acoustic, fyb);
      cons_minmod_state(g, U, i, j, 1, +1, l);
      cons_minmod_state(g, U, i, j + 1, 1, -1, r);
      lf_flux(l, r, 1, g->gamma, acoustic, fyt);
      const long k = (long)j * nx + i;
      for (int c = 0; c < 4; ++c) {
        S[c * plane + k] =
            U[c * plane + k] - dt * ((fxr[c] - fxl[c]) * inv_dx + (fyt[c] - fyb[c]) * inv_dy);
      }
    }
  }
}

/* --------------------------------------------------- boundary mappings */
typedef struct {
  const ecor_grid* g;
  const double* star;    /* 4 planes, valid cells */
  const double* mom;     /* 2 planes: lagged momentum */
  const double* p;       /* pressure */
  double fixed_p[4];     /* Dirichlet pressure of a FIXED side */
} SmCtx;

static int wrap(int i, int n) { return ((i % n) + n) % n; }
static int clampi(int i, int n) { return i < 0 ? 0 : (i > n - 1 ? n - 1 : i); }

/* CoupledHostGrid::map_i/map_j (imex_source_map_helpers.hpp:79-93) */
static int map_state(int i, int n, int periodic) { return periodic ? wrap(i, n) : clampi(i, n); }

/* source_map_index (imex_source_map_helpers.hpp:169-188): even reflection */
static int map_pressure(int i, int n, int periodic)
{
  if (periodic) return wrap(i, n);
  if (i < 0) { const int m = -i - 1; return m < n - 1 ? m : n - 1; }
  if (i >= n) { const int m = n - (i - n + 1); return m > 0 ? m : 0; }
  return i;
}

static int fixed_side(const ecor_grid* g, int i, int j)
{
  if (i < 0 && g->bc[0] == ECOR_BC_FIXED) return 0;
  if (i >= g->nx && g->bc[1] == ECOR_BC_FIXED) return 1;
  if (j < 0 && g->bc[2] == ECOR_BC_FIXED) return 2;
  if (j >= g->ny && g->bc[3] == ECOR_BC_FIXED) return 3;
  return -1;
}

/* Row/column of a cell read through the state rule; *neg_my set when a
 * reflecting side flips the normal momentum (imex_source_map_helpers.hpp:
 * 98-115, 300-320: the mirror reads row 0 for every ghost layer). */
static void state_cell(const ecor_grid* g, int i, int j, int* ii, int* jj, int* neg_mx, int* neg_my)
{
  const int px = g->bc[0] == ECOR_BC_PERIODIC, py = g->bc[2] == ECOR_BC_PERIODIC;
  *neg_mx = *neg_my = 0;
  *ii = map_state(i, g->nx, px);
  *jj = map_state(j, g->ny, py);
  if (j < 0 && g->bc[2] == ECOR_BC_REFLECT) { *jj = 0; *neg_my = 1; }
  if (j >= g->ny && g->bc[3] == ECOR_BC_REFLECT) { *jj = g->ny - 1; *neg_my = 1; }
  if (i < 0 && g->bc[0] == ECOR_BC_REFLECT) { *ii = 0; *neg_mx = 1; }
  if (i >= g->nx && g->bc[1] == ECOR_BC_REFLECT) { *ii = g->nx - 1; *neg_mx = 1; }
}

static void star_at(const SmCtx* c, int i, int j, double* u)
{
  const ecor_grid* g = c->g;
  const int fs = fixed_side(g, i, j);
  if (fs >= 0) { memcpy(u, g->fixed[fs], 4 * sizeof(double)); return; }
  int ii, jj, nmx, nmy;
  state_cell(g, i, j, &ii, &jj, &nmx, &nmy);
  const long plane = (long)g->nx * g->ny, k = (long)jj * g->nx + ii;
  for (int q = 0; q < 4; ++q) u[q] = c->star[q * plane + k];
  if (nmx) u[1] = -u[1];
  if (nmy) u[2] = -u[2];
}

static void mom_at(const SmCtx* c, int i, int j, double* m)
{
  const ecor_grid* g = c->g;
  const int fs = fixed_side(g, i, j);
  if (fs >= 0) { m[0] = g->fixed[fs][1]; m[1] = g->fixed[fs][2]; return; }
  int ii, jj, nmx, nmy;
  state_cell(g, i, j, &ii, &jj, &nmx, &nmy);
  const long plane = (long)g->nx * g->ny, k = (long)jj * g->nx + ii;
  m[0] = c->mom[k];
  m[1] = c->mom[plane + k];
  if (nmx) m[0] = -m[0];
  if (nmy) m[1] = -m[1];
}

static long pressure_col(const ecor_grid* g, int i, int j)
{
  const int ii = map_pressure(i, g->nx, g->bc[0] == ECOR_BC_PERIODIC);
  const int jj = map_pressure(j, g->ny, g->bc[2] == ECOR_BC_PERIODIC);
  return (long)jj * g->nx + ii;
}

/* imex_source_map_helpers.hpp:215-245 */
static double p_at(const SmCtx* c, int i, int j)
{
  const int fs = fixed_side(c->g, i, j);
  if (fs >= 0) return c->fixed_p[fs];
  return c->p[pressure_col(c->g, i, j)];
}

/* imex_source_map_helpers.hpp:359-402 */
static void face_detail(const SmCtx* c, int il, int jl, int ir, int jr, int axis, double* h_hat,
                        double* flux, int* guard)
{
  const double gamma = c->g->gamma, gm1 = gamma - 1.0;
  double L[4], R[4], ml[2], mr[2];
  star_at(c, il, jl, L);
  star_at(c, ir, jr, R);
  mom_at(c, il, jl, ml);
  mom_at(c, ir, jr, mr);
  const double pl = p_at(c, il, jl), pr = p_at(c, ir, jr);
  const double hl = gamma * pl / (gm1 * L[0]);
  const double hr = gamma * pr / (gm1 * R[0]);
  const double ql = axis == 0 ? ml[0] : ml[1];
  const double qr = axis == 0 ? mr[0] : mr[1];
  const double denom = ql + qr;
  const double scale = smax(smax(fabs(ql), fabs(qr)), 1.0);
  const int gd = fabs(denom) <= 1.0e-14 * scale;
  const double hh = gd ? 0.5 * (hl + hr) : (hl * ql + hr * qr) / denom;
  if (h_hat) *h_hat = hh;
  if (flux) *flux = hh * (0.5 * (ql + qr));
  if (guard) *guard = gd;
}

/* ------------------------------------------------------------ assembly */
typedef struct {
  long col;
  double val;
} Term;

static int cmp_term(const void* a, const void* b)
{
  const long x = ((const Term*)a)->col, y = ((const Term*)b)->col;
  return (x > y) - (x < y);
}

/* CSR of the pressure rows; row_ptr has n+1 entries, cols/vals 11*n. */
static int assemble(const SmCtx* c, double dt, long* row_ptr, long* cols, double* vals, double* rhs)
{
  const ecor_grid* g = c->g;
  const int nx = g->nx, ny = g->ny;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
  const double gm1 = g->gamma - 1.0;
  const double cx = 0.25 * dt * dt * inv_dx * inv_dx;
  const double cy = 0.25 * dt * dt * inv_dy * inv_dy;
  const long plane = (long)nx * ny;
  int nonfinite = 0;
  long nnz = 0;
  row_ptr[0] = 0;
  for (int j = 0; j < ny; ++j) {
    for (int i = 0; i < nx; ++i) {
      double he, hw, hn, hs;
      face_detail(c, i, j, i + 1, j, 0, &he, NULL, NULL);
      face_detail(c, i - 1, j, i, j, 0, &hw, NULL, NULL);
      face_detail(c, i, j, i, j + 1, 1, &hn, NULL, NULL);
      face_detail(c, i, j - 1, i, j, 1, &hs, NULL, NULL);
      if (!isfinite(he) || !isfinite(hw) || !isfinite(hn) || !isfinite(hs)) ++nonfinite;
      const int ti[11] = {i, i + 2, i + 1, i, i - 1, i - 2, i, i, i, i, i};
      const int tj[11] = {j, j, j, j, j, j, j + 2, j + 1, j, j - 1, j - 2};
      const double tv[11] = {1.0 / gm1, -cx * he, -cx * (he - hw), cx * (he + hw), cx * (he - hw),
                             -cx * hw, -cy * hn, -cy * (hn - hs), cy * (hn + hs), cy * (hn - hs),
                             -cy * hs};
      double r = 0.0;
      Term row[11];
      int nt = 0;
      for (int t = 0; t < 11; ++t) {
        const int fs = fixed_side(g, ti[t], tj[t]);
        if (fs >= 0) {  /* Dirichlet pressure moves to the right-hand side */
          const double contrib = -tv[t] * c->fixed_p[fs];
          r += contrib;
          if (!isfinite(contrib)) ++nonfinite;
          continue;
        }
        if (!isfinite(tv[t])) ++nonfinite;
        if (tv[t] == 0.0) continue;
        const long col = pressure_col(g, ti[t], tj[t]);
        int found = 0;
        for (int q = 0; q < nt; ++q) {
          if (row[q].col == col) { row[q].val += tv[t]; found = 1; break; }
        }
        if (!found) { row[nt].col = col; row[nt].val = tv[t]; ++nt; }
      }
      qsort(row, (size_t)nt, sizeof(Term), cmp_term);
      for (int q = 0; q < nt; ++q) { cols[nnz] = row[q].col; vals[nnz] = row[q].val; ++nnz; }
      const long k = (long)j * nx + i;
      row_ptr[k + 1] = nnz;
      /* right-hand side (imex_source_map_helpers.hpp:545-566) */
      double m[2], sp[4], sm[4];
      mom_at(c, i, j, m);
      const double rho = c->star[k], E = c->star[3 * plane + k];
      const double kinetic = 0.5 * (m[0] * m[0] + m[1] * m[1]) / rho;
      const double mx_c = c->star[plane + k], my_c = c->star[2 * plane + k];
      double terms[8];
      terms[0] = E;
      terms[1] = -kinetic;
      star_at(c, i + 1, j, sp);
      star_at(c, i - 1, j, sm);
      terms[2] = -0.5 * dt * inv_dx * he * sp[1];
      terms[3] = -0.5 * dt * inv_dx * (he - hw) * mx_c;
      terms[4] = 0.5 * dt * inv_dx * hw * sm[1];
      star_at(c, i, j + 1, sp);
      star_at(c, i, j - 1, sm);
      terms[5] = -0.5 * dt * inv_dy * hn * sp[2];
      terms[6] = -0.5 * dt * inv_dy * (hn - hs) * my_c;
      terms[7] = 0.5 * dt * inv_dy * hs * sm[2];
      /* Row terms with a fixed boundary pressure were added first (the
       * reference adds them while building the left-hand side). */
      for (int t = 0; t < 8; ++t) {
        if (!isfinite(terms[t])) ++nonfinite;
        r += terms[t];
      }
      if (!isfinite(r)) ++nonfinite;
      rhs[k] = r;
    }
  }
  return nonfinite;
}

static void spmv(long n, const long* rp, const long* cols, const double* vals, const double* x,
                 double* y)
{
  for (long r = 0; r < n; ++r) {
    double t = 0.0;
    for (long k = rp[r]; k < rp[r + 1]; ++k) t += vals[k] * x[cols[k]];
    y[r] = t;
  }
}

static double norm_inf_diff(long n, const double* a, const double* b)
{
  double m = 0.0;
  for (long i = 0; i < n; ++i) {
    const double d = fabs(a[i] - b[i]);
    if (m < d || isnan(d)) m = d;
  }
  return m;
}

static int all_finite(long n, const double* x)
{
  for (long i = 0; i < n; ++i) if (!isfinite(x[i])) return 0;
  return 1;
}

/* ---------------------------------------------------------------- GMRES */
/* imex_pressure_helpers.hpp:92-217.  result fields: iterations of the
 * returned iterate, converged/breakdown/nonfinite flags, residuals. */
void ecor_gmres_full(long n, const long* rp, const long* cols, const double* vals,
                     const double* b, const double* x0, int max_iterations, double tol,
                     double* x, ecor_gmres_result* res)
{
  memset(res, 0, sizeof *res);
  res->residual_linf = INFINITY;
  res->relative_residual_linf = INFINITY;
  memcpy(x, x0, (size_t)n * sizeof(double));
  if (!all_finite(n, b) || !all_finite(n, x0)) { res->nonfinite = 1; return; }
  double bnorm = 0.0;
  for (long i = 0; i < n; ++i) bnorm = smax(bnorm, fabs(b[i]));
  const double rhs_norm_inf = smax(bnorm, 1.0);
  double* ax = (double*)malloc((size_t)n * sizeof(double));
  double* trial = (double*)malloc((size_t)n * sizeof(double));

#define UPDATE_BEST(vec, its)                                                  \
  do {                                                                         \
    if (!all_finite(n, vec)) { res->nonfinite = 1; break; }                    \
    spmv(n, rp, cols, vals, vec, ax);                                          \
    const double rl = norm_inf_diff(n, ax, b);                                 \
    if (!isfinite(rl)) { res->nonfinite = 1; break; }                          \
    const double rel = rl / rhs_norm_inf;                                      \
    if (rl < res->residual_linf) {                                             \
      memcpy(x, vec, (size_t)n * sizeof(double));                              \
      res->iterations = (its); res->residual_linf = rl;                        \
      res->relative_residual_linf = rel;                                       \
    }                                                                          \
    if (rl <= tol || rel <= tol) {                                             \
      memcpy(x, vec, (size_t)n * sizeof(double));                              \
      res->iterations = (its); res->residual_linf = rl;                        \
      res->relative_residual_linf = rel; res->converged = 1;                   \
    }                                                                          \
  } while (0)

  UPDATE_BEST(x0, 0);
  if (res->converged || res->nonfinite) { free(ax); free(trial); return; }
  const long lim_l = max_iterations < n ? max_iterations : n;
  const int lim = lim_l > 0 ? (int)lim_l : 0;
  if (lim == 0) { free(ax); free(trial); return; }
  double* r0 = (double*)malloc((size_t)n * sizeof(double));
  spmv(n, rp, cols, vals, x0, ax);
  double beta2 = 0.0;
  for (long i = 0; i < n; ++i) { r0[i] = b[i] - ax[i]; beta2 += r0[i] * r0[i]; }
  const double beta = sqrt(beta2);
  if (!all_finite(n, r0) || !isfinite(beta)) {
    res->nonfinite = 1; free(ax); free(trial); free(r0); return;
  }
  if (beta == 0.0) {
    res->converged = 1; res->residual_linf = 0.0; res->relative_residual_linf = 0.0;
    free(ax); free(trial); free(r0); return;
  }
  double* V = (double*)malloc((size_t)(lim + 1) * (size_t)n * sizeof(double));
  double* H = (double*)calloc((size_t)(lim + 1) * (size_t)lim, sizeof(double)); /* H[i*lim+j] */
  double* R = (double*)calloc((size_t)(lim + 1) * (size_t)lim, sizeof(double));
  double* cs = (double*)malloc((size_t)lim * sizeof(double));
  double* sn = (double*)malloc((size_t)lim * sizeof(double));
  double* gvec = (double*)calloc((size_t)(lim + 1), sizeof(double));
  double* y = (double*)malloc((size_t)lim * sizeof(double));
  double* w = (double*)malloc((size_t)n * sizeof(double));
  for (long i = 0; i < n; ++i) V[i] = r0[i] / beta;
  int nbasis = 1;
  const double breakdown_tol = 2.220446049250313e-16 * 100.0 * smax(beta, 1.0);
  gvec[0] = beta;
  for (int j = 0; j < lim; ++j) {
    spmv(n, rp, cols, vals, V + (long)j * n, w);
    if (!all_finite(n, w)) { res->nonfinite = 1; break; }
    for (int i = 0; i <= j; ++i) {
      const double* v = V + (long)i * n;
      double h = 0.0;
      for (long k = 0; k < n; ++k) h += v[k] * w[k];
      H[i * lim + j] = h;
      for (long k = 0; k < n; ++k) w[k] -= h * v[k];
    }
    double nn = 0.0;
    for (long k = 0; k < n; ++k) nn += w[k] * w[k];
    const double next_norm = sqrt(nn);
    if (!isfinite(next_norm)) { res->nonfinite = 1; break; }
    H[(j + 1) * lim + j] = next_norm;
    if (next_norm > breakdown_tol) {
      for (long k = 0; k < n; ++k) V[(long)nbasis * n + k] = w[k] / next_norm;
      ++nbasis;
    }
    /* least squares min ||H(0:j+2,0:j+1) y - beta e1|| by Givens */
    for (int i = 0; i <= j + 1; ++i) R[i * lim + j] = H[i * lim + j];
    for (int i = 0; i < j; ++i) {
      const double a = R[i * lim + j], bb = R[(i + 1) * lim + j];
      R[i * lim + j] = cs[i] * a + sn[i] * bb;
      R[(i + 1) * lim + j] = -sn[i] * a + cs[i] * bb;
    }
    {
      const double a = R[j * lim + j], bb = R[(j + 1) * lim + j];
      const double rr = hypot(a, bb);
      cs[j] = rr == 0.0 ? 1.0 : a / rr;
      sn[j] = rr == 0.0 ? 0.0 : bb / rr;
      R[j * lim + j] = rr;
      R[(j + 1) * lim + j] = 0.0;
      gvec[j + 1] = -sn[j] * gvec[j];
      gvec[j] = cs[j] * gvec[j];
    }
    int ok = 1;
    for (int i = j; i >= 0; --i) {
      double s = gvec[i];
      for (int k = i + 1; k <= j; ++k) s -= R[i * lim + k] * y[k];
      y[i] = s / R[i * lim + i];
      if (!isfinite(y[i])) ok = 0;
    }
    if (!ok) { res->nonfinite = 1; break; }
    memcpy(trial, x0, (size_t)n * sizeof(double));
    for (int i = 0; i <= j && i < nbasis; ++i) {
      const double* v = V + (long)i * n;
      for (long k = 0; k < n; ++k) trial[k] += y[i] * v[k];
    }
    UPDATE_BEST(trial, j + 1);
    if (res->converged || res->nonfinite) break;
    if (next_norm <= breakdown_tol) { res->breakdown = 1; break; }
  }
#undef UPDATE_BEST
  free(V); free(H); free(R); free(cs); free(sn); free(gvec); free(y); free(w);
  free(ax); free(trial); free(r0);
}

/* -------------------------------------------------------------- step */
static void corrected_momentum(const SmCtx* c, double dt, double* out)
{
  const ecor_grid* g = c->g;
  const long plane = (long)g->nx * g->ny;
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
  for (int j = 0; j < g->ny; ++j) {
    for (int i = 0; i < g->nx; ++i) {
      const long k = (long)j * g->nx + i;
      const double gx = 0.5 * (p_at(c, i + 1, j) - p_at(c, i - 1, j)) * inv_dx;
      const double gy = 0.5 * (p_at(c, i, j + 1) - p_at(c, i, j - 1)) * inv_dy;
      out[k] = c->star[plane + k] - dt * gx;
      out[plane + k] = c->star[2 * plane + k] - dt * gy;
    }
  }
}

static void fixed_pressures(const ecor_grid* g, double* fp)
{
  for (int s = 0; s < 4; ++s) {
    double q[4];
    ecor_to_primitive(g->fixed[s], g->gamma, q);
    fp[s] = g->bc[s] == ECOR_BC_FIXED ? q[3] : 0.0;
  }
}


/* The closed state's eq78 velocity divergence (imex/
 * imex_source_map_step.hpp:207-244): central differences, neighbours wrapped
 * on a periodic axis and mirrored at a physical side (source_map_index,
 * imex/imex_source_map_helpers.hpp:169-188), a velocity NaN where the density
 * is zero or a value non-finite; infinite when no cell has a finite value. */
static int mirror_index(int i, int n, int periodic)
{
  if (periodic) return ((i % n) + n) % n;
  if (i < 0) { const int a = -i - 1; return a < n - 1 ? a : n - 1; }
  if (i >= n) { const int a = n - (i - n + 1); return a > 0 ? a : 0; }
  return i;
}

static double closed_velocity(const ecor_grid* g, const double* U, int i, int j, int axis)
{
  const long plane = (long)g->nx * g->ny;
  const int ii = mirror_index(i, g->nx, g->bc[0] == ECOR_BC_PERIODIC);
  const int jj = mirror_index(j, g->ny, g->bc[2] == ECOR_BC_PERIODIC);
  const long k = (long)jj * g->nx + ii;
  const double rho = U[k], m = U[(axis == 0 ? 1 : 2) * plane + k];
  if (!isfinite(rho) || rho == 0.0 || !isfinite(m)) return NAN;
  return m / rho;
}

static double eq78_divergence(const ecor_grid* g, const double* U)
{
  const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
  double linf = 0.0;
  long count = 0;
  for (int j = 0; j < g->ny; ++j) {
    for (int i = 0; i < g->nx; ++i) {
      const double ue = closed_velocity(g, U, i + 1, j, 0), uw = closed_velocity(g, U, i - 1, j, 0);
      const double vn = closed_velocity(g, U, i, j + 1, 1), vs = closed_velocity(g, U, i, j - 1, 1);
      const double div = 0.5 * (ue - uw) * inv_dx + 0.5 * (vn - vs) * inv_dy;
      double scale = 1.0;
      const double a[4] = {fabs(ue), fabs(uw), fabs(vn), fabs(vs)};
      for (int q = 0; q < 4; ++q) scale = scale < a[q] ? a[q] : scale;
      const double den = scale * (inv_dx < inv_dy ? inv_dy : inv_dx);
      if (!isfinite(div) || !isfinite(den) || den <= 0.0) continue;
      linf = smax(linf, fabs(div));
      ++count;
    }
  }
  return count > 0 ? linf : INFINITY;
}

/* imex_source_map_step.hpp:300-616 */
void ecor_imex_step(const ecor_grid* g, double* U, double dt, int picard, double tol, int max_iter,
                    double* p_out, ecor_imex_result* out)
{
  const int nx = g->nx, ny = g->ny;
  const long n = (long)nx * ny, plane = n;
  memset(out, 0, sizeof *out);
  double* S = (double*)malloc((size_t)(4 * n) * sizeof(double));
  double* p = (double*)malloc((size_t)n * sizeof(double));
  double* mom = (double*)malloc((size_t)(2 * n) * sizeof(double));
  long* rp = (long*)malloc((size_t)(n + 1) * sizeof(long));
  long* cols = (long*)malloc((size_t)(11 * n) * sizeof(long));
  double* vals = (double*)malloc((size_t)(11 * n) * sizeof(double));
  double* rhs = (double*)malloc((size_t)n * sizeof(double));
  double* sol = (double*)malloc((size_t)n * sizeof(double));
  double* ax = (double*)malloc((size_t)n * sizeof(double));
  SmCtx c;
  c.g = g;
  c.star = S;
  c.mom = mom;
  c.p = p;
  fixed_pressures(g, c.fixed_p);

  /* core/imex_diagnostics.hpp:111-163 defaults */
  out->trial_density_min = out->trial_pressure_min = out->trial_internal_energy_min = NAN;
  out->solved_pressure_min = out->solved_pressure_max = NAN;
  out->eq78_divergence_linf = NAN;
  ecor_imex_star(g, U, dt, 0, S);
  /* imex/imex_source_map_step.hpp:436-444 */
  out->trial_density_min = INFINITY;
  for (long k = 0; k < n; ++k)
    if (isfinite(S[k]) && S[k] < out->trial_density_min) out->trial_density_min = S[k];
  for (long k = 0; k < n; ++k) {
    mom[k] = S[plane + k];
    mom[plane + k] = S[2 * plane + k];
    double u[4] = {U[k], U[plane + k], U[2 * plane + k], U[3 * plane + k]}, q[4];
    ecor_to_primitive(u, g->gamma, q);
    p[k] = q[3];
  }
  for (long k = 0; k < n; ++k) {
    if (!isfinite(S[k]) || S[k] <= 1.0e-14) ++out->density_floor_count;
  }
  if (out->density_floor_count > 0) { out->status = ECOR_IMEX_DENSITY_FLOOR; goto done; }

  for (int it = 0; it < picard; ++it) {
    const int nonfinite = assemble(&c, dt, rp, cols, vals, rhs);
    if (nonfinite > 0 || !all_finite(n, rhs)) { out->status = ECOR_IMEX_ROW_NONFINITE; goto done; }
    double bn = 0.0;
    for (long k = 0; k < n; ++k) bn = smax(bn, fabs(rhs[k]));
    out->rhs_norm = smax(bn, 1.0);
    ecor_gmres_result gr;
    ecor_gmres_full(n, rp, cols, vals, rhs, p, max_iter, tol, sol, &gr);
    out->iterations += gr.iterations;
    out->last_iterations = gr.iterations;
    spmv(n, rp, cols, vals, sol, ax);
    out->residual_linf = norm_inf_diff(n, ax, rhs);
    out->relative_residual_linf = out->residual_linf / out->rhs_norm;
    const int abs_ok = out->residual_linf <= tol;
    const int rel_ok = out->relative_residual_linf <= tol;
    if (!isfinite(out->residual_linf) || !all_finite(n, sol)) {
      out->status = ECOR_IMEX_RESIDUAL_NONFINITE; goto done;
    }
    if (!abs_ok && !rel_ok) {
      out->status = gr.breakdown ? ECOR_IMEX_BREAKDOWN : ECOR_IMEX_NOT_CONVERGED;
      goto done;
    }
    out->status = abs_ok ? ECOR_IMEX_CONVERGED_ABSOLUTE : ECOR_IMEX_CONVERGED_RELATIVE;
    memcpy(p, sol, (size_t)n * sizeof(double));
    corrected_momentum(&c, dt, mom);
    if (!all_finite(2 * n, mom)) { out->status = ECOR_IMEX_MOMENTUM_NONFINITE; goto done; }
    ++out->picard_solves;
  }
  /* closeout (imex_source_map_step.hpp:75-258): E from the conservative
   * enthalpy flux with the corrected momentum of the final pressure. */
  {
    const double inv_dx = 1.0 / g->dx, inv_dy = 1.0 / g->dy;
    const double gm1 = g->gamma - 1.0;
    corrected_momentum(&c, dt, mom);
    double* E = (double*)malloc((size_t)n * sizeof(double));
    double res_linf = 0.0;
    double rho_min = INFINITY, p_min = INFINITY, p_max = -INFINITY, eos_min = INFINITY,
           int_min = INFINITY, mom_corr = 0.0, e_delta = 0.0;
    for (int j = 0; j < ny; ++j) {
      for (int i = 0; i < nx; ++i) {
        double fe, fw, fn, fs;
        int gd;
        face_detail(&c, i, j, i + 1, j, 0, NULL, &fe, &gd); out->guard_count += gd;
        face_detail(&c, i - 1, j, i, j, 0, NULL, &fw, &gd); out->guard_count += gd;
        face_detail(&c, i, j, i, j + 1, 1, NULL, &fn, &gd); out->guard_count += gd;
        face_detail(&c, i, j - 1, i, j, 1, NULL, &fs, &gd); out->guard_count += gd;
        const long k = (long)j * nx + i;
        const double div = (fe - fw) * inv_dx + (fn - fs) * inv_dy;
        E[k] = S[3 * plane + k] - dt * div;
        const double mx = mom[k], my = mom[plane + k], rho = S[k];
        const double kin = 0.5 * (mx * mx + my * my) / rho;
        const double internal = E[k] - kin;
        const double eos = gm1 * internal;
        const double r = p[k] / gm1 - E[k] + kin;
        if (!isfinite(rho) || !isfinite(mx) || !isfinite(my) || !isfinite(E[k]) || !isfinite(p[k]) ||
            !isfinite(kin) || !isfinite(internal) || !isfinite(eos) || !isfinite(r) || !isfinite(div)) {
          ++out->nonfinite_count;
          continue;
        }
        rho_min = rho < rho_min ? rho : rho_min;
        p_min = p[k] < p_min ? p[k] : p_min;
        p_max = p_max < p[k] ? p[k] : p_max;
        eos_min = eos < eos_min ? eos : eos_min;
        int_min = internal < int_min ? internal : int_min;
        const double dmx = mx - S[plane + k], dmy = my - S[2 * plane + k];
        mom_corr = smax(mom_corr, sqrt(dmx * dmx + dmy * dmy));
        e_delta = smax(e_delta, fabs(E[k] - S[3 * plane + k]));
        res_linf = smax(res_linf, fabs(r));
      }
    }
    /* imex/imex_source_map_step.hpp:229-257 */
    out->coupled_residual_linf = out->nonfinite_count > 0 ? INFINITY : res_linf;
    out->trial_density_min = isfinite(rho_min) ? rho_min : NAN;
    out->solved_pressure_min = isfinite(p_min) ? p_min : NAN;
    out->solved_pressure_max = isfinite(p_max) ? p_max : NAN;
    out->trial_pressure_min = isfinite(eos_min) ? eos_min : NAN;
    out->trial_internal_energy_min = isfinite(int_min) ? int_min : NAN;
    out->momentum_correction_linf = mom_corr;
    out->energy_correction_linf = e_delta;
    for (long k = 0; k < n; ++k) {
      U[k] = S[k];
      U[plane + k] = mom[k];
      U[2 * plane + k] = mom[plane + k];
      U[3 * plane + k] = E[k];
    }
    free(E);
    out->eq78_divergence_linf = eq78_divergence(g, U);
    if (out->nonfinite_count > 0) out->eq78_divergence_linf = INFINITY;
  }
done:
  if (p_out) memcpy(p_out, p, (size_t)n * sizeof(double));
  free(S); free(p); free(mom); free(rp); free(cols); free(vals); free(rhs); free(sol); free(ax);
}

/* Assembled system for the given star state, pressure and lagged momentum
 * (for tests): returns nnz, fills CSR arrays sized for 11*n entries. */
long ecor_imex_assemble(const ecor_grid* g, const double* star, const double* p,
                        const double* mom, double dt, long* row_ptr, long* cols, double* vals,
                        double* rhs)
{
  SmCtx c;
  c.g = g;
  c.star = star;
  c.mom = mom;
  c.p = p;
  fixed_pressures(g, c.fixed_p);
  assemble(&c, dt, row_ptr, cols, vals, rhs);
  return row_ptr[(long)g->nx * g->ny];
}

/* diagnostic_metrics_timestep.hpp:262-326 with the eq79 material rate. */
double ecor_imex_dt(const ecor_grid* g, const double* U, double cfl, double time, double stop,
                    int acoustic_startup, double acoustic_cap, double* material_max,
                    double* acoustic_max)
{
  const long n = (long)g->nx * g->ny;
  const double min_dx = smin(g->dx, g->dy);
  double mat = -INFINITY, aco = -INFINITY;
  for (long k = 0; k < n; ++k) {
    const double u[4] = {U[k], U[n + k], U[2 * n + k], U[3 * n + k]};
    double q[4];
    ecor_to_primitive(u, g->gamma, q);
    const double c = sqrt(g->gamma * q[3] / q[0]);
    const double speed = sqrt(q[1] * q[1] + q[2] * q[2]);
    const double m = 2.0 * speed / min_dx;
    const double a = (fabs(q[1]) + c) / g->dx + (fabs(q[2]) + c) / g->dy;
    mat = smax(mat, m);
    aco = smax(aco, a);
  }
  double rate = mat;
  if (acoustic_startup != 0 && mat <= 1.0e-12 && aco > 0.0) rate = aco;
  double dt = stop - time;
  if (rate > 0.0) dt = cfl / rate;
  if (acoustic_cap > 0.0 && aco > 0.0) {
    const double lim = acoustic_cap / aco;
    if (lim < dt) dt = lim;
  }
  if (!isfinite(dt) || dt <= 0.0) dt = stop - time;
  if (material_max) *material_max = mat;
  if (acoustic_max) *acoustic_max = aco;
  return smin(dt, stop - time);
}
