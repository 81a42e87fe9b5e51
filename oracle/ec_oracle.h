/*
 * ORACLE -- test infrastructure only.
 *
 * A serial CPU restatement of the euler_compare update path, written in plain
 * C from the reference's formulas and operation order.  It exists to check the
 * HIP product path (tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg are its only callers).  Nothing in the product links it.
 *
 * Layout: every field is compact SoA, comp-major: U[c*ny*nx + j*nx + i],
 * c in {0:rho, 1:mx, 2:my, 3:E} (amrex/apps/euler_compare/src/
 * euler_compare_app.cpp:47).  Ghost cells are never stored: they are produced
 * on read by the boundary rule of each side (see ecor_load).
 */
#ifndef EC_ORACLE_H
#define EC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ECOR_RUSANOV = 0, ECOR_HLLC = 1, ECOR_LOWMACH_HLLC = 2, ECOR_XIE_HLLCP = 3 };
enum { ECOR_MINMOD = 0, ECOR_MC = 1, ECOR_VANLEER = 2 };
enum { ECOR_BC_PERIODIC = 0, ECOR_BC_OUTFLOW = 1, ECOR_BC_REFLECT = 2, ECOR_BC_FIXED = 3 };

/* Exact-Dirichlet ghost states, euler.field_boundary=exact_dirichlet
 * (problems/boundary_conditions.hpp:133-219; problems/
 * problem_initialization.hpp:209-263, 308-375). */
enum { ECOR_EXACT_NONE = 0, ECOR_EXACT_GRESHO = 1, ECOR_EXACT_TORO = 2,
       ECOR_EXACT_QUADRANT = 3, ECOR_EXACT_BLOB = 4 };

typedef struct {
  int kind;                      /* ECOR_EXACT_* */
  int periodic[2];               /* geometry.is_periodic */
  double prob_lo[2], prob_hi[2]; /* cell size (hi - lo) / n */
  double gamma;
  double vortex_center[2], density_outer, pressure, mach; /* Gresho; blob outer/pressure */
  double interface_x, interface_y;                        /* quadrant; Toro x0 */
  double ql[4], qr[4];                                    /* Toro primitive states */
  double blob_center[2], blob_radius, density_inner, velocity[2];
} ecor_exact;

typedef struct {
  int nx, ny;
  double dx, dy;
  double gamma;
  int riemann;   /* ECOR_RUSANOV .. ECOR_XIE_HLLCP */
  int limiter;   /* ECOR_MINMOD .. ECOR_VANLEER */
  int order;     /* 1 or 2 */
  int bc[4];     /* xlo, xhi, ylo, yhi */
  double fixed[4][4]; /* conserved state used by an ECOR_BC_FIXED side */
  const ecor_exact* exact; /* NULL, or the exact ghost states (priority over bc) */
  double time;             /* time of the exact ghost states */
} ecor_grid;

/* ---- pointwise algebra (state_primitives.hpp / reconstruction_fluxes.hpp) */
void ecor_to_primitive(const double* u, double gamma, double* q);
void ecor_to_conserved(const double* q, double gamma, double* u);
double ecor_limited_slope(double backward, double forward, int limiter);
void ecor_physical_flux(const double* q, int axis, double gamma, double* f);
void ecor_interface_flux(const double* qL, const double* qR, int axis, double gamma,
                         int riemann, double sensor_f, double* f);
int ecor_hllc_star_degenerate(const double* qL, const double* qR, int axis, double gamma);
double ecor_pressure_ratio(double pL, double pR);
void ecor_exact_riemann_star(const double* qL, const double* qR, double gamma,
                             double* p_star, double* u_star);
void ecor_exact_riemann_sample(double x, double t, const double* qL, const double* qR,
                               double gamma, double x0, double* q);

/* ---- grid operations (explicit_update.hpp, diagnostic_metrics_timestep.hpp) */
void ecor_load(const ecor_grid* g, const double* U, int i, int j, double* u);
/* Whether ghost (i, j) holds an exact-Dirichlet state, and that state. */
int ecor_exact_ghost(const ecor_grid* g, int i, int j);
void ecor_exact_state(const ecor_exact* e, int nx, int ny, double time, int i, int j, double* u);
/* The problem's exact primitive state at a point (no periodic image). */
void ecor_exact_prim_at(const ecor_exact* e, double x, double y, double time, double* q);
void ecor_reconstruct(const ecor_grid* g, const double* U, int i, int j, int axis,
                      int side, double* q);
void ecor_face_flux(const ecor_grid* g, const double* U, int face_i, int face_j,
                    int axis, double* f);
int ecor_fv_step(const ecor_grid* g, const double* U, double* Unew, double dt);
double ecor_max_rate(const ecor_grid* g, const double* U);
double ecor_euler_dt(double max_rate, double cfl, double time, double stop_time);
void ecor_diagnostics(const ecor_grid* g, const double* U, double* rho_min,
                      double* p_min, int* nonfinite);

typedef struct {
  int steps;
  double time;
  long long degenerate_total;
  int status; /* 0 ok/capped, 1 initial_nonphysical_state, 2 invalid_time_step,
                 3 nonfinite_state, 4 nonpositive_density, 5 nonpositive_pressure */
  double rho_min, p_min;
  int nonfinite;
  double dt_last;
} ecor_run_result;

/* The explicit branch of the driver loop (driver_main.hpp:216-333), run on
 * U in place for at most max_step steps or until stop_time. dts (optional,
 * length max_step) receives the dt of every step. */
void ecor_run_explicit(const ecor_grid* g, double* U, double cfl, double stop_time,
                       int max_step, double* dts, ecor_run_result* out);

/* ---- initial data (problem_initialization.hpp, problem_state_functions.hpp) */
void ecor_init_gresho(int nx, int ny, double xlo, double ylo, double dx, double dy,
                      double gamma, double cx, double cy, double rho, double mach,
                      double p0, double* U);
void ecor_init_toro(int nx, int ny, double xlo, double dx, double gamma, const double* qL,
                    const double* qR, double x0, double* U);

/* ---- IMEX T1/S2 BDLTV20 source-map Picard step (ec_oracle_imex.c) */
enum {
  ECOR_IMEX_CONVERGED_ABSOLUTE = 0,
  ECOR_IMEX_CONVERGED_RELATIVE = 1,
  ECOR_IMEX_NOT_CONVERGED = 2,      /* linear_solve_not_converged */
  ECOR_IMEX_BREAKDOWN = 3,          /* gmres_breakdown_not_converged */
  ECOR_IMEX_DENSITY_FLOOR = 4,      /* density_denominator_floor_used */
  ECOR_IMEX_ROW_NONFINITE = 5,      /* source_map_pressure_row_nonfinite */
  ECOR_IMEX_RESIDUAL_NONFINITE = 6, /* linear_residual_nonfinite */
  ECOR_IMEX_MOMENTUM_NONFINITE = 7  /* source_map_lagged_momentum_nonfinite */
};

typedef struct {
  int iterations, converged, breakdown, nonfinite;
  double residual_linf, relative_residual_linf;
} ecor_gmres_result;

typedef struct {
  int status;
  int iterations;        /* summed over Picard solves */
  int last_iterations;
  int picard_solves;
  int density_floor_count;
  int guard_count;
  int nonfinite_count;
  double residual_linf, relative_residual_linf, rhs_norm;
  double coupled_residual_linf;
  /* evaluate_source_map_final (imex/imex_source_map_step.hpp:75-258) */
  double trial_density_min, trial_pressure_min, trial_internal_energy_min;
  double solved_pressure_min, solved_pressure_max;
  double momentum_correction_linf, energy_correction_linf, eq78_divergence_linf;
} ecor_imex_result;

void ecor_imex_star(const ecor_grid* g, const double* U, double dt, int acoustic, double* S);
long ecor_imex_assemble(const ecor_grid* g, const double* star, const double* p,
                        const double* mom, double dt, long* row_ptr, long* cols, double* vals,
                        double* rhs);
void ecor_gmres_full(long n, const long* rp, const long* cols, const double* vals,
                     const double* b, const double* x0, int max_iterations, double tol,
                     double* x, ecor_gmres_result* res);
/* One IMEX step of U in place (on success); p_out (optional) receives the
 * final pressure. */
void ecor_imex_step(const ecor_grid* g, double* U, double dt, int picard, double tol, int max_iter,
                    double* p_out, ecor_imex_result* out);
double ecor_imex_dt(const ecor_grid* g, const double* U, double cfl, double time, double stop,
                    int acoustic_startup, double acoustic_cap, double* material_max,
                    double* acoustic_max);

/* ---- whole-field metrics (diagnostic_metrics_timestep.hpp, output_cases.hpp) */
void ecor_admissibility(const ecor_grid* g, const double* U, double* out, int* counts);
void ecor_gresho_metrics(const ecor_grid* g, const double* U, double xlo, double ylo, double cx,
                         double cy, double rho, double mach, double p0, double* out);
void ecor_shock_bubble_metrics(const ecor_grid* g, const double* U, double xlo, double ylo,
                               double threshold, double* out);

/* ---- the cylindrical shock bubble's geometric source, in place
 * (problems/boundary_conditions.hpp:75-131) */
void ecor_cylindrical_source(const ecor_grid* g, double* U, double dt, double ylo);

/* ---- direct BDLTV20 T1/S2 step (ec_oracle_direct.c) */
typedef struct {
  ecor_exact exact; /* problem and periodicity; kind TORO or GRESHO gives the
                       exact-Dirichlet boundary, NONE a periodic case */
  double epsilon;   /* euler.bdltv20_paper_epsilon */
  int x_only_row;   /* Riemann strips: pressure row and correction in x only */
} ecor_direct;

typedef struct {
  int solver_failures, iterations_max, iterations, guard_count, nonfinite_count, nonpositive_count;
  double residual_linf_max, relative_residual_linf_max, last_residual_linf;
} ecor_direct_result;

void ecor_eps_to_primitive(const double* u, double gamma, double eps, double* q);
void ecor_eps_to_conserved(const double* q, double gamma, double eps, double* u);
void ecor_direct_boundary_state(const ecor_grid* g, const ecor_direct* d, int i, int j, double t,
                                double* u);
/* One step of U (epsilon-scaled conserved, in place) from time to time+dt. */
void ecor_direct_step(const ecor_grid* g, const ecor_direct* d, double* U, double time, double dt,
                      int picard, double tol, int max_iter, double* p_out, ecor_direct_result* out);
double ecor_direct_dt(const ecor_grid* g, const ecor_direct* d, const double* U, double imex_cfl,
                      double fixed_dt, double time, double stop_time);

#ifdef __cplusplus
}
#endif
#endif
