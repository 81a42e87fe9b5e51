/*
 * ec_amd.h -- C-ABI of the MI355X-native euler_compare update path.
 *
 * Plain C: pointers, sizes and POD structs only.  Every entry point returns an
 * ec_status code and never aborts.  Device pointers are hipMalloc'd memory on
 * the context's device; the caller owns them.  One context per host thread;
 * a context is not re-entrant.  All work is enqueued on the context's HIP
 * stream; entry points that return host values synchronise that stream.
 *
 * State layout (any AMReX FArrayBox-compatible strides): a state pointer
 * addresses component 0 of valid cell (0,0); cell (i,j) of component c is at
 * ptr[c*comp_stride + j*row_stride + i], components (rho, mx, my, E)
 * (amrex/apps/euler_compare/src/euler_compare_app.cpp:47).  Ghost cells, if
 * the caller's layout has them, are neither read nor written: the boundary
 * rule of each side is applied on read (wrap, clamp, mirror or fixed state).
 *
 * Reference citations below are relative to
 * amrex/apps/euler_compare/src/euler_compare/.
 */
#ifndef EC_AMD_H
#define EC_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  EC_OK = 0,
  EC_ERR_INVALID_ARGUMENT = 1,
  EC_ERR_HIP = 2,
  EC_ERR_OUT_OF_MEMORY = 3,
  EC_ERR_UNSUPPORTED = 4,
  EC_ERR_TOO_LARGE = 5      /* beyond the fixed workspace of ec_explicit_step_noalloc */
} ec_status;

/* Riemann solver selector: the euler.riemann values of core/solver_types.hpp:11
 * (parsed at core/config_parse_helpers.hpp:44-61). */
typedef enum {
  EC_RIEMANN_RUSANOV = 0,
  EC_RIEMANN_HLLC = 1,
  EC_RIEMANN_LOW_MACH_HLLC = 2,
  EC_RIEMANN_XIE_AM_HLLC_P = 3
} ec_riemann;

/* euler.slope_limiter (core/solver_types.hpp:14). */
typedef enum { EC_LIMITER_MINMOD = 0, EC_LIMITER_MC = 1, EC_LIMITER_VAN_LEER = 2 } ec_limiter;

/* Per-side ghost rule.  PERIODIC and OUTFLOW reproduce
 * problems/boundary_conditions.hpp:2-25; FIXED (inflow state, priority over
 * the other axis) and REFLECT (mirror, normal momentum negated) reproduce the
 * shock-bubble fill at problems/boundary_conditions.hpp:27-73.  EXACT (the
 * problem's exact state, priority over the other axis) is set only by
 * ec_set_exact_boundary. */
typedef enum {
  EC_BC_PERIODIC = 0,
  EC_BC_OUTFLOW = 1,
  EC_BC_REFLECT = 2,
  EC_BC_FIXED = 3,
  EC_BC_EXACT = 4
} ec_bc;

typedef struct {
  int nx, ny;               /* valid cells */
  double dx, dy;            /* cell sizes */
  double gamma;
  int riemann;              /* ec_riemann */
  int limiter;              /* ec_limiter */
  int spatial_order;        /* 1 or 2 */
  int bc[4];                /* ec_bc for x-low, x-high, y-low, y-high */
  double fixed_state[4][4]; /* conserved state of an EC_BC_FIXED side */
} ec_grid_desc;

typedef struct {
  long long row_stride;  /* doubles between (i,j) and (i,j+1) */
  long long comp_stride; /* doubles between component planes */
} ec_layout;

typedef struct ec_ctx ec_ctx;

/* Diagnostics of one state (core/basic_diagnostics.hpp: Diagnostics) plus the
 * CFL spectral-rate maximum and the step's degenerate-star count. */
typedef struct {
  double rho_min;            /* min rho over valid cells, -inf where non-finite */
  double pressure_min;       /* min p, -inf where non-finite */
  int nonfinite_count;
  int degenerate_count;      /* faces counted per adjacent valid cell */
  double max_rate;           /* max (|u|+c)/dx + (|v|+c)/dy */
} ec_step_diag;

/* Device-resident run state of ec_explicit_advance, mirroring the explicit
 * branch of the driver loop (app/driver_main.hpp:224-333). */
typedef struct {
  double time;
  double stop_time;
  double cfl;
  double dt_last, dt_min, dt_max, dt_sum;
  double rho_min, pressure_min, max_rate;
  long long degenerate_total;
  int step;
  int max_step;
  int nonfinite_count;
  /* 0 running/ok, 1 initial_nonphysical_state, 2 invalid_time_step,
   * 3 nonfinite_state, 4 nonpositive_density, 5 nonpositive_pressure */
  int status;
} ec_run_state;

/* ---------------------------------------------------------------- context */
/* Creates a context on the current HIP device.  stream: a hipStream_t, or
 * NULL for the default stream.  No reference counterpart: it replaces the
 * MultiFab/Geometry/RunConfig triple the reference threads through every call
 * (app/driver_main.hpp:17-23). */
ec_status ec_create(const ec_grid_desc* desc, void* stream, ec_ctx** out);
ec_status ec_destroy(ec_ctx* ctx);
const char* ec_status_string(ec_status s);
/* Text of the last HIP error seen by this context (empty if none). */
const char* ec_last_error(const ec_ctx* ctx);
/* Device buffer helpers (optional; any hipMalloc'd memory works). */
ec_status ec_alloc(ec_ctx* ctx, size_t bytes, void** dev_ptr);
ec_status ec_free(ec_ctx* ctx, void* dev_ptr);
ec_status ec_synchronize(ec_ctx* ctx);

/* ----------------------------------------------------- explicit O2 step */
/* One explicit step, out of place: U_out = U_in - dt*div F(U_in).
 * Replaces int euler_fv_one_step(MultiFab&, const Geometry&, const RunConfig&,
 * Real dt, Real time) (numerics/explicit_update.hpp:58-102): fused ghost rule,
 * MUSCL reconstruction, every face flux once, the degenerate-star count
 * (:2-56) and, as an epilogue on U_out, compute_diagnostics
 * (numerics/diagnostic_metrics_timestep.hpp:18-47) and the max-rate reduction
 * of compute_euler_dt (:227-260).  diag may be NULL (no synchronisation);
 * otherwise the stream is synchronised and diag filled. */
ec_status ec_explicit_step(ec_ctx* ctx, const double* U_in, double* U_out,
                           const ec_layout* layout, double dt, ec_step_diag* diag);

/* ec_explicit_step without a context and without allocating anything, for
 * callers that cannot allocate device or host memory (code under HIP graph
 * capture, serving loops that forbid hipMalloc).  No reference counterpart.
 * desc: as for ec_create (no exact-Dirichlet sides); stream: a hipStream_t or
 * NULL; the default 16x15 tile shape.  Results equal ec_explicit_step's for
 * the same grid.  The workspace (run control block and one partial per tile)
 * is a fixed device array of the library, one per GPU, holding
 * EC_NOALLOC_MAX_TILES tile partials: grids needing more tiles than that
 * (ceil(nx/16) * ceil(ny/15) > 65536, e.g. beyond 4096 x 3840) are rejected
 * with EC_ERR_TOO_LARGE.  Calls that overlap on one GPU race on that array.
 * diag may be NULL (the call only enqueues work); otherwise the stream is
 * synchronised and diag filled. */
#define EC_NOALLOC_MAX_TILES 65536
ec_status ec_explicit_step_noalloc(const ec_grid_desc* desc, const double* U_in, double* U_out,
                                   const ec_layout* layout, double dt, ec_step_diag* diag,
                                   void* stream);

/* The same step in place, as the reference mutates its MultiFab: the result
 * goes through a context-owned buffer and is copied back (the reference's
 * MultiFab::Copy, numerics/explicit_update.hpp:100). */
ec_status ec_explicit_step_inplace(ec_ctx* ctx, double* U, const ec_layout* layout,
                                   double dt, ec_step_diag* diag);

/* ---------------------------------------------------- CFL / diagnostics */
/* The standalone max-rate + diagnostics reduction (wavefront shuffle).
 * Replaces the reductions of compute_euler_dt and compute_diagnostics
 * (numerics/diagnostic_metrics_timestep.hpp:18-47, :227-246). */
ec_status ec_state_diagnostics(ec_ctx* ctx, const double* U, const ec_layout* layout,
                               ec_step_diag* diag);
/* dt from a max rate: compute_euler_dt's closing arithmetic
 * (numerics/diagnostic_metrics_timestep.hpp:252-259).  Host-only, no context. */
double ec_euler_dt(double max_rate, double cfl, double time, double stop_time);

/* ------------------------------------------------- device-resident loop */
/* Runs up to nsteps iterations of the explicit driver loop entirely on the
 * device (dt from the previous step's fused max-rate, no host round trip),
 * ping-ponging between U_a and U_b.  The loop stops stepping, exactly as
 * app/driver_main.hpp:224-310 does, on a failure status, at max_step or at
 * stop_time.  On return *result_in_b says which buffer holds the final state
 * and io holds the updated run state.  io->status must be 0 and the initial
 * diagnostics are evaluated on the first call (io->step == 0). */
ec_status ec_explicit_advance(ec_ctx* ctx, double* U_a, double* U_b, const ec_layout* layout,
                              int nsteps, ec_run_state* io, int* result_in_b);

/* Step-kernel timing (measurement hook; no reference counterpart).
 * ec_step_kernel_timing(ctx, 1, N) arms up to N timed launches: each explicit
 * step kernel launched afterwards is bracketed by HIP events on the
 * context's stream (the finalize kernel is outside the bracket).
 * ec_step_kernel_time returns the summed duration and the launch count since
 * arming (or the last read) and rearms; enable = 0 disarms. */
ec_status ec_step_kernel_timing(ec_ctx* ctx, int enable, int max_launches);
ec_status ec_step_kernel_time(ec_ctx* ctx, double* total_ms, int* launches);

/* Kernel tile shape of the explicit step (autotuner hook).  tile_id indexes
 * the compiled shapes; ec_explicit_tile_count() says how many exist and
 * ec_explicit_tile_shape() describes one.  A new context starts with the
 * 16x15 tile (id 0), or with the 24x4 strip tile (id 3) when the grid's 16x15
 * tiles number fewer than 1.5 per CU (up to about 80,000 cells on 256 CUs):
 * a grid that small leaves the chip waiting on one tile's latency, which the
 * strip tile shortens.  Every shape gives the same results bit for bit. */
int ec_explicit_tile_count(void);
ec_status ec_explicit_tile_shape(int tile_id, int* tx, int* ty, int* threads);
ec_status ec_set_explicit_tile(ec_ctx* ctx, int tile_id);

/* ------------------------------------------------ IMEX T1/S2 BDLTV20 step */
/* solver_status of the reference (imex/imex_source_map_step.hpp:398-567). */
typedef enum {
  EC_IMEX_CONVERGED_ABSOLUTE = 0,
  EC_IMEX_CONVERGED_RELATIVE = 1,
  EC_IMEX_NOT_CONVERGED = 2,        /* linear_solve_not_converged */
  EC_IMEX_BREAKDOWN = 3,            /* gmres_breakdown_not_converged */
  EC_IMEX_DENSITY_FLOOR = 4,        /* density_denominator_floor_used */
  EC_IMEX_ROW_NONFINITE = 5,        /* source_map_pressure_row_nonfinite */
  EC_IMEX_RESIDUAL_NONFINITE = 6,   /* linear_residual_nonfinite */
  EC_IMEX_MOMENTUM_NONFINITE = 7,   /* source_map_lagged_momentum_nonfinite */
  EC_IMEX_NOT_RUN = 8               /* not_run */
} ec_imex_status;

/* FP64: restarted GMRES(restart) with CGS2, all in FP64.  MIXED: FP64
 * flexible GMRES(20) whose preconditioner is an FP32 GMRES(restart) with one
 * reduction per Arnoldi step.  FP64_FLEXIBLE: the MIXED structure with the
 * inner solve in FP64 (the same kernels instantiated for double), the
 * comparator that isolates what the FP32 inner precision contributes. */
typedef enum {
  EC_IMEX_PRECISION_FP64 = 0,
  EC_IMEX_PRECISION_MIXED = 1,
  EC_IMEX_PRECISION_FP64_FLEXIBLE = 2
} ec_imex_precision;

typedef struct {
  int picard_iterations;     /* euler.imex_picard_iterations */
  double tol;                /* euler.imex_solver_tol (absolute or relative-to-max(|b|inf,1)) */
  int max_iter;              /* euler.imex_solver_max_iter: Krylov iterations per solve (the
                                flexible forms: inner iterations; an outer cycle stops once
                                they are spent) */
  int restart;               /* GMRES restart length, 1..64 (mixed: of the FP32 inner cycles) */
  int precision;             /* ec_imex_precision: FP64 GMRES, or MIXED: FP64 flexible outer
                                GMRES(20) preconditioned by FP32 GMRES(restart) */
  int predictor_dissipation; /* euler.imex_predictor_dissipation: 0 material, 1 acoustic */
  double inner_reduction;    /* mixed: relative 2-norm reduction asked of each FP32 inner solve */
} ec_imex_opts;

typedef struct {
  int status;                /* ec_imex_status */
  int iterations;            /* Krylov iterations summed over the Picard solves (mixed: FP32) */
  int last_iterations;
  int picard_solves;
  int refinements;           /* mixed: FP64 outer steps (one FP32 inner solve each), last solve */
  int density_floor_count;
  int guard_count;           /* face-enthalpy denominator guards in the closeout */
  int nonfinite_count;
  double residual_linf;      /* |A p - b|_inf of the last solve */
  double relative_residual_linf;
  double rhs_norm;           /* max(|b|_inf, 1) */
  double coupled_residual_linf; /* max |p/(g-1) - E + K| after the closeout (inf once a
                                   cell is not finite) */
  /* the closeout's evaluation (evaluate_source_map_final,
   * imex/imex_source_map_step.hpp:75-258), over the cells whose values are
   * finite; NaN where no cell is (or the step ended before the closeout) */
  double trial_density_min;       /* min rho (before the closeout: of the star) */
  double trial_pressure_min;      /* min (g-1)(E - K) */
  double trial_internal_energy_min; /* min E - K */
  double solved_pressure_min, solved_pressure_max;
  double momentum_correction_linf; /* max |m - m*| (0 before the closeout) */
  double energy_correction_linf;  /* max |E - E*|, also the closeout delta (0 before) */
  double eq78_divergence_linf;    /* max |central velocity divergence| (inf once a cell
                                     is not finite) */
} ec_imex_diag;

/* Defaults of core/run_config.hpp:2-55 (2 Picard solves, tol 1e-10,
 * max_iter 100) with restart 30 and FP64 iterations. */
ec_status ec_imex_default_opts(ec_imex_opts* opts);

/* One IMEX step of U in place.  Replaces ImexStepDiagnostics
 * euler_imex_bdltv20_o1_source_map_picard_one_step(MultiFab&, const
 * Geometry&, const RunConfig&, Real dt, MultiFab*) (imex/
 * imex_source_map_step.hpp:300-616).  As there, U changes only when the step
 * succeeds (diag->status 0 or 1).  p_out (optional, device, nx*ny compact)
 * receives the final pressure (the reference's solved_pressure_check). */
ec_status ec_imex_step(ec_ctx* ctx, double* U, const ec_layout* layout, double dt,
                       const ec_imex_opts* opts, ec_imex_diag* diag, double* p_out);

/* compute_imex_dt (numerics/diagnostic_metrics_timestep.hpp:262-326): the eq79
 * material rate 2|u|/min(dx,dy) and the acoustic rate, reduced on the device. */
ec_status ec_imex_dt(ec_ctx* ctx, const double* U, const ec_layout* layout, double cfl,
                     double time, double stop_time, int acoustic_startup,
                     double acoustic_cfl_cap, double* dt, double* material_rate_max,
                     double* acoustic_rate_max);

/* ------------------------------------------------------------- adjoint */
/* Adjoint of one explicit O2 step (Rusanov, HLLC, low-Mach HLLC or HLLC-P;
 * any limiter and ghost rule), with dt and every branch frozen at the
 * forward state U_n:  lambda_n = (dU_{n+1}/dU_n)^T lambda_next.  No reference
 * counterpart (the reference has no adjoint); it differentiates
 * euler_fv_one_step (numerics/explicit_update.hpp:58-102). */
ec_status ec_adjoint_step(ec_ctx* ctx, const double* U_n, const double* lambda_next,
                          double* lambda_n, const ec_layout* layout, double dt);

/* Forward driver loop (device dt, as ec_explicit_advance) for nsteps steps,
 * storing the state every checkpoint_every steps into checkpoints (device,
 * compact 4*nx*ny doubles each, ceil(nsteps/checkpoint_every) of them) and
 * every step's dt into dts (host, nsteps). */
ec_status ec_forward_checkpointed(ec_ctx* ctx, const double* U0, double* U_final,
                                  const ec_layout* layout, int nsteps, int checkpoint_every,
                                  double cfl, double stop_time, double* checkpoints, double* dts,
                                  int* steps_done);

/* Reverse sweep over nsteps steps from the checkpoints: each segment is
 * recomputed once from its checkpoint with the recorded dts, then the
 * adjoint steps run backwards.  lambda_N and lambda_0 use layout. */
ec_status ec_adjoint_checkpointed(ec_ctx* ctx, const double* checkpoints, const ec_layout* layout,
                                  int nsteps, int checkpoint_every, const double* dts,
                                  const double* lambda_N, double* lambda_0);

/* ------------------------------------------- pointwise verification hooks */
/* Host-side evaluation of one face flux (the device code's arithmetic) and of
 * its hand-written vector-Jacobian product: qL/qR primitive (rho,u,v,p),
 * F = interface_flux (numerics/reconstruction_fluxes.hpp:373-388), and
 * qLbar/qRbar/sensor_bar = (dF/d.)^T Fbar at the forward state's branches.
 * No reference counterpart: they let a caller check the adjoint against
 * finite differences without a GPU. */
ec_status ec_flux_host(const double* qL, const double* qR, int axis, double gamma, int riemann,
                       double sensor_f, double* F);
/* The four adjoint passes of ec_adjoint_step run on the host, on host
 * memory (compact layout), with the device code's arithmetic. */
ec_status ec_adjoint_step_host(const ec_grid_desc* desc, const double* U_n,
                               const double* lambda_next, double* lambda_n, double dt);
ec_status ec_flux_vjp_host(const double* qL, const double* qR, int axis, double gamma,
                           int riemann, double sensor_f, const double* Fbar, double* qLbar,
                           double* qRbar, double* sensor_bar);

/* -------------------------------------------------------------- ensembles */
/* A batch of independent explicit runs of uneven resolution (Riemann strips,
 * Gresho boxes, ...) stepped together on one GPU.  No reference counterpart:
 * the reference runs one problem per process (app/driver_main.hpp:17-333).
 * Every member follows that driver loop exactly as ec_explicit_advance does
 * on the member alone (same dt sequence, statuses and bit-identical states);
 * the members share gamma, riemann, limiter and spatial order. */
typedef struct {
  int nx, ny;               /* valid cells */
  double dx, dy;
  int bc[4];                /* ec_bc for x-low, x-high, y-low, y-high */
  double fixed_state[4][4]; /* conserved state of an EC_BC_FIXED side */
  double cfl;               /* euler.cfl */
  double stop_time;         /* the run starts at time 0 */
  int max_step;
} ec_member_desc;

typedef struct ec_batch ec_batch;

/* How rounds (one step of every running member) are separated:
 * GRAPH: one kernel launch per round, rounds captured in a hipGraph and
 * replayed; PERSISTENT: one cooperative launch that loops over rounds with
 * bounded per-member waits; AUTO: the autotuner's choice. */
typedef enum { EC_BATCH_AUTO = 0, EC_BATCH_GRAPH = 1, EC_BATCH_PERSISTENT = 2 } ec_batch_mode;

typedef struct {
  int mode;     /* ec_batch_mode */
  int tile_id;  /* ec_batch_tile_shape index */
  double ms;    /* wall time of one complete run, -1 if the pair cannot run */
} ec_batch_timing;

/* Creates the batch on the current HIP device; member states start at zero.
 * stream: the caller's hipStream_t (or NULL); the batch orders its own
 * stream after it on every call. */
ec_status ec_batch_create(const ec_member_desc* members, int n_members, double gamma,
                          int riemann, int limiter, int spatial_order, void* stream,
                          ec_batch** out);
ec_status ec_batch_destroy(ec_batch* b);
const char* ec_batch_last_error(const ec_batch* b);
/* Copies a member's state in or out; the pointer may be host or device
 * memory with the given layout.  Synchronous. */
ec_status ec_batch_set_state(ec_batch* b, int member, const double* src, const ec_layout* layout);
ec_status ec_batch_get_state(ec_batch* b, int member, double* dst, const ec_layout* layout);
/* Compiled tile shapes of the ensemble kernels. */
int ec_batch_tile_count(void);
ec_status ec_batch_tile_shape(int tile_id, int* tx, int* ty, int* threads);
/* mode: ec_batch_mode; tile_id -1 lets the autotuner choose; rounds_per_graph
 * (even, 0 keeps the current value) is the number of rounds per hipGraph. */
ec_status ec_batch_configure(ec_batch* b, int mode, int tile_id, int rounds_per_graph);
/* Graph mode: the members split into `groups` (1..4, balanced by tile count),
 * each group's rounds replayed on a stream of its own so that the groups'
 * rounds overlap; the autotuner picks the count (no reference counterpart:
 * the reference runs one grid at a time). */
ec_status ec_batch_set_groups(ec_batch* b, int groups);
int ec_batch_groups(const ec_batch* b);
/* Times one complete run of every graph-mode tile shape on a copy of the
 * current states (which are restored) and keeps the fastest.  Persistent mode
 * is not timed (it measured about twice graph mode's time on every shape: its
 * per-member waits need device-scope fences); ec_batch_configure still
 * selects it.  table may be NULL; *n_entries receives the number of pairs
 * timed. */
ec_status ec_batch_autotune(ec_batch* b, ec_batch_timing* table, int table_cap, int* n_entries);
/* Runs every member from its current state at time 0 until stop_time,
 * max_step or a failure status; the final states replace the current ones.
 * results (n_members entries, may be NULL) receive the members' run states;
 * used (may be NULL) the mode, tile shape and wall time of the run.
 * Persistent mode returns EC_ERR_HIP if a member wait timed out or a member
 * still ran after the largest max_step of the batch (the launch ends either
 * way). */
ec_status ec_batch_run(ec_batch* b, ec_run_state* results, ec_batch_timing* used);

/* ---------------------------------------------------- host initial data */
/* Test problems of problems/problem_initialization.hpp:420-435 (selector
 * euler.problem, core/config_parse_helpers.hpp:11-42), plus
 * EC_PROBLEM_SHOCK_BUBBLE_MACH: a shock of Mach number shock_mach at
 * x = shock_x running into gas at rest (density_outer, pressure) with a
 * circular bubble of density density_inner at blob_center, radius
 * blob_radius (the configurable generalisation of the reference's fixed
 * shock-density bubble, problems/problem_initialization.hpp:2-102). */
typedef enum {
  EC_PROBLEM_ADVECTION_BLOB = 0,
  EC_PROBLEM_TORO = 1,
  EC_PROBLEM_GRESHO = 2,
  EC_PROBLEM_RIEMANN_QUADRANT = 3,
  EC_PROBLEM_ISENTROPIC_VORTEX = 4,
  EC_PROBLEM_SHOCK_DENSITY_BUBBLE = 5,
  EC_PROBLEM_SHOCK_BUBBLE_MACH = 6
} ec_problem_kind;

typedef struct {
  int kind;                 /* ec_problem_kind */
  int nx, ny;
  double prob_lo[2], prob_hi[2];
  double gamma;
  int toro_test;            /* prob.test */
  double riemann_interface_x, riemann_interface_y;
  double velocity[2];
  double density_inner, density_outer, pressure, mach;
  double blob_center[2], blob_radius;
  double vortex_center[2];
  double isentropic_vortex_strength;
  double shock_mach, shock_x;
} ec_problem_desc;

/* Defaults of core/run_config.hpp:2-55 for the given problem kind. */
ec_status ec_problem_defaults(int kind, ec_problem_desc* desc);
/* Writes the initial conserved state into HOST memory with the given layout
 * (initialize_state, problems/problem_initialization.hpp:420-435; cell
 * centres prob_lo + (i+1/2) dx with dx = (prob_hi - prob_lo)/n). */
ec_status ec_init_state(const ec_problem_desc* desc, double* host_U, const ec_layout* layout);
/* The problem's exact primitive state (rho, u, v, p) at (x, y) and time t:
 * exact_riemann_sample for Toro (problems/exact_riemann.hpp:114-205),
 * gresho_state, the isentropic vortex advected with the flow
 * (problems/problem_state_functions.hpp:44-71), exact_advection_blob_state
 * (problems/problem_initialization.hpp:186-207; periodic: geometry.is_periodic)
 * and the quadrant's initial state.  Host-only. */
ec_status ec_exact_state_at(const ec_problem_desc* problem, const int* periodic, double x, double y,
                            double t, double* prim_out);
/* Post-shock primitive state (rho, u, v, p) behind a shock of the given Mach
 * number moving into gas at rest (Rankine-Hugoniot). */
ec_status ec_post_shock_state(double gamma, double mach, double rho1, double p1, double* prim_out);

/* ------------------------------------------- exact-Dirichlet ghost cells */
/* euler.field_boundary=exact_dirichlet for the context's grid.  Replaces the
 * exact branches of void fill_problem_ghosts(MultiFab&, const Geometry&,
 * const RunConfig&, Real time) (problems/boundary_conditions.hpp:202-219):
 * fill_gresho_exact_dirichlet_ghosts (:133-163),
 * fill_toro_x_exact_dirichlet_ghosts, fill_riemann_quadrant_exact_dirichlet_ghosts
 * and fill_advection_blob_exact_dirichlet_ghosts
 * (problems/problem_initialization.hpp:308-342, :344-375, :209-263).
 * problem: kind GRESHO, TORO, RIEMANN_QUADRANT or ADVECTION_BLOB with the
 * context's nx, ny and gamma; periodic: geometry.is_periodic (2 ints).  With
 * an axis not periodic, the sides of every non-periodic axis (and the x sides
 * of a Toro run) become EC_BC_EXACT and the others EC_BC_PERIODIC; with both
 * axes periodic every side becomes EC_BC_PERIODIC, as in the reference.
 * Ghost states are evaluated on read: Gresho and the quadrant from a table
 * built here, the Toro exact solution and the advected blob at the step's
 * time -- the time set by ec_set_boundary_time for ec_explicit_step and
 * ec_explicit_step_inplace, the run's own time in ec_explicit_advance.
 * ec_imex_step reads EC_BC_EXACT sides as outflow (periodic on a periodic
 * axis), as the source-map step clamps the physical ghosts before its star
 * update whatever the field boundary (imex/imex_source_map_step.hpp:405-409);
 * the adjoint entry points return EC_ERR_UNSUPPORTED on such a context.
 * Synchronises the context's stream. */
ec_status ec_set_exact_boundary(ec_ctx* ctx, const ec_problem_desc* problem, const int* periodic);
/* The time argument of euler_fv_one_step (numerics/explicit_update.hpp:58-65)
 * for the exact ghost states of the next host-dt steps (default 0). */
ec_status ec_set_boundary_time(ec_ctx* ctx, double time);
/* The same exact-Dirichlet sides for one member of a batched ensemble
 * (problem: the member's nx, ny and the batch's gamma).  The ghost states
 * take each step's start time, as the explicit branch of the driver loop
 * passes it (app/driver_main.hpp:289, numerics/explicit_update.hpp:58-65).
 * Replaces the member's bc; synchronises the batch's stream. */
ec_status ec_batch_set_exact_boundary(ec_batch* b, int member, const ec_problem_desc* problem,
                                      const int* periodic);

/* ------------------------------------------ direct BDLTV20 T1/S2 driver */
/* The cases of euler.bdltv20_paper_t1_s2 (core/config_reader.hpp:136-142),
 * each with the periodicity the reference requires of it
 * (core/config_reader.hpp:228-250). */
typedef enum {
  EC_DIRECT_TORO_X_EXACT_Y_PERIODIC = 0, /* toro_x_exact_dirichlet_y_periodic_2d: is_periodic 0 1 */
  EC_DIRECT_TORO_XY_EXACT = 1,           /* toro_xy_exact_dirichlet_2d: 0 0 */
  EC_DIRECT_ISENTROPIC_VORTEX_PERIODIC = 2, /* isentropic_vortex_periodic_2d: 1 1 */
  EC_DIRECT_GRESHO_EXACT = 3,            /* gresho_exact_dirichlet_2d: 0 0 */
  EC_DIRECT_ADVECTION_BLOB_PERIODIC = 4  /* advection_blob_periodic_2d: 1 1 */
} ec_direct_case;

/* Per-step counters of the direct driver (imex/imex_bdltv20_direct_driver.hpp:
 * 140-147); a run passes when every count is zero. */
typedef struct {
  int solver_failures;   /* Picard solves not converged; the step continues, as the reference does */
  int iterations_max;    /* largest Krylov iteration count of one solve (mixed: FP32) */
  int iterations;        /* summed over the Picard solves */
  int guard_count;       /* face-enthalpy zero-denominator guards */
  int nonfinite_count;
  int nonpositive_count;
  double residual_linf_max;
  double relative_residual_linf_max;
} ec_direct_diag;

/* Prepares the context for the direct step of one case: problem (kind TORO,
 * ISENTROPIC_VORTEX, GRESHO or ADVECTION_BLOB to match the case, with the
 * context's nx, ny and gamma) gives the exact solution sampled at the domain
 * edge for the exact-Dirichlet cases (:101-131); epsilon is
 * euler.bdltv20_paper_epsilon.  The context's ghost-rule sides are not used
 * by the direct step.  No reference counterpart: it replaces the set-up part
 * of int run_bdltv20_paper_t1_s2(const RunConfig&) (:44-138). */
ec_status ec_direct_setup(ec_ctx* ctx, int direct_case, const ec_problem_desc* problem,
                          double epsilon);
/* One step of U (epsilon-scaled conserved state, E = p/(g-1) + eps rho |u|^2/2)
 * from time to time + dt: the loop body of run_bdltv20_paper_t1_s2
 * (imex/imex_bdltv20_direct_driver.hpp:148-601) -- star predictor, Picard
 * iterations of face enthalpies, rows with the exact-Dirichlet pressures
 * folded into the right-hand side and the pressure solve (opts: Picard count,
 * tol, max_iter, restart, precision), the momentum correction and the
 * pressure closeout.  p_out (optional, device, nx*ny) receives the pressure. */
ec_status ec_direct_step(ec_ctx* ctx, double* U, const ec_layout* layout, double time, double dt,
                         const ec_imex_opts* opts, ec_direct_diag* diag, double* p_out);
/* compute_bdltv20_paper_t1_s2_dt (:8-41): fixed_dt > 0 (euler.bdltv20_paper_t1_s2_dt)
 * clipped to the stop time, else imex_cfl / max(2 |u| / min(dx, dy)). */
ec_status ec_direct_dt(ec_ctx* ctx, const double* U, const ec_layout* layout, double imex_cfl,
                       double fixed_dt, double time, double stop_time, double* dt);

/* ------------------------------------------------ whole-field metrics */
/* compute_admissibility_diagnostics (numerics/diagnostic_metrics_timestep.hpp:
 * 49-102): minima of rho, p and internal energy (-inf where a cell fails) and
 * the failure counts, as the IMEX loop takes them after every step
 * (app/driver_main.hpp:278-282). */
typedef struct {
  double rho_min, pressure_min, internal_energy_min;
  int nonfinite_count, density_failure_count, pressure_failure_count, internal_energy_failure_count;
} ec_admissibility_diag;
ec_status ec_admissibility(ec_ctx* ctx, const double* U, const ec_layout* layout,
                           ec_admissibility_diag* out);

/* compute_gresho_metrics (numerics/diagnostic_metrics_timestep.hpp:147-225)
 * against the Gresho vortex of problem (kind GRESHO, the context's grid). */
typedef struct {
  double density_l1_error, velocity_l1_error, pressure_l1_error;
  double pressure_perturbation_l1_error, pressure_perturbation_l1_relative_error;
  double kinetic_energy_initial, kinetic_energy_final, kinetic_energy_ratio;
  double reference_sound_speed, velocity_scale;
} ec_gresho_metrics_out;
ec_status ec_gresho_metrics(ec_ctx* ctx, const double* U, const ec_layout* layout,
                            const ec_problem_desc* problem, ec_gresho_metrics_out* out);

/* compute_shock_density_bubble_metrics (io/output_cases.hpp:329-407): extrema,
 * the area and centroid of rho < threshold (the reference uses 0.55) and the
 * largest density gradient, the neighbours read through the context's ghost
 * rule.  prob_lo: the domain's low corner (2 doubles). */
typedef struct {
  double rho_min, rho_max, pressure_min, pressure_max, internal_energy_min;
  int nonfinite_count;
  double density_bubble_area_threshold, density_bubble_area;
  double density_bubble_centroid_x, density_bubble_centroid_y;
  double schlieren_max;
} ec_shock_bubble_metrics_out;
ec_status ec_shock_bubble_metrics(ec_ctx* ctx, const double* U, const ec_layout* layout,
                                  const double* prob_lo, double threshold,
                                  ec_shock_bubble_metrics_out* out);

/* ----------------------------------------------------- geometric source */
/* Replaces void apply_shock_density_bubble_cylindrical_source(MultiFab&,
 * const Geometry&, const RunConfig&, Real dt)
 * (problems/boundary_conditions.hpp:75-131), which the driver calls after
 * every update of the cylindrical shock-density bubble
 * (app/driver_main.hpp:291).  Reads y as the radius r = y_lo + (j + 1/2) dy
 * and integrates dq/dt = -(rho v, rho u v, rho v^2, v (E + p)) / r over dt
 * in place with the midpoint rule; rows with r <= 0 are left alone.  Queued
 * on the context's stream; returns without synchronising. */
ec_status ec_cylindrical_source(ec_ctx* ctx, double* U, const ec_layout* layout, double dt, double y_lo);

#ifdef __cplusplus
}
#endif
#endif
