"""ctypes signatures of the entry points of ec_amd.h beyond the explicit step."""
from __future__ import annotations

import ctypes as C


class ProblemDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
                ("prob_lo", C.c_double * 2), ("prob_hi", C.c_double * 2), ("gamma", C.c_double),
                ("toro_test", C.c_int), ("riemann_interface_x", C.c_double),
                ("riemann_interface_y", C.c_double), ("velocity", C.c_double * 2),
                ("density_inner", C.c_double), ("density_outer", C.c_double),
                ("pressure", C.c_double), ("mach", C.c_double), ("blob_center", C.c_double * 2),
                ("blob_radius", C.c_double), ("vortex_center", C.c_double * 2),
                ("isentropic_vortex_strength", C.c_double), ("shock_mach", C.c_double),
                ("shock_x", C.c_double)]


class ImexOpts(C.Structure):
    _fields_ = [("picard_iterations", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int),
                ("restart", C.c_int), ("precision", C.c_int), ("predictor_dissipation", C.c_int),
                ("inner_reduction", C.c_double)]


class ImexDiag(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("last_iterations", C.c_int),
                ("picard_solves", C.c_int), ("refinements", C.c_int),
                ("density_floor_count", C.c_int), ("guard_count", C.c_int),
                ("nonfinite_count", C.c_int), ("residual_linf", C.c_double),
                ("relative_residual_linf", C.c_double), ("rhs_norm", C.c_double),
                ("coupled_residual_linf", C.c_double), ("trial_density_min", C.c_double),
                ("trial_pressure_min", C.c_double), ("trial_internal_energy_min", C.c_double),
                ("solved_pressure_min", C.c_double), ("solved_pressure_max", C.c_double),
                ("momentum_correction_linf", C.c_double), ("energy_correction_linf", C.c_double),
                ("eq78_divergence_linf", C.c_double)]


class DirectDiag(C.Structure):
    _fields_ = [("solver_failures", C.c_int), ("iterations_max", C.c_int), ("iterations", C.c_int),
                ("guard_count", C.c_int), ("nonfinite_count", C.c_int),
                ("nonpositive_count", C.c_int), ("residual_linf_max", C.c_double),
                ("relative_residual_linf_max", C.c_double)]


DIRECT_CASES = {"toro_x_exact_dirichlet_y_periodic_2d": 0, "toro_xy_exact_dirichlet_2d": 1,
                "isentropic_vortex_periodic_2d": 2, "gresho_exact_dirichlet_2d": 3,
                "advection_blob_periodic_2d": 4}


class AdmissibilityDiag(C.Structure):
    _fields_ = [("rho_min", C.c_double), ("pressure_min", C.c_double),
                ("internal_energy_min", C.c_double), ("nonfinite_count", C.c_int),
                ("density_failure_count", C.c_int), ("pressure_failure_count", C.c_int),
                ("internal_energy_failure_count", C.c_int)]


class GreshoMetricsOut(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "density_l1_error", "velocity_l1_error", "pressure_l1_error",
        "pressure_perturbation_l1_error", "pressure_perturbation_l1_relative_error",
        "kinetic_energy_initial", "kinetic_energy_final", "kinetic_energy_ratio",
        "reference_sound_speed", "velocity_scale")]


class ShockBubbleMetricsOut(C.Structure):
    _fields_ = [("rho_min", C.c_double), ("rho_max", C.c_double), ("pressure_min", C.c_double),
                ("pressure_max", C.c_double), ("internal_energy_min", C.c_double),
                ("nonfinite_count", C.c_int), ("density_bubble_area_threshold", C.c_double),
                ("density_bubble_area", C.c_double), ("density_bubble_centroid_x", C.c_double),
                ("density_bubble_centroid_y", C.c_double), ("schlieren_max", C.c_double)]


class MemberDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("bc", C.c_int * 4), ("fixed_state", (C.c_double * 4) * 4), ("cfl", C.c_double),
                ("stop_time", C.c_double), ("max_step", C.c_int)]


class BatchTiming(C.Structure):
    _fields_ = [("mode", C.c_int), ("tile_id", C.c_int), ("ms", C.c_double)]


IMEX_STATUS = {0: "converged_absolute", 1: "converged_relative", 2: "linear_solve_not_converged",
               3: "gmres_breakdown_not_converged", 4: "density_denominator_floor_used",
               5: "source_map_pressure_row_nonfinite", 6: "linear_residual_nonfinite",
               7: "source_map_lagged_momentum_nonfinite", 8: "not_run"}


def declare(L) -> None:
    """Attach argtypes/restype for the entry points beyond the explicit step."""
    from ._lib import Layout
    vp, st, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
    L.ec_problem_defaults.argtypes = [C.c_int, C.POINTER(ProblemDesc)]
    L.ec_problem_defaults.restype = st
    L.ec_init_state.argtypes = [C.POINTER(ProblemDesc), dp, C.POINTER(Layout)]
    L.ec_init_state.restype = st
    L.ec_post_shock_state.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp]
    L.ec_set_exact_boundary.argtypes = [vp, C.POINTER(ProblemDesc), C.POINTER(C.c_int)]
    L.ec_set_exact_boundary.restype = st
    L.ec_set_boundary_time.argtypes = [vp, C.c_double]
    L.ec_set_boundary_time.restype = st
    L.ec_post_shock_state.restype = st
    from ._lib import GridDesc
    L.ec_adjoint_step_host.argtypes = [C.POINTER(GridDesc), dp, dp, dp, C.c_double]
    L.ec_adjoint_step_host.restype = st
    L.ec_flux_host.argtypes = [dp, dp, C.c_int, C.c_double, C.c_int, C.c_double, dp]
    L.ec_flux_host.restype = st
    L.ec_flux_vjp_host.argtypes = [dp, dp, C.c_int, C.c_double, C.c_int, C.c_double, dp, dp, dp, dp]
    L.ec_flux_vjp_host.restype = st
    L.ec_adjoint_step.argtypes = [vp, vp, vp, vp, C.POINTER(Layout), C.c_double]
    L.ec_adjoint_step.restype = st
    L.ec_forward_checkpointed.argtypes = [vp, vp, vp, C.POINTER(Layout), C.c_int, C.c_int,
                                          C.c_double, C.c_double, vp, dp, C.POINTER(C.c_int)]
    L.ec_forward_checkpointed.restype = st
    L.ec_adjoint_checkpointed.argtypes = [vp, vp, C.POINTER(Layout), C.c_int, C.c_int, dp, vp, vp]
    L.ec_adjoint_checkpointed.restype = st
    L.ec_imex_default_opts.argtypes = [C.POINTER(ImexOpts)]
    L.ec_imex_default_opts.restype = st
    L.ec_imex_step.argtypes = [vp, vp, C.POINTER(Layout), C.c_double, C.POINTER(ImexOpts),
                               C.POINTER(ImexDiag), vp]
    L.ec_imex_step.restype = st
    L.ec_imex_dt.argtypes = [vp, vp, C.POINTER(Layout), C.c_double, C.c_double, C.c_double, C.c_int,
                             C.c_double, dp, dp, dp]
    L.ec_imex_dt.restype = st
    L.ec_admissibility.argtypes = [vp, vp, C.POINTER(Layout), C.POINTER(AdmissibilityDiag)]
    L.ec_admissibility.restype = st
    L.ec_gresho_metrics.argtypes = [vp, vp, C.POINTER(Layout), C.POINTER(ProblemDesc),
                                    C.POINTER(GreshoMetricsOut)]
    L.ec_gresho_metrics.restype = st
    L.ec_shock_bubble_metrics.argtypes = [vp, vp, C.POINTER(Layout), dp, C.c_double,
                                          C.POINTER(ShockBubbleMetricsOut)]
    L.ec_shock_bubble_metrics.restype = st
    L.ec_cylindrical_source.argtypes = [vp, vp, C.POINTER(Layout), C.c_double, C.c_double]
    L.ec_cylindrical_source.restype = st
    L.ec_exact_state_at.argtypes = [C.POINTER(ProblemDesc), C.POINTER(C.c_int), C.c_double,
                                    C.c_double, C.c_double, dp]
    L.ec_exact_state_at.restype = st
    L.ec_direct_setup.argtypes = [vp, C.c_int, C.POINTER(ProblemDesc), C.c_double]
    L.ec_direct_setup.restype = st
    L.ec_direct_step.argtypes = [vp, vp, C.POINTER(Layout), C.c_double, C.c_double,
                                 C.POINTER(ImexOpts), C.POINTER(DirectDiag), vp]
    L.ec_direct_step.restype = st
    L.ec_direct_dt.argtypes = [vp, vp, C.POINTER(Layout), C.c_double, C.c_double, C.c_double,
                               C.c_double, dp]
    L.ec_direct_dt.restype = st
    from ._lib import RunState
    ip = C.POINTER(C.c_int)
    L.ec_batch_create.argtypes = [C.POINTER(MemberDesc), C.c_int, C.c_double, C.c_int, C.c_int,
                                  C.c_int, vp, C.POINTER(vp)]
    L.ec_batch_create.restype = st
    L.ec_batch_destroy.argtypes = [vp]
    L.ec_batch_destroy.restype = st
    L.ec_batch_last_error.argtypes = [vp]
    L.ec_batch_last_error.restype = C.c_char_p
    L.ec_batch_set_state.argtypes = [vp, C.c_int, vp, C.POINTER(Layout)]
    L.ec_batch_set_state.restype = st
    L.ec_batch_set_exact_boundary.argtypes = [vp, C.c_int, C.POINTER(ProblemDesc), C.POINTER(C.c_int)]
    L.ec_batch_set_exact_boundary.restype = st
    L.ec_batch_get_state.argtypes = [vp, C.c_int, vp, C.POINTER(Layout)]
    L.ec_batch_get_state.restype = st
    L.ec_batch_tile_count.argtypes = []
    L.ec_batch_tile_count.restype = C.c_int
    L.ec_batch_tile_shape.argtypes = [C.c_int, ip, ip, ip]
    L.ec_batch_tile_shape.restype = st
    L.ec_batch_configure.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.ec_batch_configure.restype = st
    L.ec_batch_set_groups.argtypes = [vp, C.c_int]
    L.ec_batch_set_groups.restype = st
    L.ec_batch_groups.argtypes = [vp]
    L.ec_batch_groups.restype = C.c_int
    L.ec_batch_autotune.argtypes = [vp, C.POINTER(BatchTiming), C.c_int, ip]
    L.ec_batch_autotune.restype = st
    L.ec_batch_run.argtypes = [vp, C.POINTER(RunState), C.POINTER(BatchTiming)]
    L.ec_batch_run.restype = st
