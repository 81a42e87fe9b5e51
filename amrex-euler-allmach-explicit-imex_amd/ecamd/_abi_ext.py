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


def declare(L) -> None:
    """Attach argtypes/restype for the entry points beyond the explicit step."""
    from ._lib import Layout
    vp, st, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
    L.ec_problem_defaults.argtypes = [C.c_int, C.POINTER(ProblemDesc)]
    L.ec_problem_defaults.restype = st
    L.ec_init_state.argtypes = [C.POINTER(ProblemDesc), dp, C.POINTER(Layout)]
    L.ec_init_state.restype = st
    L.ec_post_shock_state.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, dp]
    L.ec_post_shock_state.restype = st
