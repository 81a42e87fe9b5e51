"""Host initial data of the euler_compare test problems (libec_amd.so).

``initial_state(kind, nx, ny, **params)`` returns a numpy float64 array of
shape (4, ny, nx) filled by ec_init_state with the reference's formulas
(problems/problem_initialization.hpp:420-435); params override the defaults of
core/run_config.hpp:2-55 (citations relative to
amrex/apps/euler_compare/src/euler_compare/).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi_ext import ProblemDesc
from ._lib import Layout, check, lib

KINDS = {"advection_blob": 0, "blob": 0, "toro1": 1, "sod": 1, "gresho_vortex": 2, "gresho": 2,
         "riemann_quadrant": 3, "isentropic_vortex": 4, "shock_density_bubble_2d": 5,
         "shock_density_bubble": 5, "shock_bubble_mach": 6}


def problem_desc(kind: str | int, nx: int, ny: int, **params) -> ProblemDesc:
    k = KINDS[kind] if isinstance(kind, str) else int(kind)
    d = ProblemDesc()
    check(None, lib().ec_problem_defaults(k, C.byref(d)), "ec_problem_defaults")
    d.nx, d.ny = nx, ny
    for key, val in params.items():
        cur = getattr(d, key)
        if isinstance(cur, C.Array):
            for i, x in enumerate(val):
                cur[i] = x
        else:
            setattr(d, key, val)
    return d


def initial_state(kind: str | int, nx: int, ny: int, **params) -> np.ndarray:
    d = problem_desc(kind, nx, ny, **params)
    U = np.zeros((4, ny, nx), dtype=np.float64)
    lay = Layout(nx, nx * ny)
    check(None, lib().ec_init_state(C.byref(d), U.ctypes.data_as(C.POINTER(C.c_double)),
                                    C.byref(lay)), "ec_init_state")
    return U


def gresho(n: int, mach: float = 0.01, gamma: float = 1.4) -> np.ndarray:
    """Periodic Gresho vortex on [0,1]^2 centred at (0.5,0.5), rho=1,
    p0 = 1/(gamma M^2) (the benchmark workload)."""
    return initial_state("gresho_vortex", n, n, gamma=gamma, mach=mach,
                         pressure=1.0 / (gamma * mach * mach), density_outer=1.0,
                         vortex_center=(0.5, 0.5))


def post_shock_state(gamma: float, mach: float, rho1: float = 1.0, p1: float = 1.0) -> np.ndarray:
    q = np.zeros(4)
    check(None, lib().ec_post_shock_state(gamma, mach, rho1, p1,
                                          q.ctypes.data_as(C.POINTER(C.c_double))),
          "ec_post_shock_state")
    return q
