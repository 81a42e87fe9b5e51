"""ORACLE -- test infrastructure only.

ctypes binding of the CPU restatement in ``oracle/`` (``libec_oracle.so``,
built by ``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` import this module, and only as the
checker.  Fields are numpy float64 arrays of shape (4, ny, nx), comp-major,
components (rho, mx, my, E).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# EC_ORACLE_LIB: another build of the same sources (the sanitizer build)
LIB_PATH = Path(os.environ.get("EC_ORACLE_LIB", HERE / "libec_oracle.so"))

RUSANOV, HLLC, LOWMACH_HLLC, XIE_HLLCP = 0, 1, 2, 3
MINMOD, MC, VANLEER = 0, 1, 2
BC_PERIODIC, BC_OUTFLOW, BC_REFLECT, BC_FIXED = 0, 1, 2, 3

RIEMANN_NAMES = {"rusanov": RUSANOV, "hllc": HLLC, "low_mach_hllc": LOWMACH_HLLC,
                 "xie_am_hllc_p": XIE_HLLCP}

_dp = C.POINTER(C.c_double)


EXACT_NONE, EXACT_GRESHO, EXACT_TORO, EXACT_QUADRANT, EXACT_BLOB = 0, 1, 2, 3, 4


class Exact(C.Structure):
    _fields_ = [("kind", C.c_int), ("periodic", C.c_int * 2), ("prob_lo", C.c_double * 2),
                ("prob_hi", C.c_double * 2), ("gamma", C.c_double),
                ("vortex_center", C.c_double * 2), ("density_outer", C.c_double),
                ("pressure", C.c_double), ("mach", C.c_double), ("interface_x", C.c_double),
                ("interface_y", C.c_double), ("ql", C.c_double * 4), ("qr", C.c_double * 4),
                ("blob_center", C.c_double * 2), ("blob_radius", C.c_double),
                ("density_inner", C.c_double), ("velocity", C.c_double * 2)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("gamma", C.c_double), ("riemann", C.c_int), ("limiter", C.c_int),
                ("order", C.c_int), ("bc", C.c_int * 4), ("fixed", (C.c_double * 4) * 4),
                ("exact", C.POINTER(Exact)), ("time", C.c_double)]


class Direct(C.Structure):
    _fields_ = [("exact", Exact), ("epsilon", C.c_double), ("x_only_row", C.c_int)]


class DirectResult(C.Structure):
    _fields_ = [("solver_failures", C.c_int), ("iterations_max", C.c_int), ("iterations", C.c_int),
                ("guard_count", C.c_int), ("nonfinite_count", C.c_int),
                ("nonpositive_count", C.c_int), ("residual_linf_max", C.c_double),
                ("relative_residual_linf_max", C.c_double), ("last_residual_linf", C.c_double)]


def make_exact(kind, periodic, prob_lo=(0.0, 0.0), prob_hi=(1.0, 1.0), gamma=1.4, **kw) -> Exact:
    """Exact-Dirichlet ghost description; keyword names are the Exact fields."""
    e = Exact()
    e.kind = kind
    e.periodic[0], e.periodic[1] = int(periodic[0]), int(periodic[1])
    e.prob_lo[0], e.prob_lo[1] = prob_lo
    e.prob_hi[0], e.prob_hi[1] = prob_hi
    e.gamma = gamma
    for k, v in kw.items():
        f = getattr(e, k)
        if isinstance(f, float) or isinstance(f, int):
            setattr(e, k, v)
        else:
            for c, x in enumerate(v):
                f[c] = float(x)
    return e


def set_exact(g: Grid, e: Exact | None, time: float = 0.0) -> Grid:
    """Attach (or detach) exact ghost states and their time to a grid."""
    g.exact = C.pointer(e) if e is not None else C.POINTER(Exact)()
    g.time = float(time)
    return g


class RunResult(C.Structure):
    _fields_ = [("steps", C.c_int), ("time", C.c_double), ("degenerate_total", C.c_longlong),
                ("status", C.c_int), ("rho_min", C.c_double), ("p_min", C.c_double),
                ("nonfinite", C.c_int), ("dt_last", C.c_double)]


class GmresResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("breakdown", C.c_int),
                ("nonfinite", C.c_int), ("residual_linf", C.c_double),
                ("relative_residual_linf", C.c_double)]


class ImexResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("last_iterations", C.c_int),
                ("picard_solves", C.c_int), ("density_floor_count", C.c_int),
                ("guard_count", C.c_int), ("nonfinite_count", C.c_int),
                ("residual_linf", C.c_double), ("relative_residual_linf", C.c_double),
                ("rhs_norm", C.c_double), ("coupled_residual_linf", C.c_double),
                ("trial_density_min", C.c_double), ("trial_pressure_min", C.c_double),
                ("trial_internal_energy_min", C.c_double), ("solved_pressure_min", C.c_double),
                ("solved_pressure_max", C.c_double), ("momentum_correction_linf", C.c_double),
                ("energy_correction_linf", C.c_double), ("eq78_divergence_linf", C.c_double)]


IMEX_STATUS = {0: "converged_absolute", 1: "converged_relative", 2: "linear_solve_not_converged",
               3: "gmres_breakdown_not_converged", 4: "density_denominator_floor_used",
               5: "source_map_pressure_row_nonfinite", 6: "linear_residual_nonfinite",
               7: "source_map_lagged_momentum_nonfinite"}


def build() -> None:
    """Compile libec_oracle.so with make (no-op when up to date)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        G = C.POINTER(Grid)
        L.ecor_to_primitive.argtypes = [_dp, C.c_double, _dp]
        L.ecor_to_conserved.argtypes = [_dp, C.c_double, _dp]
        L.ecor_limited_slope.argtypes = [C.c_double, C.c_double, C.c_int]
        L.ecor_limited_slope.restype = C.c_double
        L.ecor_physical_flux.argtypes = [_dp, C.c_int, C.c_double, _dp]
        L.ecor_interface_flux.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_int, C.c_double, _dp]
        L.ecor_hllc_star_degenerate.argtypes = [_dp, _dp, C.c_int, C.c_double]
        L.ecor_hllc_star_degenerate.restype = C.c_int
        L.ecor_pressure_ratio.argtypes = [C.c_double, C.c_double]
        L.ecor_pressure_ratio.restype = C.c_double
        L.ecor_exact_riemann_star.argtypes = [_dp, _dp, C.c_double, _dp, _dp]
        L.ecor_exact_riemann_sample.argtypes = [C.c_double, C.c_double, _dp, _dp, C.c_double,
                                                C.c_double, _dp]
        L.ecor_fv_step.argtypes = [G, _dp, _dp, C.c_double]
        L.ecor_fv_step.restype = C.c_int
        L.ecor_face_flux.argtypes = [G, _dp, C.c_int, C.c_int, C.c_int, _dp]
        L.ecor_reconstruct.argtypes = [G, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.ecor_load.argtypes = [G, _dp, C.c_int, C.c_int, _dp]
        D = C.POINTER(Direct)
        L.ecor_eps_to_primitive.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.ecor_eps_to_conserved.argtypes = [_dp, C.c_double, C.c_double, _dp]
        L.ecor_direct_boundary_state.argtypes = [G, D, C.c_int, C.c_int, C.c_double, _dp]
        L.ecor_direct_step.argtypes = [G, D, _dp, C.c_double, C.c_double, C.c_int, C.c_double,
                                       C.c_int, _dp, C.POINTER(DirectResult)]
        L.ecor_direct_dt.argtypes = [G, D, _dp, C.c_double, C.c_double, C.c_double, C.c_double]
        L.ecor_direct_dt.restype = C.c_double
        L.ecor_admissibility.argtypes = [G, _dp, _dp, C.POINTER(C.c_int)]
        L.ecor_gresho_metrics.argtypes = [G, _dp] + [C.c_double] * 7 + [_dp]
        L.ecor_shock_bubble_metrics.argtypes = [G, _dp, C.c_double, C.c_double, C.c_double, _dp]
        L.ecor_cylindrical_source.argtypes = [G, _dp, C.c_double, C.c_double]
        L.ecor_exact_ghost.argtypes = [G, C.c_int, C.c_int]
        L.ecor_exact_ghost.restype = C.c_int
        L.ecor_exact_state.argtypes = [C.POINTER(Exact), C.c_int, C.c_int, C.c_double, C.c_int,
                                       C.c_int, _dp]
        L.ecor_max_rate.argtypes = [G, _dp]
        L.ecor_max_rate.restype = C.c_double
        L.ecor_euler_dt.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
        L.ecor_euler_dt.restype = C.c_double
        L.ecor_diagnostics.argtypes = [G, _dp, _dp, _dp, C.POINTER(C.c_int)]
        L.ecor_run_explicit.argtypes = [G, _dp, C.c_double, C.c_double, C.c_int, _dp,
                                        C.POINTER(RunResult)]
        L.ecor_init_gresho.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double, _dp]
        L.ecor_init_toro.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                                     _dp, C.c_double, _dp]
        _lp = C.POINTER(C.c_long)
        L.ecor_imex_star.argtypes = [G, _dp, C.c_double, C.c_int, _dp]
        L.ecor_imex_assemble.argtypes = [G, _dp, _dp, _dp, C.c_double, _lp, _lp, _dp, _dp]
        L.ecor_imex_assemble.restype = C.c_long
        L.ecor_gmres_full.argtypes = [C.c_long, _lp, _lp, _dp, _dp, _dp, C.c_int, C.c_double,
                                      _dp, C.POINTER(GmresResult)]
        L.ecor_imex_step.argtypes = [G, _dp, C.c_double, C.c_int, C.c_double, C.c_int, _dp,
                                     C.POINTER(ImexResult)]
        L.ecor_imex_dt.argtypes = [G, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                   C.c_double, _dp, _dp]
        L.ecor_imex_dt.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _vec(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def make_grid(nx, ny, dx, dy, gamma=1.4, riemann=HLLC, limiter=MINMOD, order=2,
              bc=(BC_PERIODIC,) * 4, fixed=None) -> Grid:
    g = Grid()
    g.nx, g.ny, g.dx, g.dy, g.gamma = nx, ny, dx, dy, gamma
    g.riemann, g.limiter, g.order = riemann, limiter, order
    for s in range(4):
        g.bc[s] = bc[s]
        if fixed is not None and fixed[s] is not None:
            for c in range(4):
                g.fixed[s][c] = float(fixed[s][c])
    return g


# ---- pointwise -------------------------------------------------------------
def to_primitive(u, gamma=1.4):
    u = _vec(u); q = np.zeros(4)
    lib().ecor_to_primitive(_ptr(u), gamma, _ptr(q)); return q


def to_conserved(q, gamma=1.4):
    q = _vec(q); u = np.zeros(4)
    lib().ecor_to_conserved(_ptr(q), gamma, _ptr(u)); return u


def limited_slope(b, f, limiter=MINMOD):
    return lib().ecor_limited_slope(float(b), float(f), int(limiter))


def physical_flux(q, axis, gamma=1.4):
    q = _vec(q); f = np.zeros(4)
    lib().ecor_physical_flux(_ptr(q), axis, gamma, _ptr(f)); return f


def interface_flux(qL, qR, axis, gamma=1.4, riemann=HLLC, sensor_f=1.0):
    qL = _vec(qL); qR = _vec(qR); f = np.zeros(4)
    lib().ecor_interface_flux(_ptr(qL), _ptr(qR), axis, gamma, riemann, sensor_f, _ptr(f))
    return f


def hllc_star_degenerate(qL, qR, axis, gamma=1.4):
    qL = _vec(qL); qR = _vec(qR)
    return bool(lib().ecor_hllc_star_degenerate(_ptr(qL), _ptr(qR), axis, gamma))


def pressure_ratio(pl, pr):
    return lib().ecor_pressure_ratio(float(pl), float(pr))


def exact_riemann_star(qL, qR, gamma=1.4):
    qL = _vec(qL); qR = _vec(qR); p = C.c_double(); u = C.c_double()
    lib().ecor_exact_riemann_star(_ptr(qL), _ptr(qR), gamma, C.byref(p), C.byref(u))
    return p.value, u.value


def exact_riemann_sample(x, t, qL, qR, gamma=1.4, x0=0.5):
    qL = _vec(qL); qR = _vec(qR); q = np.zeros(4)
    lib().ecor_exact_riemann_sample(float(x), float(t), _ptr(qL), _ptr(qR), gamma, x0, _ptr(q))
    return q


# ---- grid ------------------------------------------------------------------
def fv_step(g: Grid, U: np.ndarray, dt: float):
    U = np.ascontiguousarray(U, dtype=np.float64)
    out = np.empty_like(U)
    deg = lib().ecor_fv_step(C.byref(g), _ptr(U), _ptr(out), float(dt))
    return out, deg


def face_flux(g: Grid, U, fi, fj, axis):
    U = np.ascontiguousarray(U, dtype=np.float64); f = np.zeros(4)
    lib().ecor_face_flux(C.byref(g), _ptr(U), fi, fj, axis, _ptr(f)); return f


def load(g: Grid, U, i, j):
    U = np.ascontiguousarray(U, dtype=np.float64); u = np.zeros(4)
    lib().ecor_load(C.byref(g), _ptr(U), i, j, _ptr(u)); return u


def exact_ghost(g: Grid, i, j) -> bool:
    return bool(lib().ecor_exact_ghost(C.byref(g), i, j))


def exact_state(e: Exact, nx, ny, time, i, j) -> np.ndarray:
    u = np.zeros(4)
    lib().ecor_exact_state(C.byref(e), nx, ny, time, i, j, _ptr(u))
    return u


def make_direct(exact: Exact, epsilon=1.0, x_only_row=False) -> Direct:
    d = Direct()
    d.exact = exact
    d.epsilon = float(epsilon)
    d.x_only_row = int(bool(x_only_row))
    return d


def eps_to_conserved(q, gamma=1.4, eps=1.0):
    q = _vec(q); u = np.zeros(4)
    lib().ecor_eps_to_conserved(_ptr(q), gamma, eps, _ptr(u))
    return u


def eps_to_primitive(u, gamma=1.4, eps=1.0):
    u = _vec(u); q = np.zeros(4)
    lib().ecor_eps_to_primitive(_ptr(u), gamma, eps, _ptr(q))
    return q


def direct_step(g: Grid, d: Direct, U, time, dt, picard=4, tol=1e-10, max_iter=1000):
    """One direct BDLTV20 step on a copy of U (epsilon-scaled conserved);
    returns (U_new, pressure, result)."""
    U = np.array(U, dtype=np.float64, order="C", copy=True)
    p = np.zeros(g.nx * g.ny)
    res = DirectResult()
    lib().ecor_direct_step(C.byref(g), C.byref(d), _ptr(U), time, dt, picard, tol, max_iter,
                           _ptr(p), C.byref(res))
    return U, p.reshape(g.ny, g.nx), res


def direct_dt(g: Grid, d: Direct, U, imex_cfl, fixed_dt, time, stop_time):
    U = np.ascontiguousarray(U, dtype=np.float64)
    return lib().ecor_direct_dt(C.byref(g), C.byref(d), _ptr(U), imex_cfl, fixed_dt, time,
                                stop_time)


def direct_boundary_state(g: Grid, d: Direct, i, j, t):
    u = np.zeros(4)
    lib().ecor_direct_boundary_state(C.byref(g), C.byref(d), i, j, t, _ptr(u))
    return u


def admissibility(g: Grid, U):
    """(rho_min, p_min, ie_min), (nonfinite, density, pressure, ie failure counts)"""
    U = np.ascontiguousarray(U, dtype=np.float64)
    o = np.zeros(3); c = (C.c_int * 4)()
    lib().ecor_admissibility(C.byref(g), _ptr(U), _ptr(o), c)
    return tuple(o), tuple(c)


GRESHO_METRIC_KEYS = ("density_l1_error", "velocity_l1_error", "pressure_l1_error",
                      "pressure_perturbation_l1_error", "pressure_perturbation_l1_relative_error",
                      "kinetic_energy_initial", "kinetic_energy_final", "kinetic_energy_ratio",
                      "reference_sound_speed", "velocity_scale")


def gresho_metrics(g: Grid, U, xlo, ylo, cx, cy, rho, mach, p0) -> dict:
    U = np.ascontiguousarray(U, dtype=np.float64)
    o = np.zeros(10)
    lib().ecor_gresho_metrics(C.byref(g), _ptr(U), xlo, ylo, cx, cy, rho, mach, p0, _ptr(o))
    return dict(zip(GRESHO_METRIC_KEYS, o))


SB_METRIC_KEYS = ("rho_min", "rho_max", "pressure_min", "pressure_max", "internal_energy_min",
                  "density_bubble_area", "density_bubble_centroid_x", "density_bubble_centroid_y",
                  "schlieren_max", "nonfinite_count")


def shock_bubble_metrics(g: Grid, U, xlo, ylo, threshold=0.55) -> dict:
    U = np.ascontiguousarray(U, dtype=np.float64)
    o = np.zeros(10)
    lib().ecor_shock_bubble_metrics(C.byref(g), _ptr(U), xlo, ylo, threshold, _ptr(o))
    return dict(zip(SB_METRIC_KEYS, o))


def cylindrical_source(g: Grid, U, dt, ylo=0.0):
    """The cylindrical shock bubble's geometric source over dt (a new array)."""
    U = np.array(U, dtype=np.float64, order="C", copy=True)
    lib().ecor_cylindrical_source(C.byref(g), _ptr(U), dt, ylo)
    return U


def max_rate(g: Grid, U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    return lib().ecor_max_rate(C.byref(g), _ptr(U))


def euler_dt(max_rate_, cfl, time, stop_time):
    return lib().ecor_euler_dt(max_rate_, cfl, time, stop_time)


def diagnostics(g: Grid, U):
    U = np.ascontiguousarray(U, dtype=np.float64)
    r = C.c_double(); p = C.c_double(); n = C.c_int()
    lib().ecor_diagnostics(C.byref(g), _ptr(U), C.byref(r), C.byref(p), C.byref(n))
    return r.value, p.value, n.value


def run_explicit(g: Grid, U: np.ndarray, cfl: float, stop_time: float, max_step: int):
    """Run the explicit driver loop in place on a copy; returns (U, result, dts)."""
    U = np.array(U, dtype=np.float64, order="C", copy=True)
    dts = np.zeros(max(max_step, 1))
    res = RunResult()
    lib().ecor_run_explicit(C.byref(g), _ptr(U), cfl, stop_time, max_step, _ptr(dts),
                            C.byref(res))
    return U, res, dts[:res.steps].copy()


def init_gresho(nx, ny, xlo=0.0, ylo=0.0, dx=None, dy=None, gamma=1.4, cx=0.5, cy=0.5,
                rho=1.0, mach=0.01, p0=None):
    dx = 1.0 / nx if dx is None else dx
    dy = 1.0 / ny if dy is None else dy
    p0 = 1.0 / (gamma * mach * mach) if p0 is None else p0
    U = np.zeros((4, ny, nx))
    lib().ecor_init_gresho(nx, ny, xlo, ylo, dx, dy, gamma, cx, cy, rho, mach, p0, _ptr(U))
    return U


def init_toro(nx, ny, qL, qR, xlo=0.0, dx=None, gamma=1.4, x0=0.5):
    dx = 1.0 / nx if dx is None else dx
    qL = _vec(qL); qR = _vec(qR)
    U = np.zeros((4, ny, nx))
    lib().ecor_init_toro(nx, ny, xlo, dx, gamma, _ptr(qL), _ptr(qR), x0, _ptr(U))
    return U


def rel_l1(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Relative L1 per conserved variable: sum|a-b| / sum|b| (mean|diff|/mean|ref|)."""
    a = a.reshape(4, -1); b = b.reshape(4, -1)
    den = np.abs(b).sum(axis=1)
    den = np.where(den == 0.0, 1.0, den)
    return np.abs(a - b).sum(axis=1) / den


# ---- IMEX ------------------------------------------------------------------
def imex_star(g: Grid, U, dt, acoustic=False):
    U = np.ascontiguousarray(U, dtype=np.float64); S = np.empty_like(U)
    lib().ecor_imex_star(C.byref(g), _ptr(U), float(dt), int(acoustic), _ptr(S)); return S


def imex_assemble(g: Grid, star, p, mom, dt):
    """CSR (indptr, indices, data) of the pressure rows and the right-hand side."""
    n = g.nx * g.ny
    star = np.ascontiguousarray(star, dtype=np.float64)
    p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
    mom = np.ascontiguousarray(mom, dtype=np.float64)
    rp = np.zeros(n + 1, dtype=np.int64); cols = np.zeros(11 * n, dtype=np.int64)
    vals = np.zeros(11 * n); rhs = np.zeros(n)
    lp = C.POINTER(C.c_long)
    nnz = lib().ecor_imex_assemble(C.byref(g), _ptr(star), _ptr(p), _ptr(mom), float(dt),
                                   rp.ctypes.data_as(lp), cols.ctypes.data_as(lp), _ptr(vals),
                                   _ptr(rhs))
    return rp, cols[:nnz].copy(), vals[:nnz].copy(), rhs


def gmres_full(indptr, indices, data, b, x0, max_iter, tol):
    n = len(b)
    indptr = np.ascontiguousarray(indptr, dtype=np.int64)
    indices = np.ascontiguousarray(indices, dtype=np.int64)
    data = np.ascontiguousarray(data, dtype=np.float64)
    b = _vec(b); x0 = _vec(x0); x = np.zeros(n); r = GmresResult()
    lp = C.POINTER(C.c_long)
    lib().ecor_gmres_full(n, indptr.ctypes.data_as(lp), indices.ctypes.data_as(lp), _ptr(data),
                          _ptr(b), _ptr(x0), int(max_iter), float(tol), _ptr(x), C.byref(r))
    return x, r


def imex_step(g: Grid, U, dt, picard=2, tol=1e-8, max_iter=1000):
    """One IMEX step; returns (U_new, p, result).  U_new is U when the step fails."""
    U = np.array(U, dtype=np.float64, order="C", copy=True)
    p = np.zeros(g.nx * g.ny); r = ImexResult()
    lib().ecor_imex_step(C.byref(g), _ptr(U), float(dt), int(picard), float(tol), int(max_iter),
                         _ptr(p), C.byref(r))
    return U, p.reshape(g.ny, g.nx), r


def imex_dt(g: Grid, U, cfl, time, stop, acoustic_startup=0, acoustic_cap=0.0):
    U = np.ascontiguousarray(U, dtype=np.float64)
    m = C.c_double(); a = C.c_double()
    dt = lib().ecor_imex_dt(C.byref(g), _ptr(U), cfl, time, stop, acoustic_startup,
                            acoustic_cap, C.byref(m), C.byref(a))
    return dt, m.value, a.value
