"""ctypes binding of the C-ABI in include/ec_amd.h (libec_amd.so).

The library is built in-tree by ``build.py``; this module never builds and
never falls back: a missing library or a failing call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent
# EC_AMD_LIB: another build of the same sources (the sanitizer build)
LIB_PATH = Path(os.environ.get("EC_AMD_LIB", PKG_DIR / "libec_amd.so"))

EC_OK = 0
STATUS_TEXT = {0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "out of device memory",
               4: "unsupported", 5: "grid exceeds the fixed workspace of the allocation-free entry point"}
EC_ERR_TOO_LARGE = 5

# euler.riemann spellings of core/config_parse_helpers.hpp:44-61
RIEMANN = {"rusanov": 0, "local_lax_friedrichs": 0, "llf": 0, "hllc": 1, "lowmach_hllc": 2,
           "lm_hllc": 2, "low_mach_hllc": 2, "xie_am_hllc_p": 3, "am_hllc_p": 3,
           "xie_hllc_p": 3}
LIMITER = {"minmod": 0, "mc": 1, "monotonized_central": 1, "van_leer": 2, "vanleer": 2}
BC = {"periodic": 0, "outflow": 1, "reflect": 2, "fixed": 3, "exact": 4}

RUN_STATUS = {0: "ok", 1: "initial_nonphysical_state", 2: "invalid_time_step",
              3: "nonfinite_state", 4: "nonpositive_density", 5: "nonpositive_pressure"}


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("gamma", C.c_double), ("riemann", C.c_int), ("limiter", C.c_int),
                ("spatial_order", C.c_int), ("bc", C.c_int * 4),
                ("fixed_state", (C.c_double * 4) * 4)]


class Layout(C.Structure):
    _fields_ = [("row_stride", C.c_longlong), ("comp_stride", C.c_longlong)]


class StepDiag(C.Structure):
    _fields_ = [("rho_min", C.c_double), ("pressure_min", C.c_double),
                ("nonfinite_count", C.c_int), ("degenerate_count", C.c_int),
                ("max_rate", C.c_double)]


class RunState(C.Structure):
    _fields_ = [("time", C.c_double), ("stop_time", C.c_double), ("cfl", C.c_double),
                ("dt_last", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
                ("dt_sum", C.c_double), ("rho_min", C.c_double), ("pressure_min", C.c_double),
                ("max_rate", C.c_double), ("degenerate_total", C.c_longlong),
                ("step", C.c_int), ("max_step", C.c_int), ("nonfinite_count", C.c_int),
                ("status", C.c_int)]


class EcError(RuntimeError):
    pass


_lib = None
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)


def lib() -> C.CDLL:
    """Load libec_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise EcError(f"{LIB_PATH} is missing: build it with build.py (no fallback path exists)")
    L = C.CDLL(str(LIB_PATH))
    st = C.c_int
    L.ec_create.argtypes = [C.POINTER(GridDesc), _vp, C.POINTER(_vp)]
    L.ec_create.restype = st
    L.ec_destroy.argtypes = [_vp]
    L.ec_destroy.restype = st
    L.ec_status_string.argtypes = [st]
    L.ec_status_string.restype = C.c_char_p
    L.ec_last_error.argtypes = [_vp]
    L.ec_last_error.restype = C.c_char_p
    L.ec_alloc.argtypes = [_vp, C.c_size_t, C.POINTER(_vp)]
    L.ec_alloc.restype = st
    L.ec_free.argtypes = [_vp, _vp]
    L.ec_free.restype = st
    L.ec_synchronize.argtypes = [_vp]
    L.ec_synchronize.restype = st
    L.ec_explicit_step.argtypes = [_vp, _vp, _vp, C.POINTER(Layout), C.c_double, C.POINTER(StepDiag)]
    L.ec_explicit_step.restype = st
    L.ec_explicit_step_noalloc.argtypes = [C.POINTER(GridDesc), _vp, _vp, C.POINTER(Layout),
                                           C.c_double, C.POINTER(StepDiag), _vp]
    L.ec_explicit_step_noalloc.restype = st
    L.ec_explicit_step_inplace.argtypes = [_vp, _vp, C.POINTER(Layout), C.c_double,
                                           C.POINTER(StepDiag)]
    L.ec_explicit_step_inplace.restype = st
    L.ec_state_diagnostics.argtypes = [_vp, _vp, C.POINTER(Layout), C.POINTER(StepDiag)]
    L.ec_state_diagnostics.restype = st
    L.ec_euler_dt.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
    L.ec_euler_dt.restype = C.c_double
    L.ec_explicit_advance.argtypes = [_vp, _vp, _vp, C.POINTER(Layout), C.c_int,
                                      C.POINTER(RunState), C.POINTER(C.c_int)]
    L.ec_explicit_advance.restype = st
    L.ec_explicit_tile_count.argtypes = []
    L.ec_explicit_tile_count.restype = C.c_int
    L.ec_explicit_tile_shape.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int)]
    L.ec_explicit_tile_shape.restype = st
    L.ec_set_explicit_tile.argtypes = [_vp, C.c_int]
    L.ec_set_explicit_tile.restype = st
    L.ec_step_kernel_timing.argtypes = [_vp, C.c_int, C.c_int]
    L.ec_step_kernel_timing.restype = st
    L.ec_step_kernel_time.argtypes = [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.ec_step_kernel_time.restype = st
    _declare_optional(L)
    _lib = L
    return L


def _declare_optional(L: C.CDLL) -> None:
    """Signatures of the IMEX, adjoint and batch entry points."""
    from . import _abi_ext
    _abi_ext.declare(L)


def check(ctx, status: int, what: str) -> None:
    if status != EC_OK:
        detail = ""
        if ctx:
            detail = (lib().ec_last_error(ctx) or b"").decode()
        raise EcError(f"{what}: {STATUS_TEXT.get(status, status)} {detail}".strip())


def exported_symbols(header: Path) -> list[str]:
    """Function names declared in a C header of this package."""
    import re
    text = header.read_text()
    return sorted(set(re.findall(r"\b(ec_[a-z0-9_]+)\s*\(", text)))
