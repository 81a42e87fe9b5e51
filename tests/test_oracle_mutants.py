"""The HLLC-P pins bite: each test here compiles a copy of the oracle with one
plausible slip planted in its HLLC-P flux, sensor or low-Mach HLLC blend (in
a temporary directory, on the CPU) and requires that at least one check of
tests/hllcp_pins.py rejects it, while the unmodified copy passes them all.
"""
import ctypes as C
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import hllcp_pins as P
from oracle_copy import build_copy, load

# name -> list of (text in ec_oracle.c, replacement); each text occurs once.
MUTANTS = {
    "control": [],
    "roe_means_arithmetic": [("const double sqrt_rho_l = sqrt(rho_l);", "const double sqrt_rho_l = 1.0;"),
                             ("const double sqrt_rho_r = sqrt(rho_r);", "const double sqrt_rho_r = 1.0;")],
    "h_hat_arithmetic": [("(sqrt_rho_l * h_l + sqrt_rho_r * h_r) / roe_denom", "0.5 * (h_l + h_r)")],
    "mach_factor_dropped": [("(c_hat_sq * (1.0 + fabs(m_hat)))", "(c_hat_sq)")],
    "beta_sign_flipped": [("(ff - 1.0) * s_l * s_r", "(1.0 - ff) * s_l * s_r")],
    "c_hat_sq_term_dropped": [("(h_hat - 0.5 * (u_hat * u_hat + v_hat * v_hat))",
                               "(h_hat - 0.5 * (u_hat * u_hat))")],
    "phi_energy_half_dropped": [("0.5 * beta_p * (u_hat * u_hat + v_hat * v_hat)",
                                 "beta_p * (u_hat * u_hat + v_hat * v_hat)")],
    "theta_normal_mach": [("clamp01(smax(speed_l / smax(c_l", "clamp01(smax(fabs(q_l) / smax(c_l")],
    "theta_blend_swapped": [("theta * p_star + (1.0 - theta) * 0.5 * (p_l + p_r)",
                             "(1.0 - theta) * p_star + theta * 0.5 * (p_l + p_r)")],
    "sensor_blend_swapped": [("ff * pss + (1.0 - ff) * p_star", "ff * p_star + (1.0 - ff) * pss")],
    "sensor_not_cubed": [("return c * c * c;", "return c;")],
    "sensor_wrong_face": [("y_face_ratio(g, U, fi - 1, fj + 1)", "y_face_ratio(g, U, fi - 1, fj - 1)")],
    "low_mach_blend_full": [("cl[1] = u_avg + 0.5 * z * (L[1] - R[1]);",
                             "cl[1] = u_avg + 0.5 * (L[1] - R[1]);")],
}


@pytest.fixture(scope="module")
def libs(tmp_path_factory, ecor):
    root = tmp_path_factory.mktemp("mutants")
    with ThreadPoolExecutor(max_workers=4) as ex:
        sos = dict(zip(MUTANTS, ex.map(lambda kv: build_copy(root, *kv), MUTANTS.items())))
    return sos


def _bind(so, ecor):
    L = load(so, ecor)
    dp = C.POINTER(C.c_double)

    def ptr(a):
        return a.ctypes.data_as(dp)

    def flux(qL, qR, axis, kind, f):
        qL = np.ascontiguousarray(qL, dtype=np.float64)
        qR = np.ascontiguousarray(qR, dtype=np.float64)
        out = np.zeros(4)
        L.ecor_interface_flux(ptr(qL), ptr(qR), axis, P.G, kind, f, ptr(out))
        return out

    def face_flux(g, U, i, j, axis):
        out = np.zeros(4)
        L.ecor_face_flux(C.byref(g), ptr(U), i, j, axis, ptr(out))
        return out

    return flux, face_flux


def _failing_checks(flux, face_flux, ecor):
    pairs = P.random_subsonic_pairs(6, seed=7)
    fails = []
    if max(P.roe_property_error(flux, L, R) for L, R in pairs) > 1e-9:
        fails.append("roe_property")
    if max(P.pressure_term_error(flux, L, R) for L, R in pairs) > 1e-9:
        fails.append("pressure_term")
    if max(P.paper_form_error(flux, L, R, a, f) for L, R in pairs for a in (0, 1)
           for f in (0.27, 0.9)) > 1e-12:
        fails.append("paper_form")
    for order in (1, 2):
        nx, ny = 5, 4
        _, U = P.sensor_field(nx, ny, seed=3 + order)
        Q = np.stack([ecor.to_primitive(U[:, j, i]) for j in range(ny) for i in range(nx)], axis=1)
        g = ecor.make_grid(nx, ny, 1.0 / nx, 1.0 / ny, riemann=ecor.XIE_HLLCP, order=order)
        if P.sensor_errors(flux, face_flux, g, U, Q.reshape(4, ny, nx), order) > 1e-9:
            fails.append(f"sensor_o{order}")
    if P.low_mach_blend_error(flux) > 1e-12:
        fails.append("low_mach_blend")
    return fails


@pytest.mark.parametrize("name", list(MUTANTS))
def test_mutant_is_rejected(libs, ecor, name):
    fails = _failing_checks(*_bind(libs[name], ecor), ecor)
    if name == "control":
        assert fails == []
    else:
        assert fails, f"no check rejects the '{name}' mutant"
