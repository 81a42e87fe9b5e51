"""Independent pins of the oracle's low-Mach HLLC-P flux, its 5-face pressure
sensor and the low-Mach HLLC velocity blend (reconstruction_fluxes.hpp:212-252,
254-371, 390-440 of the reference's euler_compare app).  The checks live in
tests/hllcp_pins.py; tests/test_oracle_mutants.py runs the same checks on
deliberately broken copies of the oracle to show that each one bites.
"""
import numpy as np
import pytest

import hllcp_pins as P

PAIRS = P.random_subsonic_pairs(24, seed=20261204)
LOW_MACH_PAIRS = [
    # Gresho-like: p ~ 1/(gamma M^2) at M = 0.01, O(1) velocities, O(1) jumps
    (np.array([1.0, 0.31, -0.42, 7142.857142857143]), np.array([1.0, 0.28, -0.37, 7141.9])),
    (np.array([1.2, 0.05, 0.01, 71.43]), np.array([1.0, -0.02, 0.03, 71.40])),
]


def oracle_flux(ecor):
    return lambda L, R, axis, kind, f: ecor.interface_flux(L, R, axis, P.G, kind, f)


@pytest.mark.parametrize("k", range(len(PAIRS)))
def test_roe_property_of_recovered_average(ecor, k):
    L, R = PAIRS[k]
    assert P.roe_property_error(oracle_flux(ecor), L, R) <= 1e-9


def test_roe_property_check_rejects_arithmetic_means():
    """The check itself: Roe's property fails for the arithmetic mean state."""
    L, R = PAIRS[0]
    u, v = 0.5 * (L[1] + R[1]), 0.5 * (L[2] + R[2])
    H = 0.5 * ((P.conserved(L)[3] + L[3]) / L[0] + (P.conserved(R)[3] + R[3]) / R[0])
    dU = P.conserved(R) - P.conserved(L)
    dF = P.euler_flux(R, 0) - P.euler_flux(L, 0)
    assert np.max(np.abs(P.flux_jacobian(u, v, H, 0) @ dU - dF)) / np.max(np.abs(dF)) > 1e-3


@pytest.mark.parametrize("k", range(len(PAIRS)))
def test_low_mach_pressure_correction(ecor, k):
    L, R = PAIRS[k]
    assert P.pressure_term_error(oracle_flux(ecor), L, R) <= 1e-9


@pytest.mark.parametrize("k", range(0, len(PAIRS), 3))
@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("f", [0.0, 0.27, 0.5, 0.9, 1.0])
def test_matches_paper_arrangement(ecor, k, axis, f):
    L, R = PAIRS[k]
    assert P.paper_form_error(oracle_flux(ecor), L, R, axis, f) <= 1e-12


@pytest.mark.parametrize("L,R", LOW_MACH_PAIRS)
@pytest.mark.parametrize("axis", [0, 1])
def test_matches_paper_arrangement_low_mach(ecor, L, R, axis):
    # The flux is dominated by the O(1/M^2) pressure; compare the O(1) part.
    got = ecor.interface_flux(L, R, axis, P.G, ecor.XIE_HLLCP, 0.3)
    want = P.paper_form(L, R, axis, 0.3)
    scale = max(abs(want[0]), 1.0)
    assert np.max(np.abs(got - want)) <= 1e-9 * scale + 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("order", [1, 2])
def test_sensor_footprint_and_cube(ecor, order):
    nx, ny = 7, 6
    Q, U = P.sensor_field(nx, ny, seed=11 + order)
    Qo = np.stack([ecor.to_primitive(U[:, j, i]) for j in range(ny) for i in range(nx)], axis=1)
    Qo = Qo.reshape(4, ny, nx)
    g = ecor.make_grid(nx, ny, 1.0 / nx, 1.0 / ny, riemann=ecor.XIE_HLLCP, order=order)
    err = P.sensor_errors(oracle_flux(ecor), lambda gr, UU, i, j, a: ecor.face_flux(gr, UU, i, j, a),
                          g, U, Qo, order)
    assert err <= 1e-9


def test_sensor_single_jump_is_cubed_ratio(ecor):
    """One cell at a lower pressure: faces whose 5-face stencil touches it get
    f = r^3, the rest f = 1 (here: flux unaffected, since L = R there)."""
    nx = ny = 6
    Q = np.zeros((4, ny, nx))
    Q[0] = 1.0
    Q[3] = 2.0
    Q[3, 3, 2] = 1.0                                    # r = 1/2
    Q[1, 3, 2] = 0.05
    U = np.stack([P.conserved(Q[:, j, i]) for j in range(ny) for i in range(nx)], axis=1)
    U = np.ascontiguousarray(U.reshape(4, ny, nx))
    g = ecor.make_grid(nx, ny, 1.0 / nx, 1.0 / ny, riemann=ecor.XIE_HLLCP, order=1)
    L, R = Q[:, 3, 2], Q[:, 3, 3]                       # x-face (3, 3): cells 2 | 3
    F = ecor.face_flux(g, U, 3, 3, 0)
    F0 = ecor.interface_flux(L, R, 0, P.G, ecor.XIE_HLLCP, 0.0)
    F1 = ecor.interface_flux(L, R, 0, P.G, ecor.XIE_HLLCP, 1.0)
    f = np.dot(F - F0, F1 - F0) / np.dot(F1 - F0, F1 - F0)
    assert f == pytest.approx(0.125, abs=1e-12)


def test_low_mach_hllc_velocity_blend(ecor):
    assert P.low_mach_blend_error(oracle_flux(ecor)) <= 1e-12
