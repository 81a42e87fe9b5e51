"""Pins of the oracle's pointwise functions against what the mathematics fixes.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, swapped operands) fails it; none retypes the
oracle's own formula.
"""
import math

import numpy as np
import pytest

G = 1.4


def conserved(q, g=G):
    r, u, v, p = q
    return np.array([r, r * u, r * v, p / (g - 1) + 0.5 * r * (u * u + v * v)])


def euler_flux(q, axis, g=G):
    r, u, v, p = q
    E = p / (g - 1) + 0.5 * r * (u * u + v * v)
    un = u if axis == 0 else v
    f = np.array([r * un, r * u * un, r * v * un, (E + p) * un])
    f[1 + axis] += p
    return f


def toro_hllc(L, R, axis, g=G):
    """HLLC in Toro's textbook form (Riemann Solvers, 3rd ed., eqs 10.38-10.40,
    10.73) with the Davis wave-speed estimates: an independent formula."""
    aL, aR = math.sqrt(g * L[3] / L[0]), math.sqrt(g * R[3] / R[0])
    uL, uR = L[1 + axis], R[1 + axis]
    SL, SR = min(uL - aL, uR - aR), max(uL + aL, uR + aR)
    if SL >= 0:
        return euler_flux(L, axis, g)
    if SR <= 0:
        return euler_flux(R, axis, g)
    Ss = (R[3] - L[3] + L[0] * uL * (SL - uL) - R[0] * uR * (SR - uR)) / (
        L[0] * (SL - uL) - R[0] * (SR - uR))
    K, SK, uK = (L, SL, uL) if Ss >= 0 else (R, SR, uR)
    UK = conserved(K, g)
    fac = K[0] * (SK - uK) / (SK - Ss)
    Us = fac * np.array([1.0, K[1], K[2], UK[3] / K[0] + (Ss - uK) * (Ss + K[3] / (K[0] * (SK - uK)))])
    Us[1 + axis] = fac * Ss
    return euler_flux(K, axis, g) + SK * (Us - UK)


def rot(q):
    """Rotate a primitive state by +90 degrees: (u, v) -> (-v, u)."""
    return np.array([q[0], -q[2], q[1], q[3]])


STATES = [
    ([1.0, 0.3, -0.2, 1.0], [0.5, -0.1, 0.4, 0.6]),
    ([1.2, 0.05, 0.01, 71.43], [1.0, -0.02, 0.03, 71.40]),   # low Mach
    ([0.3, 0.9, 0.0, 0.2], [2.0, 0.0, 0.5, 3.0]),
    ([1.0, 1.5, 0.3, 1.0], [0.8, 0.2, -0.1, 0.9]),           # one side supersonic
]


def test_primitive_conserved_roundtrip(ecor):
    q = np.array([1.3, 0.7, -0.4, 2.5])
    u = ecor.to_conserved(q)
    assert np.allclose(u, conserved(q), rtol=1e-15, atol=0)
    assert np.allclose(ecor.to_primitive(u), q, rtol=1e-14, atol=1e-15)
    # p = (gamma-1)(E - rho|u|^2/2) with another gamma
    assert ecor.to_primitive([2.0, 2.0, 0.0, 10.0], 5.0 / 3.0)[3] == pytest.approx(2.0 / 3.0 * 9.0)


@pytest.mark.parametrize("b,f,mm,mc,vl", [
    (1.0, 2.0, 1.0, 1.5, 4.0 / 3.0),
    (-3.0, -1.0, -1.0, -2.0, -1.5),
    (1.0, -2.0, 0.0, 0.0, 0.0),
    (0.0, 5.0, 0.0, 0.0, 0.0),
    (1.0, 10.0, 1.0, 2.0, 20.0 / 11.0),
])
def test_limiters_closed_form(ecor, b, f, mm, mc, vl):
    assert ecor.limited_slope(b, f, ecor.MINMOD) == pytest.approx(mm, abs=0)
    assert ecor.limited_slope(b, f, ecor.MC) == pytest.approx(mc, rel=1e-15)
    assert ecor.limited_slope(b, f, ecor.VANLEER) == pytest.approx(vl, rel=1e-15)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("axis", [0, 1])
def test_consistency_equal_states(ecor, kind, axis):
    q = np.array([0.9, 0.2, -0.3, 1.7])
    f = ecor.interface_flux(q, q, axis, G, kind, 0.4)
    assert np.allclose(f, euler_flux(q, axis), rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("L,R", STATES)
@pytest.mark.parametrize("axis", [0, 1])
def test_hllc_matches_toro_textbook_form(ecor, L, R, axis):
    f = ecor.interface_flux(L, R, axis, G, ecor.HLLC)
    assert np.allclose(f, toro_hllc(np.array(L), np.array(R), axis), rtol=1e-12, atol=1e-13)


def test_hllc_supersonic_upwinding(ecor):
    L = np.array([1.0, 3.0, 0.1, 1.0])   # u - c > 0
    R = np.array([0.5, 2.9, 0.0, 0.8])
    assert np.array_equal(ecor.interface_flux(L, R, 0, G, ecor.HLLC), ecor.physical_flux(L, 0))
    Lm = np.array([1.0, -3.0, 0.1, 1.0])
    Rm = np.array([0.5, -3.1, 0.0, 0.8])
    assert np.array_equal(ecor.interface_flux(Lm, Rm, 0, G, ecor.HLLC), ecor.physical_flux(Rm, 0))


@pytest.mark.parametrize("kind", [1, 3])
def test_stationary_contact_is_exact(ecor, kind):
    L = [1.0, 0.0, 0.0, 1.0]
    R = [0.125, 0.0, 0.0, 1.0]
    f = ecor.interface_flux(L, R, 0, G, kind, 1.0)
    assert np.allclose(f, [0.0, 1.0, 0.0, 0.0], atol=1e-15)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("L,R", STATES)
def test_rotation_symmetry(ecor, kind, L, R):
    """An x-face of (L,R) rotated by 90 degrees is a y-face of rotated states."""
    L, R = np.array(L), np.array(R)
    fx = ecor.interface_flux(L, R, 0, G, kind, 0.3)
    fy = ecor.interface_flux(rot(L), rot(R), 1, G, kind, 0.3)
    assert np.allclose(fy, [fx[0], -fx[2], fx[1], fx[3]], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("L,R", STATES)
def test_mirror_symmetry(ecor, kind, L, R):
    """x -> -x swaps the sides and flips the normal velocity and flux."""
    L, R = np.array(L), np.array(R)
    m = np.array([1.0, -1.0, 1.0, 1.0])
    f = ecor.interface_flux(L, R, 0, G, kind, 0.6)
    g = ecor.interface_flux(R * m, L * m, 0, G, kind, 0.6)
    assert np.allclose(g, [-f[0], f[1], -f[2], -f[3]], rtol=1e-12, atol=1e-13)


def test_hllcp_reduces_to_hllc_when_locally_supersonic(ecor):
    """f = 1 and theta = clamp01(max local Mach) = 1 turn every low-Mach
    correction off; the star-flux form then equals HLLC by Rankine-Hugoniot."""
    L, R = STATES[3]
    for axis in (0,):
        a = ecor.interface_flux(L, R, axis, G, ecor.XIE_HLLCP, 1.0)
        b = ecor.interface_flux(L, R, axis, G, ecor.HLLC)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-13)


def test_hllcp_is_affine_in_sensor(ecor):
    L, R = STATES[1]
    f0 = ecor.interface_flux(L, R, 0, G, ecor.XIE_HLLCP, 0.0)
    f1 = ecor.interface_flux(L, R, 0, G, ecor.XIE_HLLCP, 1.0)
    fh = ecor.interface_flux(L, R, 0, G, ecor.XIE_HLLCP, 0.5)
    assert np.allclose(fh, 0.5 * (f0 + f1), rtol=1e-12, atol=1e-12)
    # the sensor is clamped to [0,1]
    assert np.array_equal(ecor.interface_flux(L, R, 0, G, ecor.XIE_HLLCP, 3.0), f1)


def test_low_mach_hllc_centres_velocity(ecor):
    """At zero local Mach number the corrected states share the mean velocity,
    so the flux equals HLLC of the velocity-averaged states."""
    L = np.array([1.0, 1e-9, 0.0, 1.0e6])
    R = np.array([1.0, -1e-9, 0.0, 1.0e6])
    avg = np.array([1.0, 0.0, 0.0, 1.0e6])
    a = ecor.interface_flux(L, R, 0, G, ecor.LOWMACH_HLLC)
    b = ecor.interface_flux(avg, avg, 0, G, ecor.HLLC)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-6)


def test_pressure_ratio(ecor):
    assert ecor.pressure_ratio(2.0, 1.0) == 0.5
    assert ecor.pressure_ratio(1.0, 4.0) == 0.25
    assert ecor.pressure_ratio(-1.0, 1.0) == 1.0e-14
    assert ecor.pressure_ratio(float("nan"), 1.0e-14) == 1.0


def test_degenerate_star_detection(ecor):
    assert not ecor.hllc_star_degenerate([1, 0, 0, 1], [0.125, 0, 0, 0.1], 0)
    # s_l == s_r is degenerate (vacuum-like zero sound speed)
    assert ecor.hllc_star_degenerate([1, 0, 0, 0.0], [1, 0, 0, 0.0], 0)


# Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics, Table 4.2.
TORO_STAR = [
    ((1.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.1), 0.30313, 0.92745),
    ((1.0, -2.0, 0.0, 0.4), (1.0, 2.0, 0.0, 0.4), 0.00189, 0.00000),
    ((1.0, 0.0, 0.0, 1000.0), (1.0, 0.0, 0.0, 0.01), 460.894, 19.5975),
    ((1.0, 0.0, 0.0, 0.01), (1.0, 0.0, 0.0, 100.0), 46.0950, -6.19633),
    ((5.99924, 19.5975, 0.0, 460.894), (5.99242, -6.19633, 0.0, 46.0950), 1691.64, 8.68975),
]


@pytest.mark.parametrize("L,R,p,u", TORO_STAR)
def test_exact_riemann_star_matches_toro_table(ecor, L, R, p, u):
    ps, us = ecor.exact_riemann_star(L, R)
    assert ps == pytest.approx(p, rel=6e-5, abs=6e-6)
    assert us == pytest.approx(u, rel=6e-5, abs=6e-6)


def test_exact_riemann_sample_sod(ecor):
    """Sod at t=0.2: left state, fan, post-contact plateaus (Toro 4.3.3)."""
    L, R = (1.0, 0.0, 0.0, 1.0), (0.125, 0.0, 0.0, 0.1)
    assert np.allclose(ecor.exact_riemann_sample(0.1, 0.2, L, R), L)
    q = ecor.exact_riemann_sample(0.75, 0.2, L, R)     # between contact and shock
    assert q[0] == pytest.approx(0.26557, rel=1e-4) and q[3] == pytest.approx(0.30313, rel=1e-4)
    q = ecor.exact_riemann_sample(0.55, 0.2, L, R)     # between fan tail and contact
    assert q[0] == pytest.approx(0.42632, rel=1e-4)
    assert np.allclose(ecor.exact_riemann_sample(0.95, 0.2, L, R), R)
