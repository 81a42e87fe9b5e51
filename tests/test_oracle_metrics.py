"""Pins of the oracle's whole-field metrics.

* Admissibility (numerics/diagnostic_metrics_timestep.hpp:49-102): hand-made
  cells hit each branch -- a NaN cell, rho <= 0, E below the kinetic energy
  (pressure and internal energy both fail) -- with the counts and minima the
  definitions fix.
* Gresho metrics (:147-225): on the exact initial state every error is at
  rounding level, the kinetic-energy ratio is 1, and the kinetic energy
  converges to the closed form pi rho V^2 (0.01 + 1/60) under refinement; a
  uniform pressure offset dp gives pressure errors of exactly dp.
* Shock-bubble metrics (io/output_cases.hpp:329-407): a uniform field has a
  zero schlieren maximum; a disc of low density has its centroid at the disc
  centre (by symmetry) and an area converging to pi r^2; a linear density
  ramp has the schlieren value |grad rho| of the ramp.
"""
import math

import numpy as np
import pytest

GAMMA = 1.4


def cons(r, u, v, p):
    return np.stack([r, r * u, r * v, p / (GAMMA - 1) + 0.5 * r * (u * u + v * v)])


def test_admissibility_branches(ecor):
    nx, ny = 5, 2
    U = cons(*(np.full((ny, nx), x) for x in (1.0, 0.5, 0.0, 2.0)))
    U[:, 0, 1] = [np.nan, 0.0, 0.0, 1.0]          # non-finite
    U[:, 0, 2] = [-0.5, 0.0, 0.0, 1.0]            # density
    U[:, 1, 3] = [1.0, 2.0, 0.0, 1.0]             # E = 1 < kinetic 2
    U[:, 1, 4] = [0.25, 0.0, 0.0, 1.0]            # admissible, smallest density
    g = ecor.make_grid(nx, ny, 0.1, 0.1)
    (rmin, pmin, emin), counts = ecor.admissibility(g, U)
    assert counts == (1, 1, 1, 1)
    assert rmin == -np.inf and pmin == -np.inf and emin == -np.inf
    ok = U.copy(); ok[:, 0, 1] = ok[:, 0, 0]; ok[:, 0, 2] = ok[:, 0, 0]
    (rmin, pmin, emin), counts = ecor.admissibility(g, ok)
    assert counts == (0, 0, 1, 1)
    assert rmin == 0.25 and pmin == pytest.approx(0.4 * (1.0 - 2.0)) and emin == pytest.approx(-1.0)


def gresho_state(ecor, n, mach=0.1):
    p0 = 1.0 / (GAMMA * mach * mach)
    U = ecor.init_gresho(n, n, mach=mach, p0=p0)
    return U, p0


def test_gresho_metrics_on_the_exact_state(ecor):
    kes = []
    for n in (32, 64, 128):
        U, p0 = gresho_state(ecor, n)
        g = ecor.make_grid(n, n, 1 / n, 1 / n)
        m = ecor.gresho_metrics(g, U, 0.0, 0.0, 0.5, 0.5, 1.0, 0.1, p0)
        assert m["density_l1_error"] == 0.0
        assert m["velocity_l1_error"] < 1e-15 and m["pressure_l1_error"] < 1e-12
        assert m["kinetic_energy_ratio"] == pytest.approx(1.0, abs=1e-13)
        assert m["velocity_scale"] == pytest.approx(1.0, rel=1e-15)
        assert m["reference_sound_speed"] == pytest.approx(10.0, rel=1e-15)
        kes.append(m["kinetic_energy_initial"])
    exact = math.pi * (0.01 + 1.0 / 60.0)
    errs = [abs(k - exact) for k in kes]
    assert errs[2] < 2e-4 * exact and errs[0] / errs[2] > 6.0


def test_gresho_pressure_offset(ecor):
    n = 32
    U, p0 = gresho_state(ecor, n)
    dp = 0.125
    U2 = U.copy(); U2[3] += dp / (GAMMA - 1)
    g = ecor.make_grid(n, n, 1 / n, 1 / n)
    m = ecor.gresho_metrics(g, U2, 0.0, 0.0, 0.5, 0.5, 1.0, 0.1, p0)
    assert m["pressure_l1_error"] == pytest.approx(dp, rel=1e-10)
    assert m["pressure_perturbation_l1_error"] == pytest.approx(dp, rel=1e-10)
    # the relative form divides by sum |p_exact - p0| over the cells
    p_exact = (GAMMA - 1) * (U[3] - 0.5 * (U[1] ** 2 + U[2] ** 2) / U[0])
    S = np.abs(p_exact - p0).sum()
    assert m["pressure_perturbation_l1_relative_error"] == pytest.approx(dp * n * n / S, rel=1e-9)


def sb_grid(ecor, nx, ny, dx):
    fixed = cons(2.0, 0.5, 0.0, 3.0)
    return ecor.make_grid(nx, ny, dx, dx, bc=(3, 1, 2, 1), fixed=(list(fixed), None, None, None))


def test_shock_bubble_metrics_closed_forms(ecor):
    nx, ny, dx = 80, 40, 1 / 40
    g = sb_grid(ecor, nx, ny, dx)
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    # uniform (the fixed left state equals it): no gradient, no bubble
    U = cons(*(np.broadcast_to(np.array(q), (ny, nx)) for q in (2.0, 0.5, 0.0, 3.0)))
    m = ecor.shock_bubble_metrics(g, U, 0.0, 0.0)
    assert m["schlieren_max"] == 0.0 and m["density_bubble_area"] == 0.0
    assert math.isnan(m["density_bubble_centroid_x"])
    # a disc of density 0.1, radius 0.3, centred on a cell corner at (1.0, 0.5)
    rho = np.where((X - 1.0) ** 2 + (Y - 0.5) ** 2 <= 0.09, 0.1, 1.0)
    U = cons(rho, np.zeros_like(rho), np.zeros_like(rho), np.ones_like(rho))
    m = ecor.shock_bubble_metrics(g, U, 0.0, 0.0)
    assert m["density_bubble_centroid_x"] == pytest.approx(1.0, abs=1e-13)
    assert m["density_bubble_centroid_y"] == pytest.approx(0.5, abs=1e-13)
    assert m["density_bubble_area"] == pytest.approx(math.pi * 0.09, rel=0.03)
    assert m["rho_min"] == 0.1 and m["rho_max"] == 1.0 and m["nonfinite_count"] == 0
    # a density ramp 1 + 0.5 x + 0.25 y away from the boundaries
    rho = 1.0 + 0.5 * X + 0.25 * Y
    U = cons(rho, np.zeros_like(rho), np.zeros_like(rho), np.ones_like(rho))
    # (outflow sides: clamped neighbours give half the slope at the edges)
    g2 = ecor.make_grid(nx, ny, dx, dx, bc=(1, 1, 1, 1))
    m = ecor.shock_bubble_metrics(g2, U, 0.0, 0.0)
    assert m["schlieren_max"] == pytest.approx(math.hypot(0.5, 0.25), rel=1e-12)
