"""Pins of the oracle's cylindrical geometric source
(apply_shock_density_bubble_cylindrical_source,
problems/boundary_conditions.hpp:75-131).

The source integrates dq/dt = -(rho v, rho u v, rho v^2, v (E + p)) / r for
one step.  That ODE has a closed form: u and v stay constant, the density
decays as rho0 exp(-v t / r) and the pressure as p0 exp(-gamma v t / r)
(an adiabatic expansion, p ~ rho^gamma).  The midpoint rule is second order,
so its one-step error against the closed form falls as dt^3: halving dt
divides it by 8.  A first-order stage (a term of the half step wrong or
dropped) falls as dt^2, and a wrong term in the full step as dt, so each
such mistake fails the ratio.  Rows at r <= 0 and states with v = 0 are left
alone.
"""
import numpy as np
import pytest

GAMMA = 1.4


def cons(r, u, v, p):
    return np.array([r, r * u, r * v, p / (GAMMA - 1) + 0.5 * r * (u * u + v * v)])


def closed_form(rho0, u, v, p0, r, t):
    a = v / r
    return cons(rho0 * np.exp(-a * t), u, v, p0 * np.exp(-GAMMA * a * t))


def one_cell_grid(ecor, ny, dy):
    return ecor.make_grid(1, ny, 1.0, dy, bc=(1, 1, 1, 1))


@pytest.mark.parametrize("u,v", [(0.7, 0.9), (-1.3, 0.4), (0.2, -0.6)])
def test_midpoint_rule_against_the_closed_form(ecor, u, v):
    ny, dy, ylo = 3, 0.1, 0.5      # r = 0.55, 0.65, 0.75
    rho0, p0 = 1.3, 2.1
    U = np.zeros((4, ny, 1))
    for j in range(ny):
        U[:, j, 0] = cons(rho0, u, v, p0)
    g = one_cell_grid(ecor, ny, dy)
    errs = []
    for dt in (0.02, 0.01, 0.005):
        got = ecor.cylindrical_source(g, U, dt, ylo)
        e = np.zeros((4, ny))
        for j in range(ny):
            r = ylo + (j + 0.5) * dy
            e[:, j] = np.abs(got[:, j, 0] - closed_form(rho0, u, v, p0, r, dt))
        errs.append(e)
    for c in range(4):
        for j in range(ny):
            assert errs[2][c, j] < 5e-6
            for k in (0, 1):
                ratio = errs[k][c, j] / errs[k + 1][c, j]
                assert 7.5 < ratio < 8.5, (c, j, k, ratio)


def test_rest_and_axis_rows_are_left_alone(ecor):
    ny, dy = 6, 0.1
    rng = np.random.default_rng(3)
    U = np.stack([rng.uniform(0.5, 2.0, (ny, 4)), rng.uniform(-1, 1, (ny, 4)),
                  np.zeros((ny, 4)), rng.uniform(3.0, 5.0, (ny, 4))])
    g = ecor.make_grid(4, ny, 0.1, dy, bc=(1, 1, 1, 1))
    assert np.array_equal(ecor.cylindrical_source(g, U, 0.05, 0.0), U)     # v = 0
    U[2] = rng.uniform(-1, 1, (ny, 4))
    # y_lo = -0.3: rows 0..2 have r = -0.25, -0.15, -0.05
    got = ecor.cylindrical_source(g, U, 0.05, -0.3)
    assert np.array_equal(got[:, :3], U[:, :3])
    assert not np.array_equal(got[:, 3:], U[:, 3:])


def test_nonpositive_density_reads_zero_velocity(ecor):
    """inv_rho = 0 where rho <= 0 (:100-102): the momentum and energy
    sources vanish and only the mass source -m_y / r acts."""
    g = one_cell_grid(ecor, 1, 0.2)
    U = np.array([-0.5, 0.3, 0.4, 2.0]).reshape(4, 1, 1)
    dt, r = 0.01, 0.1
    got = ecor.cylindrical_source(g, U, dt, 0.0)
    assert got[1, 0, 0] == 0.3 and got[2, 0, 0] == 0.4 and got[3, 0, 0] == 2.0
    assert got[0, 0, 0] == pytest.approx(-0.5 - dt * 0.4 / r, rel=1e-15)
