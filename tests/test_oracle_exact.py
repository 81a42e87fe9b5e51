"""Pins of the oracle's exact-Dirichlet ghost states (euler.field_boundary=
exact_dirichlet).

* The exact Riemann solution the Toro ghosts sample: the Rankine-Hugoniot
  conditions across each shock (mass, momentum, energy), and across each
  rarefaction fan isentropy, the Riemann invariant and the characteristic
  u -/+ a = x/t; continuity at the fan's head and tail.
* The Gresho state: radial momentum balance dp/dr = rho v^2 / r.
* The quadrant states: the Lax-Liu configuration-3 table and its x<->y
  mirror symmetry.
* Placement and time: one step of an nx x ny grid with exact ghosts equals,
  on its cells, one step of the grid grown by two cells on every side whose
  extra cells hold the exact states (dyadic cell sizes make the two grids'
  cell centres identical, so the comparison is bitwise).  At t=0 the Toro
  ghosts are the initial left/right states, so the first step equals the
  outflow step.
"""
import numpy as np
import pytest

GAMMA = 1.4
TORO = {1: ([1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1]),
        2: ([1.0, -2.0, 0.0, 0.4], [1.0, 2.0, 0.0, 0.4]),
        3: ([1.0, 0.0, 0.0, 1000.0], [1.0, 0.0, 0.0, 0.01]),
        4: ([1.0, 0.0, 0.0, 0.01], [1.0, 0.0, 0.0, 100.0])}


def cons(q):
    r, u, v, p = q
    return np.array([r, r * u, r * v, p / (GAMMA - 1) + 0.5 * r * (u * u + v * v)])


def flux(q):
    r, u, v, p = q
    E = cons(q)[3]
    return np.array([r * u, r * u * u + p, r * u * v, (E + p) * u])


def sample(ecor, L, R, xi):
    return ecor.exact_riemann_sample(0.5 + xi, 1.0, L, R, x0=0.5)


def find_jump(ecor, L, R, a, b):
    """The xi of the density jump between a (one side) and b (other side)."""
    fa = sample(ecor, L, R, a)[0]
    for _ in range(200):
        m = 0.5 * (a + b)
        if sample(ecor, L, R, m)[0] == fa:
            a = m
        else:
            b = m
    return 0.5 * (a + b)


@pytest.mark.parametrize("test", [1, 3, 4])
def test_shock_rankine_hugoniot(ecor, test):
    L, R = TORO[test]
    ps, us = ecor.exact_riemann_star(L, R)
    right_shock = ps > R[3]
    side, far = (R, 50.0) if right_shock else (L, -50.0)
    S = find_jump(ecor, L, R, far, us + (1e-9 if right_shock else -1e-9))
    eps = 1e-7
    behind = sample(ecor, L, R, S - eps if right_shock else S + eps)
    ahead = sample(ecor, L, R, S + eps if right_shock else S - eps)
    assert np.allclose(ahead, side, rtol=0, atol=0)
    assert behind[1] == pytest.approx(us, rel=1e-12) and behind[3] == pytest.approx(ps, rel=1e-12)
    # F(behind) - S U(behind) = F(ahead) - S U(ahead), each component
    jb = flux(behind) - S * cons(behind)
    ja = flux(ahead) - S * cons(ahead)
    scale = np.abs(flux(ahead)).max() + abs(S) * np.abs(cons(ahead)).max() + 1.0
    assert np.abs(jb - ja).max() < 1e-7 * scale


@pytest.mark.parametrize("test", [1, 2, 3, 4])
def test_rarefaction_fan_invariants(ecor, test):
    L, R = TORO[test]
    ps, us = ecor.exact_riemann_star(L, R)
    for q, sgn in ((L, -1), (R, +1)):
        if ps > q[3]:
            continue  # shock on this side
        a0 = np.sqrt(GAMMA * q[3] / q[0])
        head = q[1] + sgn * a0
        a_star = a0 * (ps / q[3]) ** ((GAMMA - 1) / (2 * GAMMA))
        tail = us + sgn * a_star
        lo, hi = sorted((head, tail))
        for xi in np.linspace(lo, hi, 9)[1:-1]:
            r, u, _, p = sample(ecor, L, R, xi)
            a = np.sqrt(GAMMA * p / r)
            assert p / r ** GAMMA == pytest.approx(q[3] / q[0] ** GAMMA, rel=1e-12)
            assert u - sgn * 2 * a / (GAMMA - 1) == pytest.approx(
                q[1] - sgn * 2 * a0 / (GAMMA - 1), rel=1e-12, abs=1e-12)
            assert u + sgn * a == pytest.approx(xi, rel=1e-12, abs=1e-12)
        # continuous at the head and the tail
        e = 1e-9
        assert np.allclose(sample(ecor, L, R, head - sgn * e), q, rtol=1e-7, atol=1e-7)
        assert np.allclose(sample(ecor, L, R, head + sgn * e), q, rtol=1e-7, atol=1e-7)
        t_star = sample(ecor, L, R, tail - sgn * e)
        t_fan = sample(ecor, L, R, tail + sgn * e)
        assert np.allclose(t_star, t_fan, rtol=1e-7, atol=1e-7)
        assert t_star[1] == pytest.approx(us, abs=1e-15) and t_star[3] == pytest.approx(ps, rel=1e-15)


def gresho_exact(ecor, n, cx, cy, periodic=(0, 0), prob_hi=(1.0, 1.0), mach=0.1):
    return ecor.make_exact(ecor.EXACT_GRESHO, periodic, prob_hi=prob_hi, vortex_center=(cx, cy),
                           density_outer=1.0, pressure=1.0 / (GAMMA * mach * mach), mach=mach)


def test_gresho_radial_balance(ecor):
    n = 400
    c = (200 + 0.5) / n
    e = gresho_exact(ecor, n, c, c)
    j = 200
    prim = lambda k: ecor.to_primitive(ecor.exact_state(e, n, n, 0.0, k, j))
    for i in (212, 230, 250, 270, 300, 350, 400):   # r from 0.03 to 0.5 (beyond 0.4: p flat)
        q0 = prim(i)
        r = (i - 200) / n
        d1 = (prim(i + 1)[3] - prim(i - 1)[3]) * n / 2
        d2 = (prim(i + 2)[3] - prim(i - 2)[3]) * n / 4
        dpdr = (4 * d1 - d2) / 3                            # Richardson, O(dx^4)
        assert q0[1] == pytest.approx(0.0, abs=1e-12)       # tangential on this ray
        assert dpdr == pytest.approx(q0[0] * q0[2] ** 2 / r, rel=1e-5, abs=1e-9)
    # v_theta peaks at r = 0.2 at the velocity scale mach * sqrt(gamma p0 / rho) = 1
    vpeak = prim(280)[2]
    assert vpeak == pytest.approx(1.0, rel=1e-12)


def test_quadrant_table(ecor):
    e = ecor.make_exact(ecor.EXACT_QUADRANT, (0, 0), interface_x=0.5, interface_y=0.5)
    n = 8
    prim = lambda i, j: ecor.to_primitive(ecor.exact_state(e, n, n, 0.0, i, j))
    # Lax-Liu configuration 3 (quadrants 1..4 counter-clockwise from upper right)
    table = {(9, 9): (1.5, 0, 0, 1.5), (-1, 9): (0.5323, 1.206, 0, 0.3),
             (-1, -1): (0.138, 1.206, 1.206, 0.029), (9, -1): (0.5323, 0, 1.206, 0.3)}
    for (i, j), q in table.items():
        assert np.allclose(prim(i, j), q, rtol=2e-3, atol=1e-12)
    q2, q4 = prim(-1, 9), prim(9, -1)
    assert np.array_equal(q2[[0, 1, 2, 3]], q4[[0, 2, 1, 3]])   # x <-> y mirror


def _grown_step(ecor, e, nx, ny, dx, dy, t, dt, riemann, periodic):
    """One step on the grid grown by two cells along every non-periodic axis,
    its cells holding the exact states; returns the original grid's cells."""
    gx, gy = (0 if periodic[0] else 2), (0 if periodic[1] else 2)
    Ug = np.zeros((4, ny + 2 * gy, nx + 2 * gx))
    for j in range(-gy, ny + gy):
        for i in range(-gx, nx + gx):
            Ug[:, j + gy, i + gx] = ecor.exact_state(e, nx, ny, t, i, j)
    bc = tuple(0 if periodic[a // 2] else 1 for a in range(4))
    g = ecor.make_grid(nx + 2 * gx, ny + 2 * gy, dx, dy, riemann=riemann, bc=bc)
    return ecor.fv_step(g, Ug, dt)[0][:, gy:gy + ny, gx:gx + nx], Ug[:, gy:gy + ny, gx:gx + nx]


CASES = [
    ("gresho", (0, 0)), ("quadrant", (0, 0)), ("blob", (0, 0)), ("toro", (0, 0)),
    ("blob", (0, 1)), ("gresho_cross", (0, 1)), ("toro", (0, 1)),
]


def make_case(ecor, name, periodic):
    nx, ny = 16, 8
    hi = (1.0, 0.5)
    if name.startswith("gresho"):
        e = gresho_exact(ecor, nx, 0.25, 0.25, periodic, hi)
    elif name == "quadrant":
        e = ecor.make_exact(ecor.EXACT_QUADRANT, periodic, prob_hi=hi, interface_x=0.5,
                            interface_y=0.25)
    elif name == "blob":
        e = ecor.make_exact(ecor.EXACT_BLOB, periodic, prob_hi=hi, blob_center=(-0.2, 0.05),
                            blob_radius=0.15, density_inner=2.0, density_outer=1.0,
                            pressure=1.0, velocity=(1.0, 0.5))
    else:
        L, R = TORO[1]
        e = ecor.make_exact(ecor.EXACT_TORO, periodic, prob_hi=hi, ql=L, qr=R, interface_x=0.5)
    return e, nx, ny, 1.0 / 16, 1.0 / 16


@pytest.mark.parametrize("name,periodic", CASES)
def test_exact_ghosts_equal_grown_grid(ecor, name, periodic):
    e, nx, ny, dx, dy = make_case(ecor, name, periodic)
    # HLLC-P reads corner ghosts; the Gresho and quadrant states at a corner
    # are taken at the unwrapped centre, so across a periodic axis only the
    # cross-shaped stencils (HLLC) match the grown grid.
    riemann = ecor.HLLC if name == "gresho_cross" else ecor.XIE_HLLCP
    t, dt = (0.05, 0.002)
    Ubig, U = _grown_step(ecor, e, nx, ny, dx, dy, t, dt, riemann, periodic)
    bc = tuple(0 if periodic[a // 2] else 1 for a in range(4))
    g = ecor.set_exact(ecor.make_grid(nx, ny, dx, dy, riemann=riemann, bc=bc), e, t)
    Un = ecor.fv_step(g, U, dt)[0]
    assert np.array_equal(Un, Ubig)
    # the ghosts do act: with outflow ghosts the step differs (the quadrant
    # and Toro states are piecewise constant across these boundaries, so
    # there outflow and exact ghosts agree)
    if name in ("gresho", "gresho_cross", "blob"):
        g0 = ecor.make_grid(nx, ny, dx, dy, riemann=riemann, bc=bc)
        assert not np.array_equal(ecor.fv_step(g0, U, dt)[0], Un)


def test_toro_ghosts_at_time_zero_are_the_initial_states(ecor):
    L, R = TORO[1]
    nx, ny = 40, 4
    e = ecor.make_exact(ecor.EXACT_TORO, (0, 0), prob_hi=(1.0, 0.1), ql=L, qr=R, interface_x=0.5)
    U = ecor.init_toro(nx, ny, L, R)
    g = ecor.make_grid(nx, ny, 1.0 / nx, 0.1 / ny, riemann=ecor.XIE_HLLCP, bc=(1, 1, 1, 1))
    ge = ecor.set_exact(ecor.make_grid(nx, ny, 1.0 / nx, 0.1 / ny, riemann=ecor.XIE_HLLCP,
                                       bc=(1, 1, 1, 1)), e, 0.0)
    assert np.array_equal(ecor.fv_step(ge, U, 0.001)[0], ecor.fv_step(g, U, 0.001)[0])


def test_exact_ghost_predicate(ecor):
    nx, ny = 6, 5
    cases = [  # (kind, periodic, cell, exact?)
        (ecor.EXACT_GRESHO, (0, 1), (-1, 2), True), (ecor.EXACT_GRESHO, (0, 1), (2, -1), False),
        (ecor.EXACT_GRESHO, (0, 1), (-1, -1), True), (ecor.EXACT_GRESHO, (1, 1), (-1, 2), False),
        (ecor.EXACT_TORO, (1, 0), (-1, 2), True), (ecor.EXACT_TORO, (1, 0), (2, ny), True),
        (ecor.EXACT_TORO, (0, 1), (2, ny), False), (ecor.EXACT_BLOB, (1, 0), (nx, 2), False),
    ]
    for kind, per, (i, j), want in cases:
        g = ecor.set_exact(ecor.make_grid(nx, ny, 0.1, 0.1), ecor.make_exact(kind, per), 0.0)
        assert ecor.exact_ghost(g, i, j) == want, (kind, per, i, j)


def test_blob_ghosts_follow_the_advected_centre(ecor):
    nx, ny = 20, 20
    e = ecor.make_exact(ecor.EXACT_BLOB, (0, 0), blob_center=(0.35, 0.35), blob_radius=0.15,
                        density_inner=2.0, density_outer=1.0, pressure=1.0, velocity=(1.0, 0.5))
    t = 0.7   # centre at (1.05, 0.7): just right of the domain
    i = nx      # first ghost column, x = 1.025
    for j in range(ny):
        y = (j + 0.5) / ny
        inside = (1.025 - 1.05) ** 2 + (y - 0.7) ** 2 <= 0.15 ** 2
        u = ecor.exact_state(e, nx, ny, t, i, j)
        assert u[0] == (2.0 if inside else 1.0)
        assert u[1] == u[0] * 1.0 and u[2] == u[0] * 0.5
