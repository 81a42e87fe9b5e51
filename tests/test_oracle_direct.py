"""Pins of the oracle's direct BDLTV20 T1/S2 step (imex/
imex_bdltv20_direct_driver.hpp:44-779).

* At epsilon = 1 on a periodic grid the direct step solves the same pressure
  system as the source-map step (same star, face enthalpies, rows and
  momentum correction; only the closeout differs): its pressure and momentum
  agree with the source-map oracle (itself pinned to the reference's GMRES
  iteration counts) to the solve tolerance, and its energy is the pressure
  closeout E = p/(g-1) + eps |m|^2 / (2 rho).
* A uniform flow is a fixed point for any epsilon, with exact-Dirichlet
  boundary pressures folded into the right-hand side: a wrong epsilon factor
  on the diagonal, the right-hand side or the correction, or a wrong sign in
  the fold, moves the pressure.
* Sod on a y-periodic strip with the x-only row converges to the exact
  Riemann solution under refinement and stays y-uniform.
* The eq79 time step is cfl * min(dx, dy) / (2 |u|_max), clipped to the stop.
"""
import numpy as np
import pytest

GAMMA = 1.4


def cons_eps(ecor, prim, eps):
    """(4, ny, nx) primitive -> epsilon-scaled conserved."""
    r, u, v, p = prim
    return np.stack([r, r * u, r * v, p / (GAMMA - 1) + eps * 0.5 * r * (u * u + v * v)])


def smooth_periodic(nx, ny):
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    r = 1.0 + 0.2 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y)
    u = 0.6 + 0.1 * np.cos(2 * np.pi * Y)
    v = 0.4 + 0.1 * np.sin(2 * np.pi * X)
    p = 1.0 + 0.1 * np.cos(2 * np.pi * (X + Y))
    return np.stack([r, u, v, p])


def test_eps1_periodic_matches_source_map(ecor):
    nx, ny = 24, 20
    dx, dy = 1.0 / nx, 1.0 / ny
    U = cons_eps(ecor, smooth_periodic(nx, ny), 1.0)
    g = ecor.make_grid(nx, ny, dx, dy, bc=(0, 0, 0, 0))
    d = ecor.make_direct(ecor.make_exact(ecor.EXACT_NONE, (1, 1)), 1.0)
    dt = 0.4 * dx
    Ud, pd, rd = ecor.direct_step(g, d, U, 0.0, dt, picard=3, tol=1e-13, max_iter=2000)
    Us, ps, rs = ecor.imex_step(g, U, dt, picard=3, tol=1e-13, max_iter=2000)
    assert rd.solver_failures == 0 and rs.status in (0, 1)
    assert rd.guard_count == 0 and rd.nonfinite_count == 0
    assert np.abs(pd - ps).max() <= 1e-10 * np.abs(ps).max()
    for c in (0, 1, 2):
        assert np.abs(Ud[c] - Us[c]).max() <= 1e-10 * np.abs(Us[c]).max()
    E = pd / (GAMMA - 1) + 0.5 * (Ud[1] ** 2 + Ud[2] ** 2) / Ud[0]
    assert np.abs(Ud[3] - E).max() <= 1e-14 * np.abs(E).max()


@pytest.mark.parametrize("eps", [1.0, 0.01, 1e-4])
@pytest.mark.parametrize("kind,periodic,x_only", [
    ("none", (1, 1), False), ("toro", (0, 1), True), ("toro", (0, 0), True), ("toro", (0, 0), False)])
def test_uniform_flow_is_a_fixed_point(ecor, eps, kind, periodic, x_only):
    nx, ny = 16, 6
    q = [1.3, 0.3, 0.0, 0.8]
    prim = np.broadcast_to(np.array(q)[:, None, None], (4, ny, nx))
    U = cons_eps(ecor, prim, eps)
    ek = ecor.EXACT_TORO if kind == "toro" else ecor.EXACT_NONE
    e = ecor.make_exact(ek, periodic, prob_hi=(1.0, 6 / 16), ql=q, qr=q, interface_x=0.5)
    g = ecor.make_grid(nx, ny, 1 / 16, 1 / 16, bc=(0, 0, 0, 0))
    d = ecor.make_direct(e, eps, x_only)
    dt = 0.05
    U1, p, res = ecor.direct_step(g, d, U, 0.1, dt, picard=2, tol=1e-13, max_iter=500)
    assert res.solver_failures == 0
    assert np.abs(p - 0.8).max() <= 1e-11
    assert np.abs(U1 - U).max() <= 1e-11 * np.abs(U).max()


def sod_run(ecor, nx, t_end=0.2, eps=1.0):
    ny = 2
    L, R = [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1]
    e = ecor.make_exact(ecor.EXACT_TORO, (0, 1), prob_hi=(1.0, ny / nx), ql=L, qr=R,
                        interface_x=0.5)
    g = ecor.make_grid(nx, ny, 1 / nx, 1 / nx, bc=(1, 1, 0, 0))
    d = ecor.make_direct(e, eps, True)
    x = (np.arange(nx) + 0.5) / nx
    prim = np.stack([np.where(x < 0.5, a, b) for a, b in zip(L, R)])
    U = cons_eps(ecor, np.broadcast_to(prim[:, None, :], (4, ny, nx)).copy(), eps)
    t, dt = 0.0, 0.5 / nx
    while t < t_end - 1e-14:
        h = ecor.direct_dt(g, d, U, 0.8, dt, t, t_end)
        U, _, res = ecor.direct_step(g, d, U, t, h, picard=4, tol=1e-10, max_iter=2000)
        assert res.solver_failures == 0 and res.nonpositive_count == 0
        t += h
    ex = np.array([ecor.exact_riemann_sample(xx, t_end, L, R)[0] for xx in x])
    return U, np.abs(U[0, 0] - ex).mean()


def test_sod_strip_converges_to_exact_solution(ecor):
    errs = []
    for nx in (50, 100, 200):
        U, err = sod_run(ecor, nx)
        assert np.array_equal(U[:, 0], U[:, 1])   # y-uniform
        errs.append(err)
    assert errs[2] < 0.01
    assert errs[0] / errs[1] > 1.3 and errs[1] / errs[2] > 1.3


def test_eq79_time_step(ecor):
    nx, ny = 8, 4
    prim = np.zeros((4, ny, nx)); prim[0] = 1.0; prim[3] = 1.0
    prim[1, 2, 3] = 3.0; prim[2, 2, 3] = 4.0     # |u| = 5 at one cell
    U = cons_eps(ecor, prim, 0.5)
    g = ecor.make_grid(nx, ny, 0.1, 0.05, bc=(0, 0, 0, 0))
    d = ecor.make_direct(ecor.make_exact(ecor.EXACT_NONE, (1, 1)), 0.5)
    assert ecor.direct_dt(g, d, U, 0.8, -1.0, 0.0, 1.0) == pytest.approx(0.8 * 0.05 / 10.0, rel=1e-15)
    assert ecor.direct_dt(g, d, U, 0.8, -1.0, 0.999, 1.0) == pytest.approx(0.001, rel=1e-12)
    assert ecor.direct_dt(g, d, U, 0.8, 0.01, 0.0, 1.0) == 0.01     # a fixed dt
    rest = U.copy(); rest[1:3] = 0.0
    assert ecor.direct_dt(g, d, rest, 0.8, -1.0, 0.25, 1.0) == 0.75  # no flow: the remaining time


def test_boundary_state_points(ecor):
    """Outside indices sample the exact solution at the domain edge."""
    L, R = [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1]
    e = ecor.make_exact(ecor.EXACT_TORO, (0, 0), prob_hi=(1.0, 0.1), ql=L, qr=R, interface_x=0.5)
    g = ecor.make_grid(10, 2, 0.1, 0.05)
    d = ecor.make_direct(e, 0.5)
    t = 0.3
    for i, j, x in ((-1, 0, 0.0), (-2, 1, 0.0), (10, 0, 1.0), (11, 1, 1.0), (3, -1, 3.5 * 0.1), (7, 2, 7.5 * 0.1)):
        want = ecor.eps_to_conserved(ecor.exact_riemann_sample(x, t, L, R), eps=0.5)
        assert np.array_equal(ecor.direct_boundary_state(g, d, i, j, t), want), (i, j)
