"""Exact-Dirichlet ghost states on the device (euler.field_boundary=
exact_dirichlet) against the oracle.

Bar: bit for bit wherever the ghost states come from host-evaluated
formulas or from device arithmetic in the reference's operation order
(Gresho and quadrant tables, the advected blob, the Toro solution outside its
rarefaction fan).  Inside a Toro rarefaction fan the ghost state takes two
pow() calls that the device evaluates with its own math library; there the
bar is elementwise relative 1e-13 (rel_l1 <= 1e-14), far inside the 1e-12
relative L1 bound of the explicit path.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GAMMA = 1.4
TORO = {1: ([1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1])}
FAN_RTOL = 1e-13


def rel_l1(a, b):
    a = np.asarray(a).reshape(4, -1)
    b = np.asarray(b).reshape(4, -1)
    den = np.abs(b).sum(axis=1)
    return np.abs(a - b).sum(axis=1) / np.where(den == 0, 1, den)


@pytest.fixture(scope="module")
def ecamd():
    import ecamd as m
    return m


PARAMS = {
    "gresho": dict(vortex_center=(0.25, 0.3), density_outer=1.0, mach=0.1,
                   pressure=1.0 / (GAMMA * 0.01)),
    "quadrant": dict(riemann_interface_x=0.4, riemann_interface_y=0.3),
    "blob": dict(blob_center=(-0.1, 0.2), blob_radius=0.15, density_inner=2.0, density_outer=1.0,
                 pressure=1.0, velocity=(1.0, 0.5)),
    "toro": dict(toro_test=1, riemann_interface_x=0.5),
}
# the blob entering through the bottom boundary
PARAMS["blob_bottom"] = dict(PARAMS["blob"], blob_center=(0.3, -0.3))
KIND = {"gresho": "gresho_vortex", "quadrant": "riemann_quadrant", "blob": "advection_blob",
        "blob_bottom": "advection_blob", "toro": "toro1"}


def oracle_exact(ecor, name, periodic, hi):
    p = PARAMS[name]
    if name == "gresho":
        return ecor.make_exact(ecor.EXACT_GRESHO, periodic, prob_hi=hi, vortex_center=p["vortex_center"],
                               density_outer=p["density_outer"], pressure=p["pressure"], mach=p["mach"])
    if name == "quadrant":
        return ecor.make_exact(ecor.EXACT_QUADRANT, periodic, prob_hi=hi,
                               interface_x=p["riemann_interface_x"], interface_y=p["riemann_interface_y"])
    if name.startswith("blob"):
        return ecor.make_exact(ecor.EXACT_BLOB, periodic, prob_hi=hi, blob_center=p["blob_center"],
                               blob_radius=p["blob_radius"], density_inner=p["density_inner"],
                               density_outer=p["density_outer"], pressure=p["pressure"],
                               velocity=p["velocity"])
    L, R = TORO[p["toro_test"]]
    return ecor.make_exact(ecor.EXACT_TORO, periodic, prob_hi=hi, ql=L, qr=R,
                           interface_x=p["riemann_interface_x"])


def setup(ecor, ecamd, name, periodic, riemann, nx=37, ny=29, hi=(1.0, 0.7)):
    from ecamd import problems
    pd = problems.problem_desc(KIND[name], nx, ny, prob_hi=hi, gamma=GAMMA, **PARAMS[name])
    U0 = problems.initial_state(KIND[name], nx, ny, prob_hi=hi, gamma=GAMMA, **PARAMS[name])
    dx, dy = hi[0] / nx, hi[1] / ny
    # outflow placeholders on every side; the exact set-up picks the sides
    bc = tuple(0 if periodic[a // 2] else 1 for a in range(4))
    g = ecor.set_exact(ecor.make_grid(nx, ny, dx, dy, riemann=riemann, bc=bc),
                       oracle_exact(ecor, name, periodic, hi), 0.0)
    names = {0: "rusanov", 1: "hllc", 2: "low_mach_hllc", 3: "xie_am_hllc_p"}
    grid = ecamd.Grid(nx, ny, dx, dy, GAMMA, names[riemann], "minmod", 2,
                      tuple("periodic" if b == 0 else "outflow" for b in bc))
    s = ecamd.ExplicitSolver(grid)
    s.set_exact_boundary(pd, periodic)
    return s, g, U0


def check(got, ref, bitwise):
    if bitwise:
        assert np.array_equal(got, ref), (np.abs(got - ref).max(), rel_l1(got, ref))
    else:
        assert np.allclose(got, ref, rtol=FAN_RTOL, atol=0), np.abs(got - ref).max()
        assert rel_l1(got, ref).max() <= 1e-14


STEP_CASES = [
    # (problem, periodic, riemann, time, bitwise)
    ("gresho", (0, 0), 3, 0.0, True), ("gresho", (0, 1), 3, 0.0, True),
    ("gresho", (1, 0), 1, 0.0, True),
    ("quadrant", (0, 0), 3, 0.0, True), ("quadrant", (1, 0), 3, 0.0, True),
    ("blob", (0, 0), 3, 0.1, True), ("blob", (0, 1), 3, 0.1, True), ("blob_bottom", (1, 0), 2, 0.5, True),
    ("toro", (0, 1), 3, 0.3, True),      # x ghosts: left state, and behind the right shock
    ("toro", (0, 1), 3, 0.45, False),    # the left ghosts inside the rarefaction fan
    ("toro", (0, 0), 3, 0.2, False),     # y ghost rows sample the whole solution
    ("toro", (1, 0), 1, 0.2, False),     # x periodic: x ghosts exact all the same
]


@pytest.mark.parametrize("name,periodic,riemann,t,bitwise", STEP_CASES)
def test_exact_step_matches_oracle(ecor, ecamd, gpu, name, periodic, riemann, t, bitwise):
    s, g, U = setup(ecor, ecamd, name, periodic, riemann)
    g.time = t
    dt = 0.3 * 0.8 / ecor.max_rate(g, U)
    ref, deg = ecor.fv_step(g, U, dt)
    s.set_boundary_time(t)
    Ud = torch.from_numpy(U).to(gpu)
    out = torch.empty_like(Ud)
    d = s.step(Ud, out, dt)
    check(out.cpu().numpy(), ref, bitwise)
    assert d.degenerate_count == deg
    # and the exact ghosts matter here: outflow ghosts give another step
    g_out = ecor.make_grid(g.nx, g.ny, g.dx, g.dy, riemann=riemann, bc=tuple(g.bc))
    if name != "quadrant":
        assert not np.array_equal(ecor.fv_step(g_out, U, dt)[0], ref)


@pytest.mark.parametrize("name,periodic,bitwise", [
    ("gresho", (0, 0), True), ("blob", (0, 1), True), ("toro", (0, 0), False)])
def test_exact_device_loop_matches_oracle(ecor, ecamd, gpu, name, periodic, bitwise):
    """ec_explicit_advance: the ghost time is the run's own time each step."""
    nsteps = 24
    s, g, U = setup(ecor, ecamd, name, periodic, 3, nx=48, ny=20, hi=(1.0, 0.4))
    stop = 0.3 if name != "gresho" else 1.0
    ref, res, _ = ecor.run_explicit(g, U, 0.8, stop, nsteps)
    Ua = torch.from_numpy(U).to(gpu)
    Ub = torch.empty_like(Ua)
    run = s.new_run(0.8, stop, nsteps)
    fin = s.advance(Ua, Ub, nsteps, run)
    assert run.step == res.steps == nsteps
    if bitwise:
        assert run.time == res.time
    else:
        assert run.time == pytest.approx(res.time, rel=1e-12)
    check(fin.cpu().numpy(), ref, bitwise)


def test_exact_boundary_both_periodic_is_periodic(ecor, ecamd, gpu):
    s, g, U = setup(ecor, ecamd, "gresho", (1, 1), 3)
    plain = ecamd.ExplicitSolver(ecamd.Grid(g.nx, g.ny, g.dx, g.dy, GAMMA, "xie_am_hllc_p", "minmod", 2,
                                            ("periodic",) * 4))
    Ud = torch.from_numpy(U).to(gpu)
    a, b = torch.empty_like(Ud), torch.empty_like(Ud)
    s.step(Ud, a, 1e-4)
    plain.step(Ud, b, 1e-4)
    assert torch.equal(a, b)


@pytest.mark.parametrize("nx,ny", [(1, 1), (1, 4), (2, 3), (3, 2), (5, 1)])
@pytest.mark.parametrize("name,periodic,t", [("gresho", (0, 0), 0.0), ("quadrant", (1, 0), 0.0),
                                             ("toro", (0, 1), 0.3), ("blob", (0, 0), 0.1)])
def test_exact_step_on_tiny_grids(ecor, ecamd, gpu, nx, ny, name, periodic, t):
    """Grids smaller than the ghost ring: every ghost (and every unused lane
    of the partial tile) reads a ring cell or the exact solution, bit for bit
    with the oracle."""
    print(f"exact tiny {name} {periodic} {nx}x{ny}", flush=True)
    s, g, U = setup(ecor, ecamd, name, periodic, 3, nx=nx, ny=ny, hi=(1.0, 0.7))
    g.time = t
    dt = 0.1 * 0.8 / ecor.max_rate(g, U)
    ref, deg = ecor.fv_step(g, U, dt)
    s.set_boundary_time(t)
    Ud = torch.from_numpy(U).to(gpu)
    out = torch.empty_like(Ud)
    d = s.step(Ud, out, dt)
    check(out.cpu().numpy(), ref, True)
    assert d.degenerate_count == deg


def test_exact_boundary_rejections(ecor, ecamd, gpu):
    from ecamd import problems
    from ecamd._lib import lib
    s, g, U = setup(ecor, ecamd, "gresho", (0, 0), 3)
    # a problem of another size, or one without an exact state
    bad = problems.problem_desc("gresho_vortex", g.nx + 1, g.ny)
    per = (C.c_int * 2)(0, 0)
    assert lib().ec_set_exact_boundary(s.ctx, C.byref(bad), per) == 1
    iv = problems.problem_desc("isentropic_vortex", g.nx, g.ny)
    assert lib().ec_set_exact_boundary(s.ctx, C.byref(iv), per) == 1
    # the adjoint does not take exact sides (the reference has no adjoint)
    from ecamd import adjoint
    Ud = torch.from_numpy(U).to(gpu)
    with pytest.raises(RuntimeError, match="unsupported|exact"):
        adjoint.adjoint_step(s, Ud, torch.zeros_like(Ud), 1e-4)


@pytest.mark.parametrize("name,periodic", [("gresho", (0, 0)), ("toro", (0, 1)), ("toro", (1, 0))])
def test_imex_step_reads_exact_sides_as_the_source_map_step(ecor, ecamd, gpu, name, periodic):
    """The source-map step wraps periodic ghosts and clamps physical ones
    before its star update whatever euler.field_boundary says
    (imex/imex_source_map_step.hpp:405-409), and reads valid cells only
    afterwards.  A context with exact sides therefore takes the same IMEX step
    as one with outflow sides (periodic ones on a periodic axis, Toro's x
    included): bit for bit, and the star density (no solve in it) bit for bit
    against the oracle's step with those sides."""
    s, g, U = setup(ecor, ecamd, name, periodic, 3, nx=32, ny=24)
    nx, ny = g.nx, g.ny
    bc = tuple(0 if periodic[a // 2] else 1 for a in range(4))
    plain = ecamd.ExplicitSolver(ecamd.Grid(nx, ny, g.dx, g.dy, GAMMA, "xie_am_hllc_p", "minmod", 2,
                                            tuple("periodic" if b == 0 else "outflow" for b in bc)))
    go = ecor.make_grid(nx, ny, g.dx, g.dy, riemann=3, bc=bc)
    dt, _, _ = ecor.imex_dt(go, U, 0.8, 0.0, 1.0)
    opts = dict(tol=1e-10, max_iter=3000, picard_iterations=2, restart=64)
    out = []
    for solver in (s, plain):
        Ud = torch.from_numpy(U.copy()).to(gpu)
        p = torch.empty(nx * ny, dtype=torch.float64, device=gpu)
        d = solver.imex_step(Ud, dt, solver.imex_opts(**opts), p)
        assert ecamd.IMEX_STATUS[d.status] in ("converged_absolute", "converged_relative"), d.status
        out.append((Ud.cpu().numpy(), p.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    ref, _, r = ecor.imex_step(go, U, dt, picard=2, tol=1e-10, max_iter=3000)
    assert r.status in (0, 1)
    assert np.array_equal(out[0][0][0], ref[0])
    if name != "gresho":
        return           # Toro's ghosts at t = 0 equal the clamped end states
    # the vortex reaches the sides: the exact ghosts would change the star
    ge = ecor.set_exact(ecor.make_grid(nx, ny, g.dx, g.dy, riemann=3, bc=bc),
                        oracle_exact(ecor, name, periodic, (1.0, 0.7)), 0.0)
    assert not np.array_equal(ecor.imex_star(ge, U, dt)[0], ecor.imex_star(go, U, dt)[0])
