"""The cylindrical shock bubble's geometric source on the device
(ec_cylindrical_source) against the oracle
(problems/boundary_conditions.hpp:75-131).

Bar: bit for bit.  The source is pointwise IEEE arithmetic in the
reference's order, and the explicit loop around it (dt from the maximum
rate, the update, then the source: app/driver_main.hpp:224-291) is
bit-identical to the oracle's step by step.
"""
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GAMMA = 1.4
POST = (2.8181818181818183, 1.6064386578049976, 0.0, 5.0)
EXE = Path(__file__).resolve().parent.parent / "amrex-euler-allmach-explicit-imex_amd" / "euler_compare_amd"


@pytest.fixture(scope="module")
def ecamd():
    import ecamd as m
    return m


def cons(r, u, v, p):
    return np.stack([r, r * u, r * v, p / (GAMMA - 1) + 0.5 * r * (u * u + v * v)])


def test_source_matches_oracle(ecor, ecamd, gpu):
    nx, ny = 301, 67
    rng = np.random.default_rng(11)
    U = cons(0.5 + rng.random((ny, nx)), rng.standard_normal((ny, nx)), rng.standard_normal((ny, nx)),
             0.2 + rng.random((ny, nx)))
    U[0, 5, 7] = -0.25        # rho <= 0 reads a zero velocity
    U[0, 9, 3] = 0.0
    dx, dy = 2.0 / nx, 0.5 / ny
    s = ecamd.ExplicitSolver(ecamd.Grid(nx, ny, dx, dy, GAMMA, "hllc", "minmod", 2, ("outflow",) * 4))
    g = ecor.make_grid(nx, ny, dx, dy, bc=(1, 1, 1, 1))
    for dt, ylo in ((3e-3, 0.0), (1e-2, -0.1)):      # ylo < 0: rows at r <= 0 stay
        Ud = torch.from_numpy(U.copy()).to(gpu)
        s.cylindrical_source(Ud, dt, ylo)
        torch.cuda.synchronize()
        assert np.array_equal(Ud.cpu().numpy(), ecor.cylindrical_source(g, U, dt, ylo))


def sb_setup(ecor, ecamd, riemann, nx=160, ny=40):
    from ecamd import problems
    U = problems.initial_state("shock_density_bubble", nx, ny, prob_lo=(0.0, 0.0), prob_hi=(2.0, 0.5))
    fixed = cons(*POST)
    dx, dy = 2.0 / nx, 0.5 / ny
    names = {1: "hllc", 3: "xie_am_hllc_p"}
    s = ecamd.ExplicitSolver(ecamd.Grid(nx, ny, dx, dy, GAMMA, names[riemann], "minmod", 2,
                                        ("fixed", "outflow", "reflect", "outflow"),
                                        (tuple(fixed), None, None, None)))
    g = ecor.make_grid(nx, ny, dx, dy, riemann=riemann, bc=(3, 1, 2, 1), fixed=(list(fixed), None, None, None))
    return s, g, U


@pytest.mark.parametrize("riemann", [1, 3])
def test_cylindrical_loop_matches_oracle(ecor, ecamd, gpu, riemann):
    s, g, U = sb_setup(ecor, ecamd, riemann)
    A = torch.from_numpy(U.copy()).to(gpu)
    B = torch.empty_like(A)
    t, stop = 0.0, 0.3
    for k in range(40):
        rate = s.diagnostics(A).max_rate
        assert rate == ecor.max_rate(g, U), k
        dt = ecor.euler_dt(rate, 0.45, t, stop)
        s.step(A, B, dt, diag=False)
        A, B = B, A
        s.cylindrical_source(A, dt, 0.0)
        U, _ = ecor.fv_step(g, U, dt)
        U = ecor.cylindrical_source(g, U, dt, 0.0)
        t += dt
    got = A.cpu().numpy()
    assert np.array_equal(got, U)
    # the source acted: the planar loop differs
    s2, g2, U2 = sb_setup(ecor, ecamd, riemann)
    t = 0.0
    for k in range(40):
        dt = ecor.euler_dt(ecor.max_rate(g2, U2), 0.45, t, stop)
        U2, _ = ecor.fv_step(g2, U2, dt)
        t += dt
    assert not np.array_equal(U2, U)


def test_cylindrical_driver_row(tmp_path, ecor, ecamd):
    """euler.problem=shock_density_bubble_cylindrical_clawpack through the
    drop-in driver: the shock-bubble defaults (160 x 40 on [0,2] x [0,0.5],
    CFL 0.45), the source after every update, the cylindrical labels, and the
    snapshot at t = 0.02 equal to the oracle's loop clipped at that time."""
    import csv
    (tmp_path / "inputs").write_text("max_step = 8\nstop_time = 0.05\n")
    snap = tmp_path / "snaps"
    r = subprocess.run([str(EXE), str(tmp_path / "inputs"), "euler.problem=shock_density_bubble_cylindrical",
                        "euler.riemann=hllc", "stop_time=0.02", "max_step=100000",
                        "euler.shock_density_bubble_snapshot_times=0.02",
                        f"euler.shock_density_bubble_snapshot_dir={snap}"],
                       capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    kv = dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line)
    assert kv["problem"] == "shock_density_bubble_cylindrical_clawpack"
    assert kv["source_case_id"] == "same_gamma_shock_density_bubble_clawpack_cylindrical_v1"
    assert kv["cylindrical_source_used"] == "true"
    assert kv["geometric_source_form"] == "radial_symmetry_fractional_source_step"
    assert kv["validation_claim"] == "qualitative_project_clawpack_style_cylindrical_high_speed_2d_only"
    assert kv["grid_nx"] == "160" and kv["grid_ny"] == "40"
    files = sorted(snap.iterdir())
    assert len(files) == 1
    rows = list(csv.DictReader(open(files[0])))
    assert rows[0]["cylindrical_source_used"] == "true"
    assert rows[0]["geometric_source_form"] == "clawpack_radial_symmetry_fractional_source_step"
    rho = np.full((40, 160), np.nan)
    for x in rows:                   # box by box (amr.max_grid_size): place by i, j
        rho[int(x["j"]), int(x["i"])] = float(x["rho"])
    assert not np.isnan(rho).any()
    s, g, U = sb_setup(ecor, ecamd, 1)
    t, steps = 0.0, 0
    while t < 0.02 - 1e-14:
        dt = ecor.euler_dt(ecor.max_rate(g, U), 0.45, t, 0.02)
        U, _ = ecor.fv_step(g, U, dt)
        U = ecor.cylindrical_source(g, U, dt, 0.0)
        t += dt
        steps += 1
    assert kv["completed_steps"] == str(steps)
    assert np.array_equal(rho, U[0])
