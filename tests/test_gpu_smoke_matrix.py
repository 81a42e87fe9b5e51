"""The reference's own solver smoke matrix through the drop-in driver.

The twelve rows of the reference's smoke matrix (tests/golden/
smoke_matrix_rows.json, kept as data: the command-line overrides of
scripts/run_project_smoke_matrix.py:42-276) run through euler_compare_amd
exactly as the script runs the reference executable, and each must pass the
script's rule (:342-346): return code 0, status ok or passed, failure category
ok or none.  Beside the pass rule, the final state of every explicit row is
compared with the oracle's run from the same initial data (bit for bit,
except where exact-Dirichlet ghosts sample a Toro rarefaction fan with the
device pow(): 1e-12), and the main-loop IMEX shock-bubble row (tolerance
1e-10, at most 200 Krylov iterations per solve) with the oracle's IMEX loop,
whose solve is the reference's unrestarted GMRES, at solver tolerance."""
import csv
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "amrex-euler-allmach-explicit-imex_amd" / "euler_compare_amd"
ROWS = json.loads((ROOT / "tests" / "golden" / "smoke_matrix_rows.json").read_text())["rows"]
POST_SHOCK = (2.8181818181818183, 1.6064386578049976, 0.0, 5.0)


def run_row(row, tmp_path, extra=()):
    snap = tmp_path / "snap"
    args = [a.replace("{snapdir}", str(snap)) for a in row["args"]]
    args += [f"amr.plot_file={tmp_path / 'plt'}", *extra]
    r = subprocess.run([str(EXE), *args], capture_output=True, text=True, timeout=300, cwd=tmp_path)
    kv = {}
    for line in r.stdout.splitlines():
        if "=" in line:
            k, v = line.split("=", 1)
            kv[k.strip()] = v.strip()
    return r, kv, snap


def passes(r, kv):
    """scripts/run_project_smoke_matrix.py:317-346"""
    status = kv.get("project_amrex_euler_compare_status") or kv.get("bdltv20_paper_t1_s2_status") or ""
    failure = kv.get("failure_category", "none" if status in {"ok", "passed"} else "")
    return r.returncode == 0 and status in {"ok", "passed"} and failure in {"ok", "none"}, status, failure


def read_prims(path, nx, ny, dx, dy, xlo=0.0, ylo=0.0, pcol="p"):
    """(rho, u, v, p) of a CSV placed by each row's own x, y columns (any row
    order)."""
    out = np.full((4, ny, nx), np.nan)
    for r in csv.DictReader(open(path)):
        i = int(round((float(r["x"]) - xlo) / dx - 0.5))
        j = int(round((float(r["y"]) - ylo) / dy - 0.5))
        out[:, j, i] = [float(r["rho"]), float(r["u"]), float(r["v"]), float(r[pcol])]
    assert not np.isnan(out).any()
    return out


def prims(U, gamma=1.4):
    inv = 1.0 / U[0]
    u, v = U[1] * inv, U[2] * inv
    return np.stack([U[0], u, v, (gamma - 1.0) * (U[3] - 0.5 * U[0] * (u * u + v * v))])


def rel_l1(a, b):
    return np.abs(a - b).reshape(4, -1).sum(1) / np.abs(b).reshape(4, -1).sum(1)


@pytest.mark.parametrize("row", ROWS, ids=[r["name"] for r in ROWS])
def test_smoke_row_passes(row, tmp_path):
    print("smoke row:", row["name"], flush=True)
    r, kv, _ = run_row(row, tmp_path)
    ok, status, failure = passes(r, kv)
    assert ok, (row["name"], r.returncode, status, failure, r.stderr[-2000:])


def _arg(row, key):
    for a in row["args"]:
        if a.startswith(key + "="):
            return a.split("=", 1)[1]
    return None


EXPLICIT = [r for r in ROWS if _arg(r, "euler.method") == "explicit"]


@pytest.mark.parametrize("row", EXPLICIT, ids=[r["name"] for r in EXPLICIT])
def test_explicit_row_state_matches_oracle(row, tmp_path, ecor):
    from ecamd import problems
    print("smoke row state:", row["name"], flush=True)
    name = row["name"]
    rid = {"hllc": ecor.HLLC, "xie_am_hllc_p": ecor.XIE_HLLCP}[_arg(row, "euler.riemann")]
    stop, cfl = float(_arg(row, "stop_time")), float(_arg(row, "euler.cfl"))
    final = tmp_path / "final.csv"
    r, kv, snap = run_row(row, tmp_path, [f"euler.final_csv={final}"])
    assert passes(r, kv)[0], r.stderr[-2000:]
    if name.startswith("riemann_sod"):
        nx, ny, ly = 128, 4, 0.03125
        L, R = [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1]
        e = ecor.make_exact(ecor.EXACT_TORO, (0, 0), prob_hi=(1.0, ly), ql=L, qr=R, interface_x=0.5)
        g = ecor.set_exact(ecor.make_grid(nx, ny, 1 / nx, ly / ny, riemann=rid, bc=(1, 1, 1, 1)), e, 0.0)
        U0 = ecor.init_toro(nx, ny, L, R)
        ref, res, _ = ecor.run_explicit(g, U0, cfl, stop, 100000)
        got = read_prims(final, nx, ny, 1 / nx, ly / ny)
        tol = 1e-12                                  # the fan's ghosts take the device pow()
    elif name.startswith("gresho"):
        nx = ny = 32
        # the driver's problem defaults: centre (0, 0), rho 1, p0 1, M from the row
        U0 = problems.initial_state("gresho_vortex", nx, ny, mach=0.01, pressure=1.0,
                                    density_outer=1.0, vortex_center=(0.0, 0.0))
        e = ecor.make_exact(ecor.EXACT_GRESHO, (0, 0), vortex_center=(0.0, 0.0), density_outer=1.0,
                            pressure=1.0, mach=0.01)
        g = ecor.set_exact(ecor.make_grid(nx, ny, 1 / nx, 1 / ny, riemann=rid, bc=(1, 1, 1, 1)), e, 0.0)
        ref, res, _ = ecor.run_explicit(g, U0, cfl, stop, 100000)
        got = read_prims(final, nx, ny, 1 / nx, 1 / ny, pcol="pressure")
        tol = 0.0
    elif name.startswith("advection_blob"):
        nx = ny = 32
        U0 = problems.initial_state("advection_blob", nx, ny)
        g = ecor.make_grid(nx, ny, 1 / nx, 1 / ny, riemann=rid)   # periodic
        ref, res, _ = ecor.run_explicit(g, U0, cfl, stop, 100000)
        got = read_prims(final, nx, ny, 1 / nx, 1 / ny, pcol="pressure")
        tol = 0.0
    else:                                            # the shock-density bubble on [0,2] x [0,0.5]
        nx, ny = 160, 40
        U0 = problems.initial_state("shock_density_bubble_2d", nx, ny, prob_hi=(2.0, 0.5))
        first = sorted(snap.iterdir())[0]
        assert np.array_equal(read_prims(first, nx, ny, 2.0 / nx, 0.5 / ny, pcol="pressure"), prims(U0))
        g = ecor.make_grid(nx, ny, 2.0 / nx, 0.5 / ny, riemann=rid, bc=(3, 1, 2, 1),
                           fixed=(ecor.to_conserved(POST_SHOCK), None, None, None))
        ref, res, _ = ecor.run_explicit(g, U0, cfl, stop, 100000)
        last = sorted(snap.iterdir())[-1]
        assert last.name.startswith("snapshot_01_t0p030000")
        got = read_prims(last, nx, ny, 2.0 / nx, 0.5 / ny, pcol="pressure")
        tol = 0.0
    assert int(kv.get("completed_steps", res.steps)) == res.steps
    want = prims(ref)
    err = rel_l1(got, want)
    print(name, res.steps, "steps, rel L1", err)
    if tol == 0.0:
        assert np.array_equal(got, want), err
    else:
        assert np.all(err <= tol), err


def test_imex_shock_bubble_row_matches_oracle(tmp_path, ecor):
    """The main-loop IMEX row (euler.bdltv20_paper_t1_s2=off): the device
    restarted GMRES reaches the reference's statuses within its 200
    iterations per solve, and the state at t = 0.03 agrees with the oracle's
    IMEX loop (the reference's unrestarted MGS GMRES, same 200-iteration
    limit and 1e-10 tolerance) to solver tolerance."""
    from ecamd import problems
    row = next(r for r in ROWS if r["name"] == "shock_density_bubble_imex_t1s2_bdltv20_160x40")
    r, kv, snap = run_row(row, tmp_path)
    assert passes(r, kv)[0], r.stderr[-2000:]
    nx, ny = 160, 40
    U = problems.initial_state("shock_density_bubble_2d", nx, ny, prob_hi=(2.0, 0.5))
    first = sorted(snap.iterdir())[0]
    assert np.array_equal(read_prims(first, nx, ny, 2.0 / nx, 0.5 / ny, pcol="pressure"), prims(U))
    g = ecor.make_grid(nx, ny, 2.0 / nx, 0.5 / ny, bc=(3, 1, 2, 1),
                       fixed=(ecor.to_conserved(POST_SHOCK), None, None, None))
    t, steps, its = 0.0, 0, []
    while t < 0.03 - 1e-14:
        dt = ecor.imex_dt(g, U, 0.8, t, 0.03)[0]
        U, _, res = ecor.imex_step(g, U, dt, picard=4, tol=1e-10, max_iter=200)
        assert res.status in (0, 1), res.status
        its.append(res.last_iterations)
        t += dt
        steps += 1
    print("oracle IMEX steps", steps, "GMRES iterations of the last solves", its)
    assert int(kv["imex_step_count"]) == steps
    last = sorted(snap.iterdir())[-1]
    got = read_prims(last, nx, ny, 2.0 / nx, 0.5 / ny, pcol="pressure")
    want = prims(U)
    err = rel_l1(got, want)
    print("IMEX shock bubble rel L1 vs oracle", err)
    assert err[0] <= 1e-10 and err[1] <= 1e-10 and err[3] <= 1e-10, err
    # v is near zero over most of the domain: its error against the speed
    speed = np.abs(want[1]).sum() + np.abs(want[2]).sum()
    assert np.abs(got[2] - want[2]).sum() <= 1e-10 * speed
