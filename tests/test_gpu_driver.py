"""The drop-in driver executable (euler_compare_amd) behind the reference's
inputs syntax: it must reproduce the reference's Sod row (347 steps) and
write the final CSV of the oracle's state."""
import csv
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "amrex-euler-allmach-explicit-imex_amd" / "euler_compare_amd"

# The reference's smoke inputs layout (key = value lines, '#' comments).
INPUTS = """# smoke inputs
max_step = 8
stop_time = 0.05
amr.n_cell = 64 64
amr.max_grid_size = 32
amr.plot_int = 8
geometry.prob_lo = 0.0 0.0
geometry.prob_hi = 1.0 1.0
geometry.is_periodic = 1 1
euler.gamma = 1.4
euler.cfl = 0.4
"""


def grid_field(rows, col, nx, ny, dx=None, dy=None, xlo=0.0, ylo=0.0):
    """A CSV column as an (ny, nx) field, each row placed by its own i, j (or
    x, y) columns: the writers visit the cells box by box (amr.max_grid_size),
    as the reference's MFIter loops do."""
    f = np.full((ny, nx), np.nan)
    for r in rows:
        if "i" in r:
            i, j = int(r["i"]), int(r["j"])
        else:
            i = int(round((float(r["x"]) - xlo) / dx - 0.5))
            j = int(round((float(r["y"]) - ylo) / dy - 0.5))
        f[j, i] = float(r[col])
    assert not np.isnan(f).any()
    return f


def run(tmp_path, *overrides):
    inp = tmp_path / "inputs"
    inp.write_text(INPUTS)
    r = subprocess.run([str(EXE), str(inp), *overrides], capture_output=True, text=True, timeout=120,
                       cwd=tmp_path)
    kv = dict(line.split("=", 1) for line in r.stdout.splitlines() if "=" in line)
    return r.returncode, kv, r.stderr


def test_sod_row(tmp_path, ecor):
    out_csv = tmp_path / "sod.csv"
    rc, kv, err = run(tmp_path, "amr.n_cell=400 4", "geometry.prob_hi=1 0.01",
                      "geometry.is_periodic=0 1", "euler.problem=toro1", "prob.test=1",
                      "euler.cfl=0.8", "stop_time=0.2", "max_step=100000", "euler.method=explicit",
                      "euler.spatial_order=2", "euler.slope_limiter=minmod", "euler.riemann=hllc",
                      f"euler.final_csv={out_csv}")
    assert rc == 0, err
    assert kv["project_amrex_euler_compare_status"] == "ok"
    assert kv["completed_steps"] == "347"
    assert abs(float(kv["mass_drift"])) < 1e-15
    g = ecor.make_grid(400, 4, 1 / 400, 0.01 / 4, riemann=ecor.HLLC, bc=(1, 1, 0, 0))
    ref, res, _ = ecor.run_explicit(g, ecor.init_toro(400, 4, [1, 0, 0, 1], [0.125, 0, 0, 0.1]),
                                    0.8, 0.2, 100000)
    rows = list(csv.DictReader(open(out_csv)))
    assert len(rows) == 1600
    rho = grid_field(rows, "rho", 400, 4)
    assert np.array_equal(rho, ref[0])
    # box by box, as the reference's MFIter loop writes them: 400 cells
    # across with max_grid_size 32 gives 32-wide boxes, so the row after
    # (31, 0) is (0, 1)
    assert [(int(r["i"]), int(r["j"])) for r in rows[31:33]] == [(31, 0), (0, 1)]


def test_capped_and_failure_rows(tmp_path):
    rc, kv, _ = run(tmp_path, "euler.problem=gresho_vortex", "euler.vortex_center=0.5 0.5",
                    "euler.mach=0.01", "euler.pressure=7142.857142857143", "euler.riemann=xie_am_hllc_p",
                    "euler.spatial_order=2", "max_step=5", "stop_time=1.0")
    assert rc == 0 and kv["project_amrex_euler_compare_status"] == "capped_max_step"
    assert kv["completed_steps"] == "5" and kv["riemann"] == "xie_am_hllc_p"
    rc, kv, _ = run(tmp_path, "euler.problem=gresho_vortex", "euler.pressure=-1.0", "max_step=5")
    assert rc == 2 and kv["failure_category"] == "initial_nonphysical_state"
    r = subprocess.run([str(EXE), "euler.riemann=bogus"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Unknown euler.riemann=bogus" in r.stderr


def test_imex_row(tmp_path):
    rc, kv, err = run(tmp_path, "amr.n_cell=64 64", "euler.problem=gresho_vortex",
                      "euler.vortex_center=0.5 0.5", "euler.mach=0.001",
                      "euler.pressure=714285.7142857143", "euler.method=imex", "euler.spatial_order=2",
                      "euler.imex_acoustic_startup=0", "euler.imex_solver_tol=1e-10",
                      "euler.imex_solver_max_iter=1000", "max_step=3", "stop_time=1.0",
                      "euler.imex_gmres_precision=mixed")
    assert rc == 0, err
    assert kv["imex_step_count"] == "3"
    assert kv["imex_solver_status"] in ("converged_absolute", "converged_relative")
    assert float(kv["imex_solver_relative_residual_max"]) <= 1e-10
    # the closeout's evaluation keys (app/driver_main.hpp:451-544)
    assert kv["imex_picard_pressure_solves_used"] == kv["imex_kinetic_lag_update_count"] == "2"
    assert kv["imex_timestep_rule"] == "material_speed_eq79_zdrop_2d"
    assert kv["imex_acoustic_startup_steps"] == kv["imex_acoustic_cap_steps"] == "0"
    assert kv["imex_nonfinite_count_last"] == "0"
    assert kv["imex_trial_density_min"] == kv["rho_min"]     # both of the final state
    pmin, pmax = float(kv["imex_solved_pressure_min"]), float(kv["imex_solved_pressure_max"])
    assert 714285.0 < pmin <= pmax < 714287.0
    assert float(kv["imex_trial_pressure_min"]) == pytest.approx(pmin, rel=1e-9)
    assert kv["imex_energy_closeout_delta_linf_max"] == kv["imex_energy_correction_linf_max"]
    for k in ("imex_momentum_correction_linf_last", "imex_eq78_divergence_linf_last",
              "imex_eq78_divergence_linf_max"):
        assert 0.0 < float(kv[k]) < 1e3, k
    assert kv["imex_source_space_order"] == "second_order_space_s2"
    # IMEX rows name the reference's host solve path in the inputs
    rc, _, err = run(tmp_path, "euler.method=imex", "euler.spatial_order=2", "euler.imex_acoustic_startup=0",
                     "euler.imex_pressure_solver_path=eigen")
    assert rc != 0 and "imex_pressure_solver_path" in err


def test_imex_exact_dirichlet_row_equals_outflow(tmp_path):
    """euler.method=imex with euler.field_boundary=exact_dirichlet on a
    non-periodic Gresho box: the source-map step clamps the physical ghosts
    before its star update for every problem but the shock bubble
    (imex/imex_source_map_step.hpp:405-409), so the run and its final CSV
    equal the outflow run's."""
    outs = []
    for fb in ("exact_dirichlet", "outflow"):
        out_csv = tmp_path / f"{fb}.csv"
        rc, kv, err = run(tmp_path, "amr.n_cell=48 40", "geometry.prob_hi=1.0 0.8",
                          "geometry.is_periodic=0 0", "euler.problem=gresho_vortex",
                          "euler.vortex_center=0.4 0.35", "euler.mach=0.01",
                          "euler.pressure=7142.857142857143", "euler.method=imex",
                          "euler.spatial_order=2", "euler.imex_acoustic_startup=0",
                          "euler.imex_solver_tol=1e-10", "euler.imex_solver_max_iter=2000",
                          "max_step=3", "stop_time=1.0", f"euler.field_boundary={fb}",
                          f"euler.final_csv={out_csv}")
        assert rc == 0, err
        assert kv["imex_step_count"] == "3", kv
        assert kv["field_boundary"] == fb
        outs.append(out_csv.read_text())
    assert outs[0] == outs[1]


def test_toro_exact_dirichlet_row(tmp_path, ecor):
    """A Riemann-strip row of the convergence study: geometry.is_periodic=0 0,
    euler.field_boundary=exact_dirichlet, 5 cells across (the device loop with
    the exact Toro ghost states at each step's time)."""
    out_csv = tmp_path / "toro.csv"
    nx, ny = 200, 5
    rc, kv, err = run(tmp_path, f"amr.n_cell={nx} {ny}", "geometry.prob_hi=1.0 0.1",
                      "geometry.is_periodic=0 0", "euler.problem=toro1", "prob.test=1",
                      "euler.field_boundary=exact_dirichlet", "euler.cfl=0.4", "stop_time=0.25",
                      "max_step=100000", "euler.riemann=xie_am_hllc_p", "euler.spatial_order=2",
                      f"euler.final_csv={out_csv}")
    assert rc == 0, err
    assert kv["project_amrex_euler_compare_status"] == "ok"
    assert kv["field_boundary"] == "exact_dirichlet"
    L, R = [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1]
    e = ecor.make_exact(ecor.EXACT_TORO, (0, 0), prob_hi=(1.0, 0.1), ql=L, qr=R, interface_x=0.5)
    g = ecor.set_exact(ecor.make_grid(nx, ny, 1 / nx, 0.1 / ny, riemann=ecor.XIE_HLLCP,
                                      bc=(1, 1, 1, 1)), e, 0.0)
    ref, res, _ = ecor.run_explicit(g, ecor.init_toro(nx, ny, L, R), 0.4, 0.25, 100000)
    assert int(kv["completed_steps"]) == res.steps
    rows = list(csv.DictReader(open(out_csv)))
    rho = grid_field(rows, "rho", nx, ny)
    # the fan's ghost states take the device pow(): relative 1e-12 in L1
    assert np.abs(rho - ref[0]).sum() <= 1e-12 * np.abs(ref[0]).sum()
    # the exact boundary differs from outflow (the y ghost rows)
    g0 = ecor.make_grid(nx, ny, 1 / nx, 0.1 / ny, riemann=ecor.XIE_HLLCP, bc=(1, 1, 1, 1))
    out0, _, _ = ecor.run_explicit(g0, ecor.init_toro(nx, ny, L, R), 0.4, 0.25, 100000)
    assert not np.array_equal(out0[0], ref[0])


def test_direct_bdltv20_toro_row(tmp_path, ecor):
    """A Riemann-convergence IMEX row (euler.bdltv20_paper_t1_s2=
    toro_xy_exact_dirichlet_2d, fixed dt, 4 Picard solves): the driver's
    report keys, final CSV schema and states against the oracle's direct
    step run through the same loop (solves to 1e-12: states to 1e-9)."""
    out_csv = tmp_path / "direct.csv"
    nx, ny, t_end, nsteps = 64, 5, 0.1, 20
    dt = t_end / nsteps
    rc, kv, err = run(tmp_path, f"amr.n_cell={nx} {ny}", "geometry.prob_hi=1.0 0.078125",
                      "geometry.is_periodic=0 0", "euler.problem=toro1", "prob.test=1",
                      "euler.field_boundary=exact_dirichlet", "euler.method=imex",
                      "euler.spatial_order=2", "euler.imex_acoustic_startup=0",
                      "euler.bdltv20_paper_t1_s2=toro_xy_exact_dirichlet_2d",
                      f"euler.bdltv20_paper_t1_s2_dt={dt!r}", "euler.imex_picard_iterations=4",
                      "euler.imex_solver_tol=1e-12", "euler.imex_solver_max_iter=4000",
                      f"stop_time={t_end!r}", f"euler.final_csv={out_csv}")
    assert rc == 0, err
    assert kv["bdltv20_paper_t1_s2_status"] == "passed"
    assert int(kv["bdltv20_paper_steps"]) == nsteps
    assert kv["bdltv20_paper_boundary_policy"] == "xy_exact_dirichlet_paper_2d_reduction"
    L, R = [1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1]
    e = ecor.make_exact(ecor.EXACT_TORO, (0, 0), prob_hi=(1.0, 0.078125), ql=L, qr=R, interface_x=0.5)
    g = ecor.make_grid(nx, ny, 1 / nx, 0.078125 / ny)
    d = ecor.make_direct(e, 1.0, True)
    U = ecor.init_toro(nx, ny, L, R)
    t = 0.0
    for _ in range(nsteps):
        h = ecor.direct_dt(g, d, U, 0.8, dt, t, t_end)
        U, _, res = ecor.direct_step(g, d, U, t, h, picard=4, tol=1e-12, max_iter=4000)
        assert res.solver_failures == 0
        t += h
    rows = list(csv.DictReader(open(out_csv)))
    assert list(rows[0].keys()) == ["x", "y", "rho", "u", "v", "p", "exact_rho", "exact_u",
                                    "exact_v", "exact_p"]
    dx_, dy_ = 1 / nx, 0.078125 / ny
    rho = grid_field(rows, "rho", nx, ny, dx_, dy_)
    assert np.abs(rho - U[0]).max() <= 1e-9 * np.abs(U[0]).max()
    xs = (np.arange(nx) + 0.5) * dx_
    ex = np.array([ecor.exact_riemann_sample(x, t, L, R)[0] for x in xs])
    assert np.array_equal(grid_field(rows, "exact_rho", nx, ny, dx_, dy_)[0], ex)
    l1 = np.abs(rho - ex[None, :]).mean()
    assert float(kv["bdltv20_paper_density_l1_error"]) == pytest.approx(l1, rel=1e-9)


def test_shock_bubble_snapshots_and_summary(tmp_path):
    """The shock-density bubble row's on-disk outputs (io/output_cases.hpp:
    408-592): snapshot CSVs named by index, time and step at the requested
    times, the summary CSV, and the report keys -- the summary's extrema,
    centroid and schlieren maximum (device reductions) agree with the last
    snapshot's rows (host values)."""
    import math
    snap = tmp_path / "snaps"
    summary = tmp_path / "summary.csv"
    rc, kv, err = run(tmp_path, "amr.n_cell=160 40", "euler.problem=shock_density_bubble_2d",
                      "euler.riemann=hllc", "euler.spatial_order=2", "euler.cfl=0.4", "stop_time=0.06",
                      "max_step=100000", "euler.shock_density_bubble_snapshot_times=0,0.03,0.06",
                      f"euler.shock_density_bubble_snapshot_dir={snap}", f"euler.final_csv={summary}")
    assert rc == 0, err
    assert kv["snapshot_status"] == "ok" and kv["snapshots_written"] == "3"
    assert kv["scheme_name"] == "explicit_o2_hllc" and kv["order"] == "O2"
    files = sorted(p.name for p in snap.iterdir())
    assert len(files) == 3 and files[0] == "snapshot_00_t0p000000_step000000.csv"
    assert files[1].startswith("snapshot_01_t0p030000_step") and files[2].startswith("snapshot_02_t0p060000_step")
    rows = list(csv.DictReader(open(snap / files[2])))
    assert len(rows) == 160 * 40
    s = list(csv.DictReader(open(summary)))[0]
    assert s["status"] == "ok" and int(s["steps_accepted"]) == int(files[2][-10:-4])
    rho = np.array([float(r["rho"]) for r in rows])
    assert float(s["rho_min"]) == rho.min() and float(s["rho_max"]) == rho.max()
    schl = np.array([float(r["schlieren"]) for r in rows])
    assert float(s["schlieren_max"]) == pytest.approx(schl.max(), rel=1e-15)
    ind = np.array([int(r["density_bubble_indicator"]) for r in rows]) == 1
    x = np.array([float(r["x"]) for r in rows]); y = np.array([float(r["y"]) for r in rows])
    assert ind.any()
    assert float(s["density_bubble_centroid_x"]) == pytest.approx(x[ind].mean(), rel=1e-12)
    assert float(s["density_bubble_centroid_y"]) == pytest.approx(y[ind].mean(), rel=1e-12)
    assert not math.isnan(float(kv["density_bubble_centroid_x"]))


def test_gresho_final_csv_schema(tmp_path, ecor):
    """write_gresho_final_csv (io/output_cases.hpp:16-71): the column set, the
    exact Gresho state and errors per cell, the run labels, and NaN trial
    minima for an explicit run."""
    out_csv = tmp_path / "gresho.csv"
    n, mach = 32, 0.1
    rc, kv, err = run(tmp_path, f"amr.n_cell={n} {n}", "euler.problem=gresho_vortex",
                      "euler.vortex_center=0.5 0.5", f"euler.mach={mach}",
                      f"euler.pressure={1 / (1.4 * mach * mach)!r}", "euler.density_outer=1",
                      "euler.riemann=xie_am_hllc_p", "euler.spatial_order=2", "stop_time=0.01",
                      "max_step=1000", f"euler.final_csv={out_csv}")
    assert rc == 0, err
    rows = list(csv.DictReader(open(out_csv)))
    assert list(rows[0].keys()) == [
        "x", "y", "rho", "u", "v", "pressure", "exact_rho", "exact_u", "exact_v", "exact_pressure",
        "density_error", "velocity_error", "pressure_error", "pressure_perturbation_error", "time", "step",
        "mach", "method", "riemann", "imex_form", "imex_predictor_flux", "imex_predictor_dissipation",
        "spatial_order", "slope_limiter", "imex_trial_density_run_min", "imex_high_trial_pressure_run_min",
        "imex_high_trial_internal_energy_run_min"]
    r0 = rows[0]
    assert (r0["method"], r0["riemann"], r0["imex_form"], r0["slope_limiter"]) == \
        ("explicit", "xie_am_hllc_p", "not_applicable", "minmod")
    assert r0["imex_trial_density_run_min"] == "nan"
    U0 = ecor.init_gresho(n, n, mach=mach, p0=1 / (1.4 * mach * mach))
    p_ex = grid_field(rows, "exact_pressure", n, n, 1 / n, 1 / n)
    p_init = 0.4 * (U0[3] - 0.5 * (U0[1] ** 2 + U0[2] ** 2) / U0[0])
    assert np.abs(p_ex - p_init).max() <= 1e-13 * p_init.max()
    # the errors are the differences of the columns
    e = np.array([float(r["pressure_error"]) for r in rows])
    assert np.array_equal(e, np.array([float(r["pressure"]) - float(r["exact_pressure"]) for r in rows]))
    assert float(kv["kinetic_energy_ratio"]) < 1.0


def read_plotfile(path):
    """Minimal reader of a single-level AMReX plotfile (the Header, the VisMF
    header Level_0/Cell_H and native-double FABs in Cell_D files)."""
    h = (path / "Header").read_text().split("\n")
    nvar = int(h[1])
    names = h[2:2 + nvar]
    time = float(h[2 + nvar + 1])
    dom = h[2 + nvar + 6].strip()
    lo, hi = [tuple(int(v) for v in t.strip("()").split(",")) for t in dom.strip("()").split(") (")[:2]]
    nx, ny = hi[0] + 1, hi[1] + 1
    ch = (path / "Level_0" / "Cell_H").read_text().split("\n")
    nbox = int(ch[4][1:].split()[0])
    fods = [ln.split() for ln in ch[4 + nbox + 3: 4 + nbox + 3 + nbox]]
    field = np.full((nvar, ny, nx), np.nan)
    raw = {}
    for _, fname, off in fods:
        data = raw.setdefault(fname, (path / "Level_0" / fname).read_bytes())
        off = int(off)
        end = data.index(b"\n", off)
        head = data[off:end].decode()
        assert head.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))")
        box = head[head.index(")))") + 3:].rsplit(" ", 1)
        ncomp = int(box[1])
        (i0, j0), (i1, j1) = [tuple(int(v) for v in t.split(",")) for t in
                              box[0].strip("()").split(") (")[:2]]
        cnt = (i1 - i0 + 1) * (j1 - j0 + 1)
        vals = np.frombuffer(data, dtype="<f8", count=cnt * ncomp, offset=end + 1)
        field[:, j0:j1 + 1, i0:i1 + 1] = vals.reshape(ncomp, j1 - j0 + 1, i1 - i0 + 1)
    return names, time, field, ch


def test_plotfiles(tmp_path):
    """write_plotfile (io/output_cases.hpp:2-14): plotfiles at step 0, every
    plot_int steps and at the end, over the max_grid_size box layout; the
    last one holds the final state of the CSV (bit for bit) and its VisMF
    min/max records match the data."""
    out_csv = tmp_path / "sod.csv"
    rc, kv, err = run(tmp_path, "amr.n_cell=80 8", "amr.max_grid_size=32", "amr.plot_int=5",
                      "geometry.prob_hi=1 0.1", "geometry.is_periodic=0 1", "euler.problem=toro1",
                      "prob.test=1", "euler.cfl=0.8", "stop_time=0.05", "max_step=100000",
                      "euler.riemann=hllc", "euler.spatial_order=2", f"euler.final_csv={out_csv}")
    assert rc == 0, err
    steps = int(kv["completed_steps"])
    plts = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("plt"))
    want = sorted({f"plt{s:05d}" for s in list(range(0, steps + 1, 5)) + [steps]})
    assert plts == want
    names, t, field, ch = read_plotfile(tmp_path / plts[-1])
    assert names == ["rho", "momx", "momy", "E", "pressure"]
    assert t == pytest.approx(0.05, abs=1e-15)
    # BoxList::maxSize halves 80 and 32 together to 5 blocks of 2, then
    # splits 5 into 2 + 2 + 1: boxes 32, 32 and 16 cells wide
    assert ch[4].startswith("(3 0") and ch[5] == "((0,0) (31,7) (0,0))" and ch[7] == "((64,0) (79,7) (0,0))"
    rows = list(csv.DictReader(open(out_csv)))
    rho = grid_field(rows, "rho", 80, 8)
    p = grid_field(rows, "p", 80, 8)
    assert np.array_equal(field[0], rho) and np.array_equal(field[4], p)
    nb = 3
    mins = [[float(v) for v in ln.rstrip(",").split(",")] for ln in ch[4 + nb + 3 + nb + 2: 4 + nb + 3 + nb + 2 + nb]]
    assert mins[0][0] == field[0][:, :32].min() and mins[2][4] == field[4][:, 64:].min()


def test_plot_interval_longer_than_a_device_chunk(tmp_path):
    """amr.plot_int beyond the device loop's 512-step chunks and not a
    multiple of them: a plotfile at every multiple of plot_int (write_plotfile
    at step % plot_int == 0, io/output_cases.hpp:8), at step 0 and at the
    end."""
    rc, kv, err = run(tmp_path, "amr.n_cell=64 4", "amr.plot_int=600", "geometry.prob_hi=1 0.0625",
                      "geometry.is_periodic=1 1", "euler.problem=gresho_vortex",
                      "euler.vortex_center=0.5 0.03125", "euler.mach=0.1", "euler.riemann=rusanov",
                      "euler.spatial_order=1", "stop_time=100.0", "max_step=1300")
    assert rc == 0, err
    assert kv["completed_steps"] == "1300"
    plts = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("plt"))
    assert plts == ["plt00000", "plt00600", "plt01200", "plt01300"]
