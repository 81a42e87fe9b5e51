#!/usr/bin/env python3
"""Throughput of the explicit O2 HLLC-P step on one MI355X.

Workload (the one the headline metric is quoted on): Gresho vortex, 2048^2
periodic on [0,1]^2, FP64, gamma=1.4, rho=1, centre (0.5,0.5),
p0 = 1/(gamma M^2) at M = 0.01, explicit O2 MUSCL/minmod with the Xie
AM-HLLC-P flux, CFL 0.4 (the reference's default euler.cfl).  A step is one
pass of the device-resident driver loop: dt from the previous step's fused
CFL reduction, the fused step kernel, its diagnostics epilogue.  The state
is resident in HBM before timing starts.

Prints one JSON line (rank 0).  --gpus N with N > 1 exits: this build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "amrex-euler-allmach-explicit-imex_amd"
sys.path.insert(0, str(PKG))

METRIC = "cell-updates/s & % HBM roofline (HLLC-P O2, IMEX step); adjoint/forward cost"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, vendor figure
BYTES_PER_CELL_UPDATE = 64     # read U^n (4 x 8 B) + write U^{n+1} (4 x 8 B)
TRAFFIC_FILE = ROOT / "profiles" / "explicit_step_xie_hllcp_gresho2048_traffic.json"


def roofline_sweep(dev, riemann: str, order: int, mach: float, cfl: float) -> dict:
    """The roofline sweep of the Gresho workload at 256^2 and 1024^2 (2048^2 is
    the main line): throughput and the step kernel's event-timed duration."""
    import torch
    import ecamd
    from ecamd import problems
    out = {}
    for n, steps in ((256, 400), (1024, 200)):
        grid = ecamd.Grid(n, n, 1.0 / n, 1.0 / n, 1.4, riemann, "minmod", order)
        s = ecamd.ExplicitSolver(grid, dev, stream=torch.cuda.current_stream(dev))
        A = torch.from_numpy(problems.gresho(n, mach)).to(dev)
        B = torch.empty_like(A)
        run = s.new_run(cfl, 1.0, 2**30)
        if s.advance(A, B, 20, run) is B:
            A, B = B, A
        torch.cuda.synchronize()
        s.arm_kernel_timing(steps)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        s.advance(A, B, steps, run)
        b.record()
        torch.cuda.synchronize()
        k_ms, k_n = s.kernel_time()
        s.disarm_kernel_timing()
        per = a.elapsed_time(b) / 1e3 / steps
        kern = k_ms / 1e3 / max(k_n, 1)
        out[f"{n}x{n}"] = {"ms_per_step": per * 1e3, "cell_updates_per_s": n * n / per,
                           "kernel_us_avg": kern * 1e6,
                           "hbm_frac": BYTES_PER_CELL_UPDATE * n * n / kern / 1e9 / HBM_PEAK_GBS}
        s.close()
    return out


def cpu_baseline(n: int, mach: float, cfl: float, steps: int) -> dict:
    """The oracle (serial CPU restatement) on the same workload, first `steps` steps."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import ecor  # noqa: E402  (oracle: CPU baseline leg only)
    g = ecor.make_grid(n, n, 1.0 / n, 1.0 / n, gamma=1.4, riemann=ecor.XIE_HLLCP)
    U = ecor.init_gresho(n, n, mach=mach)
    t0 = time.perf_counter()
    _, res, _ = ecor.run_explicit(g, U, cfl, 1.0, steps)
    dt = time.perf_counter() - t0
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": n * n * res.steps / dt, "unit": "cell-updates/s", "cores": 1,
            "kind": "port",
            "sample": f"first {res.steps} steps of the same {n}x{n} workload, serial oracle, "
                      f"{dt:.2f} s on 1 of {os.cpu_count()} host threads ({cpu}); "
                      "includes the per-step dt and diagnostics passes like the reference's "
                      "wall_time_sec"}


# VALU wave instructions per launch of the HLLC-P O2 step at 2048^2 (16x15
# tiles), from the committed rocprofv3 SQ counter summary.
VALU_FILE = ROOT / "profiles" / "explicit_step_xie_hllcp_gresho2048_valu.json"

# The sources the explicit step kernel is compiled from: a counter figure is
# reported as current only when it was measured on these exact bytes.
KERNEL_SOURCES = ("ec_explicit.hip", "ec_tile.hpp", "ec_flux_window.hpp", "ec_exact_ops.hpp",
                  "ec_state.hpp", "ec_ghost.hpp", "ec_wave.hpp", "ec_context.hpp")


def kernel_sources_digest() -> str:
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        h.update(name.encode() + b"\0" + (PKG / "csrc" / name).read_bytes())
    return h.hexdigest()[:16]


def profile_figure(path: Path, key: str, digest: str):
    """A per-launch counter figure from a committed profile summary, with the
    sources digest and commit it was measured on.  A figure measured on other
    sources is stale: its value is withheld (None) and reported beside."""
    if not path.exists():
        return None, None
    try:
        d = json.loads(path.read_text())
    except (OSError, ValueError):
        return None, None
    prov = {"file": str(path.relative_to(ROOT)), "kernel_sources": d.get("kernel_sources"),
            "commit": d.get("commit"), "current": d.get("kernel_sources") == digest}
    if not prov["current"]:
        prov["stale_value"] = d.get(key)
        return None, prov
    return d.get(key), prov


def valu_bound(riemann: str, order: int, n: int, kernel_s: float, digest: str):
    """The FP64 VALU issue bound of the step kernel: counted VALU wave
    instructions x 4 cycles (a wave64 instruction on a 16-lane SIMD) over the
    1024 SIMDs at 2.4 GHz, against the measured kernel time."""
    if not (n == 2048 and riemann == "xie_am_hllc_p" and order == 2):
        return None
    valu, prov = profile_figure(VALU_FILE, "valu_wave_insts_per_launch", digest)
    if prov is None:
        return None
    if valu is None:
        return {"valu_wave_insts_per_launch": None, "profile": prov}
    issue_s = valu * 4 / (1024 * 2.4e9)
    return {"valu_wave_insts_per_launch": valu, "issue_bound_us": issue_s * 1e6,
            "frac_of_issue_bound": issue_s / kernel_s, "profile": prov}


def _counter_means(run_dir: Path, kernel_substr: str) -> dict:
    """Per-dispatch means of every counter of one rocprofv3 --pmc run over the
    dispatches whose kernel name contains kernel_substr."""
    import csv
    files = sorted(run_dir.rglob("*counter_collection.csv"))
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {run_dir}")
    acc = {}
    for r in csv.DictReader(open(files[0])):
        if kernel_substr in r["Kernel_Name"]:
            acc.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def summarize_counters(spec: list, commit: str) -> None:
    """Writes the two profile summaries bench.py reads (VALU counts; HBM
    traffic with the copy calibration) from rocprofv3 --pmc run directories:
    spec holds sq=DIR (SQ_WAVES, SQ_INSTS_VALU, ...), fetch=DIR, write=DIR
    (the step), copy_fetch=DIR, copy_write=DIR (bench.py --mode copy)."""
    dirs = dict(item.split("=", 1) for item in spec)
    digest = kernel_sources_digest()
    kern = "explicit_step_kernel<16, 15, 256, 3, 2>"
    sq, passes = {}, []
    for i, d in enumerate(dirs["sq"].split(",")):   # one directory per counter pass
        m = _counter_means(Path(d), kern)
        sq.update(m)
        passes.append(sorted(m))
    wl = "gresho_vortex_2048x2048_periodic_M0.01_explicit_o2_minmod_xie_am_hllc_p_cfl0.4"
    base = "python3 bench.py --steps 10 --warmup 2 --no-cpu-baseline --no-extras"
    valu = {"kernel": "explicit_step_kernel<16,15,256,3,2>", "workload": wl,
            "kernel_sources": digest, "commit": commit,
            "valu_wave_insts_per_launch": round(sq["SQ_INSTS_VALU"]),
            "waves_per_launch": round(sq["SQ_WAVES"]),
            "counters_per_dispatch": {k: round(v, 1) for k, v in sorted(sq.items())},
            "commands": [f"rocprofv3 --pmc {' '.join(c)} -d gpurun_out/pmc_sq{i} -o p "
                         f"--output-format csv -- {base}" for i, c in enumerate(passes)]}
    VALU_FILE.write_text(json.dumps(valu, indent=2) + "\n")
    fetch = _counter_means(Path(dirs["fetch"]), kern)["FETCH_SIZE"]
    write = _counter_means(Path(dirs["write"]), kern)["WRITE_SIZE"]
    cf = _counter_means(Path(dirs["copy_fetch"]), "copyBuffer")["FETCH_SIZE"]
    cw = _counter_means(Path(dirs["copy_write"]), "copyBuffer")["WRITE_SIZE"]
    copy_kib = 4 * 2048 * 2048 * 8 / 1024
    ff, wf = copy_kib / cf, copy_kib / cw
    traffic = {"kernel": valu["kernel"], "workload": wl, "kernel_sources": digest, "commit": commit,
               "fetch_size_kib_per_dispatch": round(fetch, 1),
               "write_size_kib_per_dispatch": round(write, 1),
               "fetch_factor": round(ff, 4), "write_factor": round(wf, 4),
               "calibration": f"device copy of {int(copy_kib * 1024)} B: FETCH_SIZE {cf:.1f} KiB, "
                              f"WRITE_SIZE {cw:.1f} KiB per dispatch",
               "hbm_bytes_per_launch": round((fetch * ff + write * wf) * 1024),
               "algorithmic_bytes_per_launch": BYTES_PER_CELL_UPDATE * 2048 * 2048,
               "commands": [f"rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_fetch -o p --output-format csv -- {base}",
                            f"rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pmc_write -o p --output-format csv -- {base}",
                            "rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_copy_fetch -o p --output-format csv "
                            "-- python3 bench.py --mode copy --steps 5",
                            "rocprofv3 --pmc WRITE_SIZE -d gpurun_out/pmc_copy_write -o p --output-format csv "
                            "-- python3 bench.py --mode copy --steps 5"]}
    TRAFFIC_FILE.write_text(json.dumps(traffic, indent=2) + "\n")
    print(json.dumps({"valu": str(VALU_FILE), "traffic": str(TRAFFIC_FILE), "kernel_sources": digest}))


def copy_calibration(dev, n: int, steps: int) -> int:
    """Device-to-device copies of one n x n state (4 planes of FP64): a plain
    copy kernel of known size, so that FETCH_SIZE / WRITE_SIZE per dispatch can
    be scaled to bytes."""
    import torch
    a = torch.ones((4, n, n), dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        b.copy_(a)
    end.record()
    torch.cuda.synchronize()
    s = start.elapsed_time(end) / 1e3 / steps
    nbytes = a.numel() * 8
    print(json.dumps({"mode": "copy", "bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes,
                      "copies": steps, "us_per_copy": s * 1e6, "GB_per_s": 2 * nbytes / s / 1e9}))
    return 0


def stream_copy_gbs(dev, nbytes: int = 1 << 30, copies: int = 10, reps: int = 5) -> float:
    """Achievable HBM bandwidth on this device: a STREAM-style copy kernel
    (torch's elementwise b = 1.0 * a, so that no DMA engine takes the copy) over
    a 1 GiB buffer, four times the 256 MB Infinity Cache, so that it streams
    from and to HBM.  `copies` back-to-back launches per timed region amortise
    the launch cost; read + write bytes over the time per copy, median of
    `reps` regions."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        a0, a1 = _events()
        a0.record()
        for _ in range(copies):
            torch.mul(a, 1.0, out=b)
        a1.record()
        torch.cuda.synchronize()
        times.append(a0.elapsed_time(a1) / 1e3 / copies)
    s = sorted(times)[len(times) // 2]
    del a, b
    return 2 * nbytes / s / 1e9


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_imex(dev, n: int = 512, mach: float = 0.001, reps: int = 3, max_iter: int = 20000,
               oracle_state=None) -> dict:
    """IMEX T1/S2 BDLTV20 source-map Picard step, Gresho n x n periodic at
    Mach `mach`: device GMRES(30) to 1e-10 relative residual in three forms,
    same dt, same initial state:
      fp64           restarted GMRES(30), CGS2, all FP64;
      fp64_flexible  FP64 flexible GMRES(20) over a one-reduction FP64
                     GMRES(30) inner solve (the mixed form's structure and
                     kernels in FP64: the comparator that isolates precision);
      mixed          the same with the inner solve in FP32.
    Time per step, the differences of the results, and (when the CPU leg ran)
    each result against the oracle's step at 1e-8."""
    import torch
    import ecamd
    from ecamd import problems
    grid = ecamd.Grid(n, n, 1.0 / n, 1.0 / n, 1.4, "hllc", "minmod", 2)
    s = ecamd.ExplicitSolver(grid, dev)
    U0 = torch.from_numpy(problems.gresho(n, mach)).to(dev)
    dt, _, _ = s.imex_dt(U0, 0.8, 0.0, 1.0)
    out = {"workload": f"gresho_vortex_{n}x{n}_periodic_M{mach:g}_imex_t1s2_bdltv20_picard2",
           "dt": dt, "tol": 1e-10, "restart": 30}
    results = {}
    for prec in ("fp64", "fp64_flexible", "mixed"):
        opts = s.imex_opts(precision=prec, tol=1e-10, max_iter=max_iter, restart=30)
        U = U0.clone()
        s.imex_step(U, dt, opts)            # warm-up
        times = []
        for _ in range(reps):
            U = U0.clone()
            torch.cuda.synchronize()
            a, b = _events()
            a.record()
            d = s.imex_step(U, dt, opts)
            b.record()
            torch.cuda.synchronize()
            times.append(a.elapsed_time(b))
        results[prec] = U
        out[prec] = {"ms_per_step": sorted(times)[len(times) // 2], "status": ecamd.IMEX_STATUS[d.status],
                     "gmres_iterations": d.iterations, "refinements": d.refinements,
                     "relative_residual_linf": d.relative_residual_linf}
    def rel(a, b):
        return [float(x) for x in ((a - b).abs().reshape(4, -1).sum(1) /
                                   b.abs().reshape(4, -1).sum(1)).cpu()]

    out["mixed_vs_fp64_rel_l1"] = rel(results["mixed"], results["fp64"])
    out["fp64_flexible_vs_fp64_rel_l1"] = rel(results["fp64_flexible"], results["fp64"])
    t = {k: out[k]["ms_per_step"] for k in ("fp64", "fp64_flexible", "mixed")}
    out["speedup_mixed_over_fp64"] = t["fp64"] / t["mixed"]
    out["speedup_mixed_over_fp64_flexible"] = t["fp64_flexible"] / t["mixed"]
    out["speedup_fp64_flexible_over_fp64"] = t["fp64"] / t["fp64_flexible"]
    out["speedup_mixed_over_fastest_fp64"] = min(t["fp64"], t["fp64_flexible"]) / t["mixed"]
    if oracle_state is not None:
        ref = torch.from_numpy(oracle_state["state"]).to(dev)
        out["vs_oracle"] = {"oracle": oracle_state["label"],
                            **{f"{k}_rel_l1": rel(results[k], ref) for k in results}}
    s.close()
    return out


def cpu_imex(n: int = 512, mach: float = 0.001) -> dict:
    """The oracle's IMEX step (the reference's unrestarted host GMRES,
    restated) on the same 512^2 state, at tol 1e-8, the tightest it reaches in
    seconds: its serial time (a CPU baseline for the IMEX row) and its state,
    the reference result the device forms are compared with."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import ecor  # noqa: E402  (oracle: CPU baseline leg only)
    from ecamd import problems
    U = problems.gresho(n, mach)
    g = ecor.make_grid(n, n, 1.0 / n, 1.0 / n)
    dt, _, _ = ecor.imex_dt(g, U, 0.8, 0.0, 1.0)
    t0 = time.perf_counter()
    ref, _, r = ecor.imex_step(g, U, dt, picard=2, tol=1e-8, max_iter=1000)
    sec = time.perf_counter() - t0
    return {"state": ref, "label": f"oracle step at tol 1e-8 ({r.iterations} GMRES iterations, "
                                    f"relative residual {r.relative_residual_linf:.2e})",
            "seconds_per_step": sec, "dt": dt}


def bench_direct(dev, n: int = 256, mach: float = 0.01, reps: int = 3, cpu_n: int = 48) -> dict:
    """Direct BDLTV20 T1/S2 step (the report rows' IMEX driver), Gresho n x n
    with exact-Dirichlet boundaries at Mach `mach`, epsilon 1, 4 Picard
    solves to 1e-8 (the settings of the reference's Gresho matrix), eq79 dt:
    time per step on the device in FP64 and mixed form, and one step of the
    oracle on the host at cpu_n x cpu_n beside the device at the same size."""
    import time as _time
    import torch
    import ecamd
    from ecamd import problems
    sys.path.insert(0, str(ROOT / "oracle"))
    import ecor

    def case(m):
        params = dict(vortex_center=(0.5, 0.5), density_outer=1.0, mach=mach,
                      pressure=1.0 / (1.4 * mach * mach))
        pd = problems.problem_desc("gresho_vortex", m, m, **params)
        U = problems.initial_state("gresho_vortex", m, m, **params)
        s = ecamd.ExplicitSolver(ecamd.Grid(m, m, 1.0 / m, 1.0 / m, 1.4, "hllc", "minmod", 2), dev)
        s.direct_setup("gresho_exact_dirichlet_2d", pd, 1.0)
        return s, U, params

    def timed(s, U0, dt, prec):
        opts = s.imex_opts(precision=prec, tol=1e-8, max_iter=1000, restart=30, picard_iterations=4)
        U = torch.from_numpy(U0).to(dev)
        s.direct_step(U.clone(), 0.0, dt, opts)     # warm-up
        times = []
        for _ in range(reps):
            V = U.clone()
            torch.cuda.synchronize()
            a, b = _events()
            a.record()
            d = s.direct_step(V, 0.0, dt, opts)
            b.record()
            torch.cuda.synchronize()
            times.append(a.elapsed_time(b))
        return sorted(times)[len(times) // 2], d

    s, U0, params = case(n)
    dt = s.direct_dt(torch.from_numpy(U0).to(dev), 0.8, -1.0, 0.0, 1.0)
    out = {"workload": f"gresho_{n}x{n}_exact_dirichlet_M{mach:g}_direct_bdltv20_t1s2_eps1_picard4_tol1e-8",
           "dt": dt}
    for prec in ("fp64", "mixed"):
        ms, d = timed(s, U0, dt, prec)
        out[prec] = {"ms_per_step": ms, "gmres_iterations": d.iterations,
                     "solver_failures": d.solver_failures}
    s.close()
    # like for like with the oracle (one host thread) at a small size
    s, U0, params = case(cpu_n)
    dt_c = s.direct_dt(torch.from_numpy(U0).to(dev), 0.8, -1.0, 0.0, 1.0)
    gpu_ms, _ = timed(s, U0, dt_c, "fp64")
    s.close()
    e = ecor.make_exact(ecor.EXACT_GRESHO, (0, 0), vortex_center=(0.5, 0.5),
                        density_outer=1.0, pressure=params["pressure"], mach=mach)
    g = ecor.make_grid(cpu_n, cpu_n, 1.0 / cpu_n, 1.0 / cpu_n)
    t0 = _time.perf_counter()
    ecor.direct_step(g, ecor.make_direct(e, 1.0), U0, 0.0, dt_c, picard=4, tol=1e-8, max_iter=1000)
    cpu_ms = (_time.perf_counter() - t0) * 1e3
    out["cpu_baseline"] = {"kind": "port", "n": cpu_n, "device_fp64_ms": gpu_ms, "oracle_ms": cpu_ms,
                           "cores": 1}
    return out


def bench_adjoint(dev, nx: int = 640, ny: int = 160, nsteps: int = 768, K: int = 32,
                  ndirs: int = 16, fd_eps: float = 1e-7) -> dict:
    """Checkpointed adjoint of the explicit O2 HLLC run: single-gamma shock
    (Mach 1.22) into a light bubble on [0,4]x[0,1]; gradient of the bubble
    kinetic energy after nsteps w.r.t. the initial density; cost ratio of the
    adjoint sweep (with its recompute) to the forward run, and central
    differences on ndirs seeded smooth directions."""
    import ctypes as C
    import numpy as np
    import torch
    import ecamd
    from ecamd import problems
    from ecamd._lib import lib
    from ecamd.adjoint import ExplicitRun, conserved_from_density, kinetic_energy_in_region, run_with_dts
    lx, ly = 4.0, 1.0
    U0 = problems.initial_state("shock_bubble_mach", nx, ny, prob_hi=(lx, ly), gamma=1.4,
                                shock_mach=1.22, shock_x=0.5, density_outer=1.0, pressure=1.0,
                                density_inner=0.138, blob_center=(1.0, 0.5), blob_radius=0.25)
    post = problems.post_shock_state(1.4, 1.22)
    inflow = (post[0], post[0] * post[1], 0.0, post[3] / 0.4 + 0.5 * post[0] * post[1] ** 2)
    grid = ecamd.Grid(nx, ny, lx / nx, ly / ny, 1.4, "hllc", "minmod", 2,
                      ("fixed", "outflow", "reflect", "outflow"), (inflow, None, None, None))
    s = ecamd.Solver(grid, dev)
    rho = torch.from_numpy(U0[0]).to(dev)
    u = torch.from_numpy(U0[1] / U0[0]).to(dev)
    v = torch.from_numpy(U0[2] / U0[0]).to(dev)
    p = torch.from_numpy(0.4 * (U0[3] - 0.5 * (U0[1] ** 2 + U0[2] ** 2) / U0[0])).to(dev)
    x = (torch.arange(nx, device=dev, dtype=torch.float64) + 0.5) * lx / nx
    y = (torch.arange(ny, device=dev, dtype=torch.float64) + 0.5) * ly / ny
    X, Y = torch.meshgrid(x, y, indexing="xy")
    bubble = (((X - 1.0) ** 2 + (Y - 0.5) ** 2) < 0.25 ** 2).to(torch.float64)
    region = (((X - 1.0) ** 2 + (Y - 0.5) ** 2) < 0.4 ** 2).to(torch.float64)
    rho = rho * (1.0 + 0.05 * torch.sin(3 * X) * torch.cos(2 * Y) * bubble)
    obj = lambda U: kinetic_energy_in_region(U, region, lx / nx * ly / ny)
    cfl = 0.4
    Ub = conserved_from_density(rho, u, v, p, 1.4)
    # plain forward: the device driver loop, no checkpoints
    A, B = Ub.clone(), torch.empty_like(Ub)
    run = s.new_run(cfl, 1.0e30, nsteps)
    s.advance(A, B, nsteps, run)
    torch.cuda.synchronize()
    A = Ub.clone()
    run = s.new_run(cfl, 1.0e30, nsteps)
    a, b = _events()
    a.record()
    s.advance(A, B, nsteps, run)
    b.record()
    torch.cuda.synchronize()
    fwd_ms = a.elapsed_time(b)
    # checkpointed forward and adjoint sweep
    n = nx * ny
    chk = torch.empty(((nsteps + K - 1) // K) * 4 * n, dtype=torch.float64, device=dev)
    UN = torch.empty_like(Ub)
    dts = np.zeros(nsteps)
    done = C.c_int()
    def fwd_chk():
        lib().ec_forward_checkpointed(s.ctx, C.c_void_p(Ub.data_ptr()), C.c_void_p(UN.data_ptr()),
                                      C.byref(s.layout), nsteps, K, cfl, 1.0e30,
                                      C.c_void_p(chk.data_ptr()),
                                      dts.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done))
    fwd_chk()
    torch.cuda.synchronize()
    a, b = _events()
    a.record()
    fwd_chk()
    b.record()
    torch.cuda.synchronize()
    fwd_chk_ms = a.elapsed_time(b)
    UNg = UN.detach().clone().requires_grad_(True)
    (lamN,) = torch.autograd.grad(obj(UNg), UNg)
    lamN = lamN.contiguous()
    lam0 = torch.empty_like(lamN)
    def adj():
        lib().ec_adjoint_checkpointed(s.ctx, C.c_void_p(chk.data_ptr()), C.byref(s.layout), nsteps, K,
                                      dts.ctypes.data_as(C.POINTER(C.c_double)),
                                      C.c_void_p(lamN.data_ptr()), C.c_void_p(lam0.data_ptr()))
    adj()
    torch.cuda.synchronize()
    a, b = _events()
    a.record()
    adj()
    b.record()
    torch.cuda.synchronize()
    adj_ms = a.elapsed_time(b)
    # gradient through the autograd wrapper, checked by central differences
    rho_req = rho.clone().requires_grad_(True)
    J = ExplicitRun.apply(rho_req, u, v, p, s, nsteps, K, cfl, obj)
    (grad,) = torch.autograd.grad(J, rho_req)
    rng = np.random.default_rng(20261203)
    errs = []
    eps = fd_eps
    D = []
    for _ in range(ndirs):
        ph = rng.uniform(0, 2 * np.pi, 2)
        kx, ky = rng.integers(2, 7, 2)
        d = torch.sin(float(kx) * X + ph[0]) * torch.cos(float(ky) * Y + ph[1]) * bubble
        D.append(d)
        Jp = float(obj(run_with_dts(s, conserved_from_density(rho + eps * d, u, v, p, 1.4), dts)))
        Jm = float(obj(run_with_dts(s, conserved_from_density(rho - eps * d, u, v, p, 1.4), dts)))
        fd = (Jp - Jm) / (2 * eps)
        ad = float((grad * d).sum())
        errs.append(abs(ad - fd) / max(abs(fd), 1e-300))
    frozen = frozen_branch_fd(rho, u, v, p, Ub, dts, inflow, torch.stack(D), grad, region,
                              lx / nx, ly / ny)
    s.close()
    return {"workload": f"shock_bubble_{nx}x{ny}_M1.22_hllc_o2_cfl{cfl:g}_{nsteps}steps_K{K}",
            "forward_ms": fwd_ms, "forward_checkpointed_ms": fwd_chk_ms, "adjoint_ms": adj_ms,
            "adjoint_over_forward": adj_ms / fwd_ms, "steps": nsteps, "checkpoint_every": K,
            "fd_directions": ndirs, "fd_eps": eps, "fd_rel_err_max": max(errs),
            "fd_rel_err_median": sorted(errs)[len(errs) // 2],
            "frozen_branch_fd": frozen}


def frozen_branch_fd(rho, u, v, p, Ub, dts, inflow, D, grad, region, dx, dy, eps: float = 1e-5):
    """Central differences of the frozen-branch forward map on the same
    directions: tests/torch_step.py (an independent FP64 torch restatement of
    the step) records every branch decision of the forward trajectory, then
    replays them for all 2 x ndirs perturbed runs as one batched run.  This is
    the map the adjoint differentiates (dt and every limiter and Riemann
    branch frozen), so its differences carry no kink-crossing error."""
    import numpy as np
    import torch
    sys.path.insert(0, str(ROOT / "tests"))
    import torch_step  # noqa: E402  (the test suite's differentiable restatement of the step)
    from ecamd.adjoint import conserved_from_density
    bcs = ("fixed", "outflow", "reflect", "outflow")
    B = torch_step.Branches("record")
    with torch.no_grad():
        V = Ub[:, None].clone()
        for dt in dts:
            V = torch_step.step(V, float(dt), dx, dy, riemann="hllc", bcs=bcs,
                                fixed=(inflow, None, None, None), B=B)
        nd = D.shape[0]
        R = torch.cat([rho[None] + eps * D, rho[None] - eps * D])
        V = conserved_from_density(R, u[None], v[None], p[None], 1.4)
        B.replay()
        for dt in dts:
            V = torch_step.step(V, float(dt), dx, dy, riemann="hllc", bcs=bcs,
                                fixed=(inflow, None, None, None), B=B)
        Jv = (0.5 * (V[1] ** 2 + V[2] ** 2) / V[0] * region).sum((-2, -1)) * (dx * dy)
    fd = (Jv[:nd] - Jv[nd:]) / (2 * eps)
    ad = (grad[None] * D).sum((-2, -1))
    err = ((ad - fd).abs() / ad.abs()).cpu().numpy()
    return {"eps": eps, "rel_err_max": float(err.max()), "rel_err_median": float(np.median(err))}


def bench_ensemble(dev, riemann: str, stop_time: float = 0.1, cfl: float = 0.8,
                   mode: str = "auto", tile: int = 0) -> dict:
    """The 64 seeded Riemann strips (tests/golden/ensemble_64.json), each run
    to stop_time: autotuned (mode, tile) table and the best run's throughput
    (mode "graph" or "persistent": that mode with tile shape `tile`, no autotune)."""
    import numpy as np
    from ecamd.batch import Ensemble, riemann_strip_members, tile_shapes
    table = json.loads((ROOT / "tests" / "golden" / "ensemble_64.json").read_text())
    members = table["members"]
    ens = Ensemble(riemann_strip_members(members, cfl, stop_time, 100000), gamma=1.4,
                   riemann=riemann, device=dev)
    g = 1.4
    for k, m in enumerate(members):
        nx = m["nx"]
        xc = (np.arange(nx) + 0.5) / nx
        U = np.empty((4, 4, nx))
        for side, sel in (("left", xc < 0.5), ("right", xc >= 0.5)):
            rho, uu, vv, pp = m[side]
            cons = (rho, rho * uu, rho * vv, pp / (g - 1.0) + 0.5 * rho * (uu * uu + vv * vv))
            for c in range(4):
                U[c][:, sel] = cons[c]
        ens.set_state(k, U)
    if mode == "auto":
        tuned = ens.autotune()
    else:
        tuned = []
        ens.configure(mode, tile)
    res, used = ens.run()
    cell_updates = sum(r.step * m["nx"] * 4 for r, m in zip(res, members))
    shapes = tile_shapes()
    groups = ens.groups
    out = {"workload": f"riemann_strips_64_nx100-1600x4_o2_minmod_{riemann}_cfl{cfl:g}_t{stop_time:g}",
           "members": len(members), "cell_updates": cell_updates,
           "max_member_steps": max(r.step for r in res), "best": {"mode": used.mode,
           "tile": shapes[used.tile_id], "ms": used.ms, "stream_groups": groups},
           "cell_updates_per_s": cell_updates / (used.ms / 1e3),
           "autotune": [{"mode": t.mode, "tile": shapes[t.tile_id], "ms": t.ms} for t in tuned],
           "statuses_ok": all(r.status == 0 for r in res)}
    ens.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--riemann", default="xie_am_hllc_p")
    ap.add_argument("--order", type=int, default=2)
    ap.add_argument("--mode", choices=("step", "copy"), default="step",
                    help="copy: time a plain device copy of one state (HBM counter calibration)")
    ap.add_argument("--cfl", type=float, default=0.4)
    ap.add_argument("--mach", type=float, default=0.01)
    ap.add_argument("--tile", type=int, default=-1, help="explicit tile shape id (-1: default)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the IMEX, adjoint and ensemble measurements")
    ap.add_argument("--only", choices=("imex", "adjoint", "ensemble", "direct"),
                    help="print only that measurement")
    ap.add_argument("--ensemble-mode", choices=("auto", "graph", "persistent"), default="auto",
                    help="ensemble rounds: autotuned, or one mode with the tile shape of --ensemble-tile")
    ap.add_argument("--ensemble-tile", type=int, default=0,
                    help="ensemble tile shape id for --ensemble-mode graph or persistent")
    ap.add_argument("--summarize-counters", nargs="+", metavar="KIND=DIR",
                    help="write the profile summaries from rocprofv3 --pmc run directories "
                         "(sq=, fetch=, write=, copy_fetch=, copy_write=) and exit")
    ap.add_argument("--commit", default=None, help="commit the summarized counters were measured on")
    args = ap.parse_args()

    if args.summarize_counters:
        summarize_counters(args.summarize_counters, args.commit)
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(json.dumps({"error": "this build runs the euler_compare path on one GPU; "
                                   "multi-GPU runs are not part of it", "n_gpus": args.gpus}))
        return 2

    import torch
    import ecamd
    from ecamd import problems

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    n = args.n
    if args.only:
        fn = {"imex": lambda: bench_imex(dev), "adjoint": lambda: bench_adjoint(dev),
              "direct": lambda: bench_direct(dev),
              "ensemble": lambda: {r: bench_ensemble(dev, r, mode=args.ensemble_mode, tile=args.ensemble_tile)
                                   for r in ("xie_am_hllc_p", "hllc")}}
        print(json.dumps({args.only: fn[args.only]()}))
        return 0
    if args.mode == "copy":
        return copy_calibration(dev, n, args.steps)
    grid = ecamd.Grid(n, n, 1.0 / n, 1.0 / n, 1.4, args.riemann, "minmod", args.order)
    stream = torch.cuda.current_stream(dev)
    solver = ecamd.ExplicitSolver(grid, dev, stream=stream)
    if args.tile >= 0:
        solver.set_tile(args.tile)
    A = torch.from_numpy(problems.gresho(n, args.mach)).to(dev)
    B = torch.empty_like(A)
    run = solver.new_run(args.cfl, 1.0, 2**30)

    out = solver.advance(A, B, args.warmup, run)
    if out is B:
        A, B = B, A
    torch.cuda.synchronize()
    step0 = run.step

    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    out = solver.advance(A, B, args.steps, run)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = start.elapsed_time(end) / 1e3
    if run.step - step0 != args.steps or run.status != 0:
        print(json.dumps({"error": "the run stopped early", "status": solver.status_name(run),
                          "steps_done": run.step - step0}))
        return 1

    per_step = gpu_s / args.steps
    cells = n * n
    value = cells * args.steps / gpu_s
    # The step kernel's own average duration, from HIP events recorded around
    # each launch on the solver's stream (a second pass of the same loop; the
    # per-step finalize kernel is outside the brackets).
    if out is B:
        A, B = B, A
    solver.arm_kernel_timing(args.steps)
    out = solver.advance(A, B, args.steps, run)
    torch.cuda.synchronize()
    k_ms, k_n = solver.kernel_time()
    solver.disarm_kernel_timing()
    kernel_s = k_ms / 1e3 / max(k_n, 1)
    achieved = BYTES_PER_CELL_UPDATE * cells / kernel_s / 1e9
    copy_gbs = stream_copy_gbs(dev)
    digest = kernel_sources_digest()
    traffic, traffic_prov = None, None
    if n == 2048 and args.riemann == "xie_am_hllc_p" and args.order == 2:
        traffic, traffic_prov = profile_figure(TRAFFIC_FILE, "hbm_bytes_per_launch", digest)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Gresho vortex initial data generated on the host)",
        "config": {"workload": f"gresho_vortex_{n}x{n}_periodic_M{args.mach:g}_explicit_"
                               f"o{args.order}_minmod_{args.riemann}_cfl{args.cfl:g}",
                   "nx": n, "ny": n, "parallelism": "single_gpu"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_profile": traffic_prov, "kernel_sources": digest,
                     "kernel": f"explicit_step_kernel ({args.riemann}, O{args.order})",
                     "kernel_us_avg": kernel_s * 1e6, "kernel_launches_timed": k_n,
                     "bytes_per_cell_update": BYTES_PER_CELL_UPDATE,
                     "bytes_per_launch": BYTES_PER_CELL_UPDATE * cells,
                     "fp64_valu": valu_bound(args.riemann, args.order, n, kernel_s, digest),
                     "stream_copy": {"achieved": copy_gbs, "unit": "GB/s",
                                     "frac_of_peak": copy_gbs / HBM_PEAK_GBS,
                                     "kernel_frac_of_copy": achieved / copy_gbs}},
        "wall_s": wall,
        "sim_time_advanced": run.time,
    }
    imex_cpu = None
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(n, args.mach, args.cfl, args.cpu_steps)
        if not args.no_extras:
            imex_cpu = cpu_imex()
    if not args.no_extras:
        line["roofline_sweep"] = roofline_sweep(dev, args.riemann, args.order, args.mach, args.cfl)
        line["imex"] = bench_imex(dev, oracle_state=imex_cpu)
        if imex_cpu is not None:
            line["imex"]["cpu_baseline"] = {
                "seconds_per_step": imex_cpu["seconds_per_step"], "cores": 1, "kind": "port",
                "sample": "one IMEX step of the same 512x512 state, serial oracle, tol 1e-8 "
                          "(the reference's unrestarted MGS GMRES does not reach 1e-10 here)"}
        line["direct_bdltv20"] = bench_direct(dev)
        line["adjoint"] = bench_adjoint(dev)
        line["ensemble"] = {r: bench_ensemble(dev, r) for r in ("xie_am_hllc_p", "hllc")}
    print(json.dumps(line))
    solver.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
