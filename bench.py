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
    args = ap.parse_args()

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
    achieved = BYTES_PER_CELL_UPDATE * cells / per_step / 1e9
    traffic = None
    if TRAFFIC_FILE.exists() and n == 2048 and args.riemann == "xie_am_hllc_p" and args.order == 2:
        try:
            traffic = json.loads(TRAFFIC_FILE.read_text()).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
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
                     "kernel": f"explicit_step_kernel ({args.riemann}, O{args.order})",
                     "bytes_per_cell_update": BYTES_PER_CELL_UPDATE},
        "wall_s": wall,
        "sim_time_advanced": run.time,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(n, args.mach, args.cfl, args.cpu_steps)
    print(json.dumps(line))
    solver.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
