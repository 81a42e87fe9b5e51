"""The batched ensemble (ec_batch_*) on the GPU: the 64 seeded Riemann
strips of tests/golden/ensemble_64.json, O2 HLLC and HLLC-P.

* capped runs (10 steps): every member bit-identical to the oracle's
  driver loop, for both round modes and every compiled tile shape;
* full runs to t = 0.1: every member bit-identical to the single-grid
  device loop (ec_explicit_advance) on that member alone, which the
  explicit tests pin to the oracle;
* per-member statuses: a non-physical member stops alone;
* the autotuner times every graph-mode tile shape and leaves the states as
  they were.
"""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).parent / "golden"
GAMMA = 1.4
CFL = 0.8


def table():
    return json.loads((GOLDEN / "ensemble_64.json").read_text())["members"]


def conserved(q):
    rho, u, v, p = q
    return np.array([rho, rho * u, rho * v, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)])


def initial_state(m):
    nx = m["nx"]
    x = (np.arange(nx) + 0.5) / nx
    U = np.empty((4, 4, nx))
    uL, uR = conserved(m["left"]), conserved(m["right"])
    for c in range(4):
        U[c] = np.where(x < 0.5, uL[c], uR[c])[None, :]
    return U


def make_members(stop_time, max_step):
    from ecamd.batch import riemann_strip_members
    return riemann_strip_members(table(), CFL, stop_time, max_step)


def build(gpu, riemann, stop_time, max_step, states=None):
    from ecamd.batch import Ensemble
    ens = Ensemble(make_members(stop_time, max_step), gamma=GAMMA, riemann=riemann, device=gpu)
    for k, m in enumerate(table()):
        ens.set_state(k, states[k] if states is not None else initial_state(m))
    return ens


RIEMANN_ID = {"hllc": 1, "xie_am_hllc_p": 3}


@pytest.fixture(scope="module")
def oracle_capped(ecor):
    out = {}
    for name, rid in RIEMANN_ID.items():
        runs = []
        for m in table():
            nx = m["nx"]
            g = ecor.make_grid(nx, 4, 1.0 / nx, 1.0 / nx, riemann=rid, bc=(1, 1, 0, 0))
            U, res, _ = ecor.run_explicit(g, initial_state(m), CFL, 0.1, 10)
            runs.append((U, res))
        out[name] = runs
    return out


@pytest.mark.parametrize("riemann", list(RIEMANN_ID))
@pytest.mark.parametrize("mode", ["graph", "persistent"])
@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4])
def test_capped_members_match_oracle(gpu, oracle_capped, riemann, mode, tile):
    print(f"ensemble capped {riemann} {mode} tile {tile}", flush=True)
    ens = build(gpu, riemann, 0.1, 10)
    ens.configure(mode, tile, 4)
    res, used = ens.run()
    assert used.mode == mode and used.tile_id == tile
    for k, (U, r) in enumerate(oracle_capped[riemann]):
        got = ens.get_state(k)
        assert np.array_equal(got, U), f"member {k}"
        assert res[k].step == r.steps == 10
        assert res[k].time == r.time
        assert res[k].degenerate_total == r.degenerate_total
        assert res[k].status == 0
    ens.close()


@pytest.mark.parametrize("riemann", list(RIEMANN_ID))
@pytest.mark.parametrize("mode", ["graph", "persistent"])
def test_full_runs_match_single_grid_loop(gpu, riemann, mode):
    from ecamd import ExplicitSolver, Grid
    from ecamd.explicit import run_explicit
    print(f"ensemble full {riemann} {mode}", flush=True)
    ens = build(gpu, riemann, 0.1, 100000)
    ens.configure(mode, 0, 64)
    res, used = ens.run()
    for k, m in enumerate(table()):
        nx = m["nx"]
        solver = ExplicitSolver(Grid(nx, 4, 1.0 / nx, 1.0 / nx, GAMMA, riemann, "minmod", 2,
                                     ("outflow", "outflow", "periodic", "periodic")), gpu)
        U, run = run_explicit(solver, torch.from_numpy(initial_state(m)).to(gpu), CFL, 0.1, 100000)
        assert np.array_equal(ens.get_state(k), U.cpu().numpy()), f"member {k}"
        assert (res[k].step, res[k].time, res[k].degenerate_total, res[k].status) == \
            (run.step, run.time, run.degenerate_total, run.status)
        assert res[k].dt_min == run.dt_min and res[k].dt_max == run.dt_max
        assert abs(res[k].time - 0.1) < 1e-14
        solver.close()
    ens.close()


def test_failing_member_stops_alone(gpu):
    states = [initial_state(m) for m in table()]
    states[3] = states[3].copy()
    states[3][3, :, 5] = -1.0          # negative energy: non-physical initial state
    ens = build(gpu, "hllc", 0.01, 100000, states)
    ens.configure("persistent", 0, 0)
    res, _ = ens.run()
    assert ens.status_name(res[3]) == "initial_nonphysical_state" and res[3].step == 0
    assert np.array_equal(ens.get_state(3), states[3])
    for k in range(len(states)):
        if k != 3:
            assert res[k].status == 0 and abs(res[k].time - 0.01) < 1e-14
    ens.close()


def test_autotune_restores_states(gpu):
    from ecamd.batch import tile_shapes
    ens = build(gpu, "xie_am_hllc_p", 0.01, 100000)
    before = [ens.get_state(k) for k in range(len(table()))]
    timings = ens.autotune()
    assert len(timings) == len(tile_shapes()) and all(t.ms > 0 for t in timings)
    assert all(t.mode == "graph" for t in timings)    # graph mode only
    for k in range(len(table())):
        assert np.array_equal(ens.get_state(k), before[k])
    res, used = ens.run()
    best = min(timings, key=lambda t: t.ms)
    assert (used.mode, used.tile_id) == (best.mode, best.tile_id)
    ens.configure(best.mode, best.tile_id, 0)
    ens2 = build(gpu, "xie_am_hllc_p", 0.01, 100000)
    ens2.configure("graph", 0, 0)
    ens2.run()
    for k in range(len(table())):
        assert np.array_equal(ens.get_state(k), ens2.get_state(k))
    ens.close()
    ens2.close()


@pytest.mark.parametrize("mode", ["graph", "persistent"])
@pytest.mark.parametrize("tile", [0, 2, 4])
def test_mixed_2d_and_strip_members(ecor, gpu, mode, tile):
    """Members of different kinds in one batch: Gresho boxes of several sizes
    (periodic), a strip and a shock-bubble-rule box (fixed / outflow /
    reflect / outflow), HLLC-P O2, 15 capped steps; each member bit-identical
    to the oracle's driver loop."""
    from ecamd.batch import Ensemble, Member
    specs = [(64, 64, (0, 0, 0, 0)), (48, 40, (0, 0, 0, 0)), (100, 30, (0, 0, 0, 0)),
             (200, 4, (1, 1, 0, 0)), (36, 28, (3, 1, 2, 1))]
    members, states, grids = [], [], []
    for k, (nx, ny, bc) in enumerate(specs):
        if bc[0] == 3:
            U = ecor.init_gresho(nx, ny, mach=0.3)
            fixed = (tuple(U[:, ny // 2, 0] * np.array([1.05, 1.2, 1.0, 1.1])), None, None, None)
        elif ny == 4:
            U = ecor.init_toro(nx, ny, [1, 0.2, 0, 1], [0.3, -0.1, 0, 0.4])
            fixed = None
        else:
            U = ecor.init_gresho(nx, ny, mach=0.05)
            fixed = None
        g = ecor.make_grid(nx, ny, 1.0 / nx, 1.0 / ny, riemann=3, bc=bc, fixed=fixed)
        members.append(Member(nx=nx, ny=ny, dx=1.0 / nx, dy=1.0 / ny, bc=bc,
                              fixed=fixed if fixed else (None,) * 4, cfl=0.4, stop_time=1.0,
                              max_step=15))
        states.append(np.ascontiguousarray(U))
        grids.append(g)
    ens = Ensemble(members, gamma=1.4, riemann="xie_am_hllc_p", device=gpu)
    for k, U in enumerate(states):
        ens.set_state(k, U)
    ens.configure(mode, tile, 4)
    res, _ = ens.run()
    for k, (U, g) in enumerate(zip(states, grids)):
        ref, r, _ = ecor.run_explicit(g, U, 0.4, 1.0, 15)
        assert res[k].step == r.steps == 15, k
        assert res[k].time == r.time and res[k].degenerate_total == r.degenerate_total
        assert np.array_equal(ens.get_state(k), ref), f"member {k}"
    ens.close()
