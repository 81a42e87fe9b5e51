"""Ensembles of independent explicit runs of uneven resolution (ec_batch_*).

Each member is one euler_compare explicit run (its own grid, boundary rule,
CFL, stop time and max_step); the members share gamma, the Riemann solver, the
limiter and the spatial order.  Every member follows the driver loop of
amrex/apps/euler_compare/src/euler_compare/app/driver_main.hpp:216-333 exactly
as ExplicitSolver.advance does on the member alone, so its final state and run
state are bit-identical to that single-grid run.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from ._abi_ext import BatchTiming, MemberDesc
from ._lib import BC, LIMITER, RIEMANN, RUN_STATUS, EcError, Layout, RunState, lib

MODES = {"auto": 0, "graph": 1, "persistent": 2}
MODE_NAMES = {v: k for k, v in MODES.items()}


@dataclass
class Member:
    nx: int
    ny: int
    dx: float
    dy: float
    bc: tuple = ("outflow", "outflow", "periodic", "periodic")
    fixed: tuple = (None, None, None, None)
    cfl: float = 0.8
    stop_time: float = 0.2
    max_step: int = 100000

    def desc(self) -> MemberDesc:
        d = MemberDesc()
        d.nx, d.ny, d.dx, d.dy = self.nx, self.ny, self.dx, self.dy
        for s in range(4):
            b = self.bc[s]
            d.bc[s] = BC[b] if isinstance(b, str) else int(b)
            if self.fixed[s] is not None:
                for c in range(4):
                    d.fixed_state[s][c] = float(self.fixed[s][c])
        d.cfl, d.stop_time, d.max_step = self.cfl, self.stop_time, self.max_step
        return d


@dataclass
class Timing:
    mode: str
    tile_id: int
    ms: float


def tile_shapes() -> list[tuple[int, int, int]]:
    L = lib()
    out = []
    for t in range(L.ec_batch_tile_count()):
        tx, ty, nt = C.c_int(), C.c_int(), C.c_int()
        L.ec_batch_tile_shape(t, C.byref(tx), C.byref(ty), C.byref(nt))
        out.append((tx.value, ty.value, nt.value))
    return out


class Ensemble:
    """A batch of explicit runs on one GPU (device memory owned by the batch)."""

    def __init__(self, members: list[Member], gamma: float = 1.4, riemann: str = "hllc",
                 limiter: str = "minmod", order: int = 2, device="cuda", stream=None):
        dev = torch.device(device)
        if dev.type != "cuda":
            raise ValueError("Ensemble runs on the GPU only")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.members = list(members)
        self.stream = stream if stream is not None else torch.cuda.current_stream(dev)
        arr = (MemberDesc * len(self.members))(*[m.desc() for m in self.members])
        self._b = C.c_void_p()
        with torch.cuda.device(dev):
            st = lib().ec_batch_create(arr, len(self.members), gamma, RIEMANN[riemann],
                                       LIMITER[limiter], order, C.c_void_p(self.stream.cuda_stream),
                                       C.byref(self._b))
        if st != 0:
            raise EcError(f"ec_batch_create failed with status {st}")

    def _check(self, st: int, what: str) -> None:
        if st != 0:
            msg = (lib().ec_batch_last_error(self._b) or b"").decode()
            raise EcError(f"{what}: status {st} {msg}".strip())

    def close(self) -> None:
        if self._b:
            lib().ec_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_exact_boundary(self, m: int, problem, periodic) -> None:
        """Exact-Dirichlet sides for member m (euler.field_boundary=exact_dirichlet):
        problem is an ecamd.problems.problem_desc of the member's grid,
        periodic the geometry.is_periodic pair."""
        per = (C.c_int * 2)(*[int(x) for x in periodic])
        self._check(lib().ec_batch_set_exact_boundary(self._b, m, C.byref(problem), per),
                    "ec_batch_set_exact_boundary")

    def set_state(self, m: int, U) -> None:
        """U: (4, ny, nx) float64, numpy (host) or torch (host or this device)."""
        mem = self.members[m]
        if isinstance(U, np.ndarray):
            U = np.ascontiguousarray(U, dtype=np.float64)
            ptr = U.ctypes.data
        else:
            if U.dtype != torch.float64 or not U.is_contiguous():
                raise ValueError("state must be a contiguous float64 tensor")
            ptr = U.data_ptr()
        if tuple(U.shape) != (4, mem.ny, mem.nx):
            raise ValueError(f"member {m} state must have shape (4, {mem.ny}, {mem.nx})")
        lay = Layout(mem.nx, mem.nx * mem.ny)
        self._check(lib().ec_batch_set_state(self._b, m, C.c_void_p(ptr), C.byref(lay)),
                    "ec_batch_set_state")

    def get_state(self, m: int) -> np.ndarray:
        mem = self.members[m]
        out = np.empty((4, mem.ny, mem.nx))
        lay = Layout(mem.nx, mem.nx * mem.ny)
        self._check(lib().ec_batch_get_state(self._b, m, C.c_void_p(out.ctypes.data), C.byref(lay)),
                    "ec_batch_get_state")
        return out

    def configure(self, mode: str = "auto", tile_id: int = -1, rounds_per_graph: int = 0) -> None:
        self._check(lib().ec_batch_configure(self._b, MODES[mode], tile_id, rounds_per_graph),
                    "ec_batch_configure")

    def set_groups(self, groups: int) -> None:
        """Graph mode: members split into `groups` streams (1..4)."""
        self._check(lib().ec_batch_set_groups(self._b, groups), "ec_batch_set_groups")

    @property
    def groups(self) -> int:
        return lib().ec_batch_groups(self._b)

    def autotune(self) -> list[Timing]:
        n = 2 * lib().ec_batch_tile_count()
        table = (BatchTiming * n)()
        got = C.c_int()
        self._check(lib().ec_batch_autotune(self._b, table, n, C.byref(got)), "ec_batch_autotune")
        return [Timing(MODE_NAMES[t.mode], t.tile_id, t.ms) for t in table[:got.value]]

    def run(self) -> tuple[list[RunState], Timing]:
        res = (RunState * len(self.members))()
        used = BatchTiming()
        self._check(lib().ec_batch_run(self._b, res, C.byref(used)), "ec_batch_run")
        return list(res), Timing(MODE_NAMES[used.mode], used.tile_id, used.ms)

    @staticmethod
    def status_name(r: RunState) -> str:
        return RUN_STATUS.get(r.status, str(r.status))


def riemann_strip_members(table: list[dict], cfl: float, stop_time: float, max_step: int,
                          ny: int = 4) -> list[Member]:
    """Members of a Riemann-strip ensemble: Nx x ny cells, square cells
    dx = dy = 1/Nx, transmissive x and periodic y (the reference's Toro rows,
    problems/boundary_conditions.hpp:2-25)."""
    return [Member(nx=r["nx"], ny=ny, dx=1.0 / r["nx"], dy=1.0 / r["nx"], cfl=cfl,
                   stop_time=stop_time, max_step=max_step) for r in table]
