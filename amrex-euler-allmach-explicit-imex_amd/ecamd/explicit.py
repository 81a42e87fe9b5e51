"""Explicit O2 update on the device, through the C-ABI.

Mirrors the reference's explicit calls: ``euler_fv_one_step``
(numerics/explicit_update.hpp:58-102), ``compute_euler_dt`` and
``compute_diagnostics`` (numerics/diagnostic_metrics_timestep.hpp:18-47,
:227-260), and the explicit branch of the driver loop
(app/driver_main.hpp:224-333).  Citations are relative to
amrex/apps/euler_compare/src/euler_compare/.

States are torch float64 CUDA tensors of shape (4, ny, nx) (comp-major,
components rho, mx, my, E); torch is only the device allocator and stream
provider here, all arithmetic runs in the HIP kernels of libec_amd.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import torch

from ._lib import (BC, LIMITER, RIEMANN, RUN_STATUS, GridDesc, Layout, RunState, StepDiag,
                   check, lib)


@dataclass
class Grid:
    nx: int
    ny: int
    dx: float
    dy: float
    gamma: float = 1.4
    riemann: str = "hllc"
    limiter: str = "minmod"
    spatial_order: int = 2
    bc: tuple = ("periodic", "periodic", "periodic", "periodic")  # xlo, xhi, ylo, yhi
    fixed_state: tuple = field(default_factory=lambda: (None, None, None, None))

    def desc(self) -> GridDesc:
        d = GridDesc()
        d.nx, d.ny, d.dx, d.dy, d.gamma = self.nx, self.ny, self.dx, self.dy, self.gamma
        d.riemann = RIEMANN[self.riemann]
        d.limiter = LIMITER[self.limiter]
        d.spatial_order = self.spatial_order
        for s in range(4):
            d.bc[s] = BC[self.bc[s]] if isinstance(self.bc[s], str) else int(self.bc[s])
            fs = self.fixed_state[s]
            if fs is not None:
                for c in range(4):
                    d.fixed_state[s][c] = float(fs[c])
        return d


def _ptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


@dataclass
class StepResult:
    rho_min: float
    pressure_min: float
    nonfinite_count: int
    degenerate_count: int
    max_rate: float


class ExplicitSolver:
    """Device explicit O2 solver bound to one grid and one HIP stream."""

    def __init__(self, grid: Grid, device: torch.device | str = "cuda", stream=None):
        self.grid = grid
        dev = torch.device(device)
        if dev.type != "cuda":
            raise ValueError("ExplicitSolver runs on the GPU only")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self._ctx = C.c_void_p()
        with torch.cuda.device(self.device):
            d = grid.desc()
            check(None, lib().ec_create(C.byref(d), C.c_void_p(self.stream.cuda_stream),
                                        C.byref(self._ctx)), "ec_create")
        self.layout = Layout(grid.nx, grid.nx * grid.ny)

    def close(self) -> None:
        if self._ctx:
            lib().ec_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ctx(self):
        return self._ctx

    def empty_state(self) -> torch.Tensor:
        return torch.empty((4, self.grid.ny, self.grid.nx), dtype=torch.float64, device=self.device)

    def _check_state(self, U: torch.Tensor) -> None:
        if (U.dtype != torch.float64 or U.device != self.device or not U.is_contiguous()
                or tuple(U.shape) != (4, self.grid.ny, self.grid.nx)):
            raise ValueError(f"state must be contiguous float64 (4,{self.grid.ny},{self.grid.nx}) "
                             f"on {self.device}")

    def set_tile(self, tile_id: int) -> None:
        check(self._ctx, lib().ec_set_explicit_tile(self._ctx, tile_id), "ec_set_explicit_tile")

    @staticmethod
    def tile_shapes() -> list[tuple[int, int, int]]:
        out = []
        for t in range(lib().ec_explicit_tile_count()):
            a, b, c = C.c_int(), C.c_int(), C.c_int()
            lib().ec_explicit_tile_shape(t, C.byref(a), C.byref(b), C.byref(c))
            out.append((a.value, b.value, c.value))
        return out

    def set_exact_boundary(self, problem, periodic) -> None:
        """euler.field_boundary=exact_dirichlet: the exact branches of
        fill_problem_ghosts (problems/boundary_conditions.hpp:133-219).
        problem: a ProblemDesc (see problems.problem_desc) of this grid;
        periodic: geometry.is_periodic as two ints."""
        per = (C.c_int * 2)(int(periodic[0]), int(periodic[1]))
        check(self._ctx, lib().ec_set_exact_boundary(self._ctx, C.byref(problem), per),
              "ec_set_exact_boundary")

    def set_boundary_time(self, time: float) -> None:
        """Time of the exact ghost states for the next host-dt steps (the time
        argument of euler_fv_one_step, numerics/explicit_update.hpp:58-65)."""
        check(self._ctx, lib().ec_set_boundary_time(self._ctx, float(time)),
              "ec_set_boundary_time")

    def step(self, U_in: torch.Tensor, U_out: torch.Tensor, dt: float,
             diag: bool = True) -> StepResult | None:
        """One explicit step U_out = U_in - dt div F(U_in)."""
        self._check_state(U_in)
        self._check_state(U_out)
        d = StepDiag()
        check(self._ctx, lib().ec_explicit_step(self._ctx, _ptr(U_in), _ptr(U_out),
                                                 C.byref(self.layout), float(dt),
                                                 C.byref(d) if diag else None),
              "ec_explicit_step")
        if not diag:
            return None
        return StepResult(d.rho_min, d.pressure_min, d.nonfinite_count, d.degenerate_count,
                          d.max_rate)

    def step_inplace(self, U: torch.Tensor, dt: float) -> StepResult:
        self._check_state(U)
        d = StepDiag()
        check(self._ctx, lib().ec_explicit_step_inplace(self._ctx, _ptr(U), C.byref(self.layout),
                                                         float(dt), C.byref(d)),
              "ec_explicit_step_inplace")
        return StepResult(d.rho_min, d.pressure_min, d.nonfinite_count, d.degenerate_count,
                          d.max_rate)

    def diagnostics(self, U: torch.Tensor) -> StepResult:
        self._check_state(U)
        d = StepDiag()
        check(self._ctx, lib().ec_state_diagnostics(self._ctx, _ptr(U), C.byref(self.layout),
                                                     C.byref(d)), "ec_state_diagnostics")
        return StepResult(d.rho_min, d.pressure_min, d.nonfinite_count, d.degenerate_count,
                          d.max_rate)

    @staticmethod
    def euler_dt(max_rate: float, cfl: float, time: float, stop_time: float) -> float:
        return lib().ec_euler_dt(max_rate, cfl, time, stop_time)

    def new_run(self, cfl: float, stop_time: float, max_step: int) -> RunState:
        rs = RunState()
        rs.cfl, rs.stop_time, rs.max_step = cfl, stop_time, max_step
        return rs

    def advance(self, U_a: torch.Tensor, U_b: torch.Tensor, nsteps: int,
                run: RunState) -> torch.Tensor:
        """Device-resident driver loop for up to nsteps; returns the buffer holding the state."""
        self._check_state(U_a)
        self._check_state(U_b)
        in_b = C.c_int()
        check(self._ctx, lib().ec_explicit_advance(self._ctx, _ptr(U_a), _ptr(U_b),
                                                    C.byref(self.layout), int(nsteps),
                                                    C.byref(run), C.byref(in_b)),
              "ec_explicit_advance")
        return U_b if in_b.value else U_a

    def arm_kernel_timing(self, launches: int) -> None:
        """Bracket each of the next `launches` step kernels with HIP events."""
        check(self._ctx, lib().ec_step_kernel_timing(self._ctx, 1, int(launches)),
              "ec_step_kernel_timing")

    def kernel_time(self) -> tuple[float, int]:
        """(summed step-kernel milliseconds, launches) since arming."""
        ms, n = C.c_double(), C.c_int()
        check(self._ctx, lib().ec_step_kernel_time(self._ctx, C.byref(ms), C.byref(n)),
              "ec_step_kernel_time")
        return ms.value, n.value

    def disarm_kernel_timing(self) -> None:
        check(self._ctx, lib().ec_step_kernel_timing(self._ctx, 0, 0), "ec_step_kernel_timing")

    @staticmethod
    def status_name(run: RunState) -> str:
        return RUN_STATUS.get(run.status, str(run.status))

    # ---- IMEX T1/S2 BDLTV20 (imex/imex_source_map_step.hpp:300-616) --------
    def imex_opts(self, **kw):
        from ._abi_ext import ImexOpts
        o = ImexOpts()
        check(None, lib().ec_imex_default_opts(C.byref(o)), "ec_imex_default_opts")
        prec = kw.pop("precision", None)
        if prec is not None:
            o.precision = ({"fp64": 0, "mixed": 1, "fp64_flexible": 2}[prec] if isinstance(prec, str)
                           else int(prec))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def imex_step(self, U: torch.Tensor, dt: float, opts=None, p_out: torch.Tensor | None = None):
        """One IMEX step of U in place (only on success); returns the diagnostics."""
        from ._abi_ext import ImexDiag
        self._check_state(U)
        o = opts if opts is not None else self.imex_opts()
        d = ImexDiag()
        pp = None
        if p_out is not None:
            if (p_out.dtype != torch.float64 or p_out.numel() != self.grid.nx * self.grid.ny
                    or p_out.device != self.device or not p_out.is_contiguous()):
                raise ValueError("p_out must be a contiguous float64 nx*ny device tensor")
            pp = _ptr(p_out)
        check(self._ctx, lib().ec_imex_step(self._ctx, _ptr(U), C.byref(self.layout), float(dt),
                                             C.byref(o), C.byref(d), pp), "ec_imex_step")
        return d

    # ---- direct BDLTV20 T1/S2 driver (imex/imex_bdltv20_direct_driver.hpp) --
    def direct_setup(self, case, problem, epsilon: float = 1.0) -> None:
        """case: a euler.bdltv20_paper_t1_s2 name or ec_direct_case value;
        problem: a ProblemDesc of this grid (see problems.problem_desc)."""
        from ._abi_ext import DIRECT_CASES
        k = DIRECT_CASES[case] if isinstance(case, str) else int(case)
        check(self._ctx, lib().ec_direct_setup(self._ctx, k, C.byref(problem), float(epsilon)),
              "ec_direct_setup")

    def direct_step(self, U: torch.Tensor, time: float, dt: float, opts=None,
                    p_out: torch.Tensor | None = None):
        """One direct step of U (epsilon-scaled conserved) in place; returns the counters."""
        from ._abi_ext import DirectDiag
        self._check_state(U)
        o = opts if opts is not None else self.imex_opts()
        d = DirectDiag()
        pp = None
        if p_out is not None:
            if (p_out.dtype != torch.float64 or p_out.numel() != self.grid.nx * self.grid.ny
                    or p_out.device != self.device or not p_out.is_contiguous()):
                raise ValueError("p_out must be a contiguous float64 nx*ny device tensor")
            pp = _ptr(p_out)
        check(self._ctx, lib().ec_direct_step(self._ctx, _ptr(U), C.byref(self.layout), float(time),
                                               float(dt), C.byref(o), C.byref(d), pp),
              "ec_direct_step")
        return d

    def direct_dt(self, U: torch.Tensor, imex_cfl: float, fixed_dt: float, time: float,
                  stop_time: float) -> float:
        self._check_state(U)
        dt = C.c_double()
        check(self._ctx, lib().ec_direct_dt(self._ctx, _ptr(U), C.byref(self.layout), imex_cfl,
                                             fixed_dt, time, stop_time, C.byref(dt)),
              "ec_direct_dt")
        return dt.value

    # ---- whole-field metrics ---------------------------------------------
    def admissibility(self, U: torch.Tensor):
        """compute_admissibility_diagnostics (numerics/diagnostic_metrics_timestep.hpp:49-102)."""
        from ._abi_ext import AdmissibilityDiag
        self._check_state(U)
        d = AdmissibilityDiag()
        check(self._ctx, lib().ec_admissibility(self._ctx, _ptr(U), C.byref(self.layout), C.byref(d)),
              "ec_admissibility")
        return d

    def gresho_metrics(self, U: torch.Tensor, problem):
        """compute_gresho_metrics (numerics/diagnostic_metrics_timestep.hpp:147-225)."""
        from ._abi_ext import GreshoMetricsOut
        self._check_state(U)
        m = GreshoMetricsOut()
        check(self._ctx, lib().ec_gresho_metrics(self._ctx, _ptr(U), C.byref(self.layout),
                                                  C.byref(problem), C.byref(m)), "ec_gresho_metrics")
        return m

    def shock_bubble_metrics(self, U: torch.Tensor, prob_lo=(0.0, 0.0), threshold: float = 0.55):
        """compute_shock_density_bubble_metrics (io/output_cases.hpp:329-407)."""
        from ._abi_ext import ShockBubbleMetricsOut
        self._check_state(U)
        m = ShockBubbleMetricsOut()
        lo = (C.c_double * 2)(*prob_lo)
        check(self._ctx, lib().ec_shock_bubble_metrics(self._ctx, _ptr(U), C.byref(self.layout), lo,
                                                        float(threshold), C.byref(m)),
              "ec_shock_bubble_metrics")
        return m

    def cylindrical_source(self, U: torch.Tensor, dt: float, y_lo: float = 0.0) -> None:
        """apply_shock_density_bubble_cylindrical_source
        (problems/boundary_conditions.hpp:75-131), in place on U."""
        self._check_state(U)
        check(self._ctx, lib().ec_cylindrical_source(self._ctx, _ptr(U), C.byref(self.layout), float(dt),
                                                      float(y_lo)),
              "ec_cylindrical_source")

    def imex_dt(self, U: torch.Tensor, cfl: float, time: float, stop_time: float,
                acoustic_startup: int = 0, acoustic_cfl_cap: float = 0.0):
        self._check_state(U)
        dt, m, a = C.c_double(), C.c_double(), C.c_double()
        check(self._ctx, lib().ec_imex_dt(self._ctx, _ptr(U), C.byref(self.layout), cfl, time,
                                           stop_time, acoustic_startup, acoustic_cfl_cap,
                                           C.byref(dt), C.byref(m), C.byref(a)), "ec_imex_dt")
        return dt.value, m.value, a.value


Solver = ExplicitSolver


def run_explicit(solver: ExplicitSolver, U0: torch.Tensor, cfl: float, stop_time: float,
                 max_step: int, chunk: int = 256):
    """Run the explicit driver loop from U0 to stop_time / max_step on the device.

    Returns (U_final, RunState)."""
    A = U0.clone()
    B = torch.empty_like(A)
    run = solver.new_run(cfl, stop_time, max_step)
    while True:
        before = run.step
        res = solver.advance(A, B, chunk, run)
        if res is B:
            A, B = B, A
        done = (run.status != 0 or run.step >= max_step or
                run.time >= stop_time - 1.0e-14 or run.step == before)
        if done:
            return A, run
