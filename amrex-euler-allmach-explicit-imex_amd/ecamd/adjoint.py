"""Differentiable explicit O2 run: checkpointed forward + hand-written adjoint.

``ExplicitRun.apply(rho0, prim_rest, solver, nsteps, K, cfl, objective)``
is a ``torch.autograd.Function``: the forward pass builds the initial
conserved state from a density field at fixed velocity and pressure, runs
``nsteps`` steps of the device driver loop (dt from the CFL reduction,
recorded per step) storing a checkpoint every ``K`` steps, and returns the
objective of the final state; the backward pass runs the adjoint sweep on the
device (``ec_adjoint_checkpointed``: each K-step segment is recomputed once
from its checkpoint, then reversed with dt and every branch frozen at the
forward trajectory) and returns d objective / d rho0.

The adjoint itself is hand-written HIP (libec_amd.so); torch only holds the
device buffers and chains the objective's gradient.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import check, lib


def _ptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def conserved_from_density(rho: torch.Tensor, u: torch.Tensor, v: torch.Tensor, p: torch.Tensor,
                           gamma: float) -> torch.Tensor:
    """to_conserved (core/state_primitives.hpp:113-117) with fixed u, v, p."""
    kin = 0.5 * rho * (u * u + v * v)
    return torch.stack([rho, rho * u, rho * v, p / (gamma - 1.0) + kin]).contiguous()


def kinetic_energy_in_region(U: torch.Tensor, mask: torch.Tensor, cell_area: float) -> torch.Tensor:
    """Integrated kinetic energy |m|^2/(2 rho) dA over the masked cells."""
    return (0.5 * (U[1] ** 2 + U[2] ** 2) / U[0] * mask).sum() * cell_area


class ExplicitRun(torch.autograd.Function):
    """J(rho0) for a device explicit run; backward = checkpointed adjoint."""

    @staticmethod
    def forward(ctx, rho0, u0, v0, p0, solver, nsteps, K, cfl, objective):
        g = solver.grid
        U0 = conserved_from_density(rho0.detach(), u0, v0, p0, g.gamma)
        n = g.nx * g.ny
        nseg = (nsteps + K - 1) // K
        chk = torch.empty(nseg * 4 * n, dtype=torch.float64, device=U0.device)
        UN = torch.empty_like(U0)
        dts = np.zeros(nsteps)
        done = C.c_int()
        check(solver.ctx, lib().ec_forward_checkpointed(
            solver.ctx, _ptr(U0), _ptr(UN), C.byref(solver.layout), nsteps, K, cfl, 1.0e30,
            _ptr(chk), dts.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done)),
            "ec_forward_checkpointed")
        if done.value != nsteps:
            raise RuntimeError(f"forward run stopped after {done.value} of {nsteps} steps")
        with torch.enable_grad():
            UNg = UN.detach().requires_grad_(True)
            J = objective(UNg)
            (lamN,) = torch.autograd.grad(J, UNg)
        ctx.save_for_backward(u0, v0, p0, chk, lamN.contiguous())
        ctx.solver, ctx.nsteps, ctx.K, ctx.dts = solver, nsteps, K, dts
        ctx.UN = UN
        return J.detach()

    @staticmethod
    def backward(ctx, gJ):
        u0, v0, p0, chk, lamN = ctx.saved_tensors
        s = ctx.solver
        lam0 = torch.empty_like(lamN)
        check(s.ctx, lib().ec_adjoint_checkpointed(
            s.ctx, _ptr(chk), C.byref(s.layout), ctx.nsteps, ctx.K,
            ctx.dts.ctypes.data_as(C.POINTER(C.c_double)), _ptr(lamN), _ptr(lam0)),
            "ec_adjoint_checkpointed")
        # dU0/drho0 at fixed (u, v, p) = (1, u, v, (u^2+v^2)/2)
        grad = lam0[0] + lam0[1] * u0 + lam0[2] * v0 + lam0[3] * 0.5 * (u0 * u0 + v0 * v0)
        return gJ * grad, None, None, None, None, None, None, None, None


def run_with_dts(solver, U0: torch.Tensor, dts) -> torch.Tensor:
    """Forward run with prescribed dts (host dt per step): the frozen-dt map
    the adjoint differentiates, used for finite-difference checks."""
    A = U0.clone()
    B = torch.empty_like(A)
    for dt in dts:
        solver.step(A, B, float(dt), diag=False)
        A, B = B, A
    return A


def adjoint_step(solver, U: torch.Tensor, lam_next: torch.Tensor, dt: float) -> torch.Tensor:
    lam = torch.empty_like(lam_next)
    check(solver.ctx, lib().ec_adjoint_step(solver.ctx, _ptr(U), _ptr(lam_next), _ptr(lam),
                                             C.byref(solver.layout), float(dt)), "ec_adjoint_step")
    return lam
