"""The hand-written adjoint against a reference gradient: torch.autograd
through an independent differentiable torch version of the step
(tests/torch_step.py).  The torch forward is first checked against the
oracle's step; then the host run of the device adjoint passes
(ec_adjoint_step_host, bit-identical to the device step) must give the
autograd gradient, one step and a chain of steps, for Rusanov, HLLC and
HLLC-P with its 5-face sensor, under the periodic, outflow and shock-bubble
ghost rules.  CPU only."""
import ctypes as C

import numpy as np
import pytest
import torch

import torch_step
from test_gpu_adjoint import smooth_state

dp = C.POINTER(C.c_double)


def host_adjoint(ecor_grid, U, lam, dt):
    from ecamd import _lib
    d = _lib.GridDesc()
    g = ecor_grid
    d.nx, d.ny, d.dx, d.dy, d.gamma = g.nx, g.ny, g.dx, g.dy, g.gamma
    d.riemann, d.limiter, d.spatial_order = g.riemann, g.limiter, g.order
    for s in range(4):
        d.bc[s] = g.bc[s]
    out = np.zeros_like(U)
    U = np.ascontiguousarray(U)
    lam = np.ascontiguousarray(lam)
    assert _lib.lib().ec_adjoint_step_host(C.byref(d), U.ctypes.data_as(dp), lam.ctypes.data_as(dp),
                                           out.ctypes.data_as(dp), dt) == 0
    return out


CASES = [("hllc", 1, (True, True)), ("hllc", 1, (False, True)), ("rusanov", 0, (True, True)),
         ("rusanov", 0, (False, False)), ("xie_am_hllc_p", 3, (True, True)),
         ("xie_am_hllc_p", 3, (False, True))]


def grid_for(ecor, nx, ny, rid, periodic):
    bc = tuple(0 if periodic[k // 2] else 1 for k in range(4))
    return ecor.make_grid(nx, ny, 1.0 / nx, 1.0 / ny, riemann=rid, bc=bc)


@pytest.mark.parametrize("riemann,rid,periodic", CASES)
def test_torch_step_matches_oracle(ecor, riemann, rid, periodic):
    nx, ny = 20, 14
    U = smooth_state(nx, ny, 2)
    g = grid_for(ecor, nx, ny, rid, periodic)
    dt = 0.3 / ecor.max_rate(g, U)
    ref = ecor.fv_step(g, np.ascontiguousarray(U), dt)[0]
    got = torch_step.step(torch.from_numpy(U), dt, g.dx, g.dy, riemann=riemann, periodic=periodic).numpy()
    assert np.max(np.abs(got - ref) / (np.abs(ref) + 1e-300)) < 1e-13


@pytest.mark.parametrize("riemann,rid,periodic", CASES)
def test_one_step_adjoint_equals_autograd(ecor, riemann, rid, periodic):
    nx, ny = 20, 14
    U = smooth_state(nx, ny, 6)
    g = grid_for(ecor, nx, ny, rid, periodic)
    dt = 0.3 / ecor.max_rate(g, U)
    w = np.random.default_rng(3).standard_normal(U.shape)
    Ut = torch.from_numpy(U.copy()).requires_grad_(True)
    J = (torch.from_numpy(w) * torch_step.step(Ut, dt, g.dx, g.dy, riemann=riemann, periodic=periodic)).sum()
    (ref,) = torch.autograd.grad(J, Ut)
    lam = host_adjoint(g, U, w, dt)
    scale = np.abs(ref.numpy()).max()
    assert np.max(np.abs(lam - ref.numpy())) <= 1e-12 * scale


@pytest.mark.parametrize("riemann,rid", [("hllc", 1), ("xie_am_hllc_p", 3)])
def test_shock_bubble_rule_adjoint_equals_autograd(ecor, riemann, rid):
    """The shock-bubble ghost rule (fixed inflow, outflow, reflecting bottom):
    the torch step matches the oracle's, and the host adjoint step equals
    autograd through it."""
    from ecamd import _lib
    nx, ny = 20, 14
    U = smooth_state(nx, ny, 4)
    inflow = tuple(U[:, ny // 2, 0] * np.array([1.07, 0.9, 1.1, 1.05]))
    bcs = ("fixed", "outflow", "reflect", "outflow")
    g = ecor.make_grid(nx, ny, 1.0 / nx, 1.0 / ny, riemann=rid, bc=(3, 1, 2, 1),
                       fixed=(inflow, None, None, None))
    dt = 0.3 / ecor.max_rate(g, U)
    ref_step = ecor.fv_step(g, np.ascontiguousarray(U), dt)[0]
    fixed = (inflow, None, None, None)
    got = torch_step.step(torch.from_numpy(U), dt, g.dx, g.dy, riemann=riemann, bcs=bcs,
                          fixed=fixed).numpy()
    assert np.max(np.abs(got - ref_step) / (np.abs(ref_step) + 1e-300)) < 1e-13
    w = np.random.default_rng(2).standard_normal(U.shape)
    Ut = torch.from_numpy(U.copy()).requires_grad_(True)
    J = (torch.from_numpy(w) * torch_step.step(Ut, dt, g.dx, g.dy, riemann=riemann, bcs=bcs,
                                              fixed=fixed)).sum()
    (ref,) = torch.autograd.grad(J, Ut)
    d = _lib.GridDesc()
    d.nx, d.ny, d.dx, d.dy, d.gamma = nx, ny, g.dx, g.dy, g.gamma
    d.riemann, d.limiter, d.spatial_order = rid, 0, 2
    for k, b in enumerate((3, 1, 2, 1)):
        d.bc[k] = b
    for c in range(4):
        d.fixed_state[0][c] = inflow[c]
    lam = np.zeros_like(U)
    Uc, wc = np.ascontiguousarray(U), np.ascontiguousarray(w)
    assert _lib.lib().ec_adjoint_step_host(C.byref(d), Uc.ctypes.data_as(dp), wc.ctypes.data_as(dp),
                                           lam.ctypes.data_as(dp), dt) == 0
    assert np.max(np.abs(lam - ref.numpy())) <= 1e-12 * np.abs(ref.numpy()).max()


@pytest.mark.parametrize("riemann,rid,periodic", [CASES[0], CASES[1], CASES[4]])
def test_chained_adjoint_equals_autograd(ecor, riemann, rid, periodic):
    """Twelve steps with the oracle's CFL dts frozen: the chain of adjoint
    steps over the forward trajectory equals autograd through the chain."""
    nx, ny = 24, 16
    U0 = smooth_state(nx, ny, 9)
    g = grid_for(ecor, nx, ny, rid, periodic)
    _, res, dts = ecor.run_explicit(g, U0, 0.4, 1.0, 12)
    states = [np.ascontiguousarray(U0)]
    for dt in dts:
        states.append(ecor.fv_step(g, states[-1], dt)[0])
    w = np.random.default_rng(8).standard_normal(U0.shape)
    Ut = torch.from_numpy(U0.copy()).requires_grad_(True)
    V = Ut
    for dt in dts:
        V = torch_step.step(V, float(dt), g.dx, g.dy, riemann=riemann, periodic=periodic)
    (ref,) = torch.autograd.grad((torch.from_numpy(w) * V).sum(), Ut)
    lam = w.copy()
    for k in range(len(dts) - 1, -1, -1):
        lam = host_adjoint(g, states[k], lam, float(dts[k]))
    scale = np.abs(ref.numpy()).max()
    assert np.max(np.abs(lam - ref.numpy())) <= 1e-11 * scale
