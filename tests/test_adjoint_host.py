"""The four adjoint passes, run on the host with the device code's arithmetic
(ec_adjoint_step_host), against central differences of the oracle's step
(bitwise the device step): every cell and component on a small grid, for
each Riemann solver and both boundary families.  CPU only."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_adjoint import smooth_state

dp = C.POINTER(C.c_double)


def host_adjoint(g, U, w, dt):
    from ecamd import _lib
    d = _lib.GridDesc()
    d.nx, d.ny, d.dx, d.dy, d.gamma = g.nx, g.ny, g.dx, g.dy, g.gamma
    d.riemann, d.limiter, d.spatial_order = g.riemann, g.limiter, g.order
    for s in range(4):
        d.bc[s] = g.bc[s]
        for c in range(4):
            d.fixed_state[s][c] = g.fixed[s][c]
    lam = np.zeros_like(U)
    assert _lib.lib().ec_adjoint_step_host(C.byref(d), U.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                           lam.ctypes.data_as(dp), dt) == 0
    return lam


@pytest.mark.parametrize("riemann", [0, 1, 2, 3])
@pytest.mark.parametrize("bcs", [(0, 0, 0, 0), (3, 1, 2, 1)])
@pytest.mark.parametrize("limiter", [0, 1])
def test_adjoint_step_every_cell(ecor, riemann, bcs, limiter):
    nx, ny = 9, 7
    U = np.ascontiguousarray(smooth_state(nx, ny, 11))
    inflow = U[:, 3, 0] * np.array([1.07, 0.9, 1.1, 1.05])
    fixed = (inflow, None, None, None) if bcs[0] == 3 else None
    g = ecor.make_grid(nx, ny, 1 / nx, 1 / ny, riemann=riemann, limiter=limiter, bc=bcs, fixed=fixed)
    dt = 0.3 / ecor.max_rate(g, U)
    w = np.random.default_rng(2).standard_normal(U.shape)
    lam = host_adjoint(g, U, w, dt)
    J = lambda V: (w * ecor.fv_step(g, np.ascontiguousarray(V), dt)[0]).sum()
    fd = np.zeros_like(U)
    eps = 1e-7
    for idx in np.ndindex(U.shape):
        e = np.zeros_like(U)
        e[idx] = eps
        fd[idx] = (J(U + e) - J(U - e)) / (2 * eps)
    assert np.allclose(lam, fd, rtol=1e-6, atol=1e-6), np.abs(lam - fd).max()
