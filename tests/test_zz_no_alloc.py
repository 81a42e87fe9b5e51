"""ec_explicit_step_noalloc (the explicit step without a context or any
allocation) against ec_explicit_step through the C-ABI: the new state bit for
bit and the same diagnostics, on two calls in a row, the second on a smaller
grid, with the device's free memory unchanged by the calls; a grid beyond the
fixed workspace is rejected before any work is enqueued."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _perturbed(nx, ny, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.5 * (X > 0.55)
    u = 0.4 * np.cos(2 * np.pi * Y) + 0.05 * rng.standard_normal((ny, nx))
    v = -0.3 * np.sin(2 * np.pi * X) + 0.05 * rng.standard_normal((ny, nx))
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (X + Y)) + 2.0 * (Y < 0.3)
    return np.ascontiguousarray(np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)]))


def _compare(ecamd, _lib, grid, U, dt, gpu):
    lib = _lib.lib()
    desc = grid.desc()
    nx, ny = grid.nx, grid.ny
    lay = _lib.Layout(nx, nx * ny)
    Ud = torch.from_numpy(U).to(gpu)
    # reference: the context entry point
    s = ecamd.ExplicitSolver(grid, gpu)
    ref = torch.empty_like(Ud)
    dref = s.step(Ud, ref, dt)
    torch.cuda.synchronize()
    out = torch.full_like(Ud, float("nan"))
    d = _lib.StepDiag()
    free0 = torch.cuda.mem_get_info()[0]
    st = lib.ec_explicit_step_noalloc(C.byref(desc), C.c_void_p(Ud.data_ptr()),
                                      C.c_void_p(out.data_ptr()), C.byref(lay), dt, C.byref(d), None)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert st == 0, _lib.STATUS_TEXT.get(st, st)
    assert free1 == free0
    assert torch.equal(out, ref)
    assert (d.rho_min, d.pressure_min, d.nonfinite_count, d.degenerate_count, d.max_rate) == (
        dref.rho_min, dref.pressure_min, dref.nonfinite_count, dref.degenerate_count, dref.max_rate)
    s.close()
    return d


def test_noalloc_step_matches_context_step(gpu):
    import ecamd
    from ecamd import _lib
    lib = _lib.lib()
    # The HIP runtime may set up its own staging for the first device-to-host
    # copy into pageable memory: one unmeasured call first.
    g0 = ecamd.Grid(8, 8, 1.0 / 8, 1.0 / 8, 1.4, "rusanov", "minmod", 1)
    U0 = torch.from_numpy(_perturbed(8, 8, seed=1)).to(gpu)
    O0 = torch.empty_like(U0)
    d0 = _lib.StepDiag()
    lay0 = _lib.Layout(8, 64)
    assert lib.ec_explicit_step_noalloc(C.byref(g0.desc()), C.c_void_p(U0.data_ptr()),
                                        C.c_void_p(O0.data_ptr()), C.byref(lay0), 1e-4,
                                        C.byref(d0), None) == 0
    # first measured call: the largest grid the workspace takes (256 x 256
    # tiles of 16 x 15)
    nx, ny = 4096, 3840
    print("noalloc: 4096x3840 HLLC-P", flush=True)
    g1 = ecamd.Grid(nx, ny, 1.0 / nx, 1.0 / ny, 1.4, "xie_am_hllc_p", "minmod", 2)
    import ecor
    U1 = ecor.init_gresho(nx, ny, mach=0.01)
    dt1 = 0.4 / ecor.max_rate(ecor.make_grid(nx, ny, 1.0 / nx, 1.0 / ny, riemann=3), U1)
    _compare(ecamd, _lib, g1, U1, dt1, gpu)
    del U1
    # second call: a smaller grid with partial tiles and non-periodic sides
    print("noalloc: 37x29 HLLC", flush=True)
    g2 = ecamd.Grid(37, 29, 1.0 / 37, 0.7 / 29, 1.4, "hllc", "mc", 2,
                    ("periodic", "periodic", "outflow", "outflow"))
    U2 = _perturbed(37, 29, seed=5)
    d2 = _compare(ecamd, _lib, g2, U2, 0.05 / 40.0, gpu)
    assert d2.degenerate_count >= 0
    # beyond the workspace: rejected, nothing written
    big = g1.desc()
    big.ny = 3841
    out = torch.zeros(8, dtype=torch.float64, device=gpu)
    lay = _lib.Layout(nx, nx * 3841)
    st = lib.ec_explicit_step_noalloc(C.byref(big), C.c_void_p(out.data_ptr()),
                                      C.c_void_p(out.data_ptr() + 32), C.byref(lay), 1e-6, None, None)
    assert st == _lib.EC_ERR_TOO_LARGE
    assert lib.ec_status_string(st).startswith(b"grid exceeds")
