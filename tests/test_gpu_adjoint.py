"""The device adjoint of the explicit O2 step against central finite
differences of the frozen-dt forward map (<= 1e-6)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ecamd():
    import ecamd as m
    return m


def smooth_state(nx, ny, seed, lx=1.0, ly=1.0):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx * lx
    y = (np.arange(ny) + 0.5) / ny * ly
    X, Y = np.meshgrid(x, y)
    ph = rng.uniform(0, 2 * np.pi, 6)
    # generic state: no exact ties in the limiter arguments (a finite
    # difference must not straddle a minmod switch)
    noise = lambda: 0.02 * rng.standard_normal((ny, nx))
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X / lx + ph[0]) * np.cos(2 * np.pi * Y / ly + ph[1]) + noise()
    u = 0.4 * np.cos(2 * np.pi * Y / ly + ph[2]) + 0.2 + noise()
    v = -0.3 * np.sin(2 * np.pi * X / lx + ph[3]) + noise()
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (X / lx + Y / ly) + ph[4]) + noise()
    U = np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
    return U


@pytest.mark.parametrize("riemann", ["rusanov", "hllc", "lowmach_hllc", "xie_am_hllc_p"])
@pytest.mark.parametrize("bcs", [("periodic",) * 4, ("fixed", "outflow", "reflect", "outflow")])
def test_adjoint_step_matches_finite_differences(ecamd, gpu, riemann, bcs):
    from ecamd.adjoint import adjoint_step
    nx, ny = 20, 14
    U = smooth_state(nx, ny, 5)
    # an inflow state unlike any boundary cell (a ghost equal to its
    # neighbour would put a zero limiter argument at the base point)
    inflow = tuple(U[:, ny // 2, 0] * np.array([1.07, 0.9, 1.1, 1.05]))
    fixed = (inflow, None, None, None) if bcs[0] == "fixed" else (None,) * 4
    grid = ecamd.Grid(nx, ny, 1.0 / nx, 1.0 / ny, 1.4, riemann, "minmod", 2, bcs, fixed)
    s = ecamd.Solver(grid, gpu)
    Ud = torch.from_numpy(U).to(gpu)
    dt = 0.3 / s.diagnostics(Ud).max_rate
    rng = np.random.default_rng(1)
    w = torch.from_numpy(rng.standard_normal(U.shape)).to(gpu)
    lam = adjoint_step(s, Ud, w, dt)
    out = torch.empty_like(Ud)

    def J(V):
        s.step(V, out, dt, diag=False)
        return float((w * out).sum())

    for k in range(4):
        d = torch.from_numpy(rng.standard_normal(U.shape) * np.array([1, 0.3, 0.3, 2])[:, None, None]).to(gpu)
        eps = 1e-7
        fd = (J(Ud + eps * d) - J(Ud - eps * d)) / (2 * eps)
        ad = float((lam * d).sum())
        assert ad == pytest.approx(fd, rel=1e-6, abs=1e-9), (k, ad, fd)


def test_checkpointed_gradient_matches_fd(ecamd, gpu):
    """Small version of the adjoint configuration: a Mach 1.22 shock into a
    light bubble (HLLC O2), 48 steps with checkpoints every 16, kinetic
    energy in the bubble region; central differences on 4 smooth
    directions supported in the bubble."""
    from ecamd import problems
    from ecamd.adjoint import ExplicitRun, conserved_from_density, kinetic_energy_in_region, run_with_dts
    nx, ny = 128, 32
    lx, ly = 4.0, 1.0
    U0 = problems.initial_state("shock_bubble_mach", nx, ny, prob_hi=(lx, ly), gamma=1.4,
                                shock_mach=1.22, shock_x=0.5, density_outer=1.0, pressure=1.0,
                                density_inner=0.138, blob_center=(1.0, 0.5), blob_radius=0.25)
    post = problems.post_shock_state(1.4, 1.22)
    fixed = (tuple(np.array([post[0], post[0] * post[1], 0.0, post[3] / 0.4 + 0.5 * post[0] * post[1] ** 2])),
             None, None, None)
    grid = ecamd.Grid(nx, ny, lx / nx, ly / ny, 1.4, "hllc", "minmod", 2,
                      ("fixed", "outflow", "reflect", "reflect"), fixed)
    s = ecamd.Solver(grid, gpu)
    rho = torch.from_numpy(U0[0]).to(gpu)
    u = torch.from_numpy(U0[1] / U0[0]).to(gpu)
    v = torch.from_numpy(U0[2] / U0[0]).to(gpu)
    p = torch.from_numpy(0.4 * (U0[3] - 0.5 * (U0[1] ** 2 + U0[2] ** 2) / U0[0])).to(gpu)
    x = (torch.arange(nx, device=gpu, dtype=torch.float64) + 0.5) * lx / nx
    y = (torch.arange(ny, device=gpu, dtype=torch.float64) + 0.5) * ly / ny
    X, Y = torch.meshgrid(x, y, indexing="xy")
    bubble = (((X - 1.0) ** 2 + (Y - 0.5) ** 2) < 0.25 ** 2).to(torch.float64)
    region = (((X - 1.0) ** 2 + (Y - 0.5) ** 2) < 0.4 ** 2).to(torch.float64)
    rho = rho * (1.0 + 0.05 * torch.sin(3 * X) * torch.cos(2 * Y) * bubble)
    obj = lambda U: kinetic_energy_in_region(U, region, lx / nx * ly / ny)
    rho_req = rho.clone().requires_grad_(True)
    nsteps, K = 48, 16
    J = ExplicitRun.apply(rho_req, u, v, p, s, nsteps, K, 0.4, obj)
    (grad,) = torch.autograd.grad(J, rho_req)
    rng = np.random.default_rng(7)
    U_base = conserved_from_density(rho, u, v, p, 1.4)
    # the forward again, to fetch the dts it records (the frozen-dt map)
    import ctypes as C
    from ecamd._lib import lib
    n = nx * ny
    chk = torch.empty(((nsteps + K - 1) // K) * 4 * n, dtype=torch.float64, device=gpu)
    UN = torch.empty_like(U_base)
    dtv = np.zeros(nsteps)
    done = C.c_int()
    lib().ec_forward_checkpointed(s.ctx, C.c_void_p(U_base.data_ptr()), C.c_void_p(UN.data_ptr()),
                                  C.byref(s.layout), nsteps, K, 0.4, 1e30, C.c_void_p(chk.data_ptr()),
                                  dtv.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done))
    assert abs(float(obj(UN)) - float(J)) <= 1e-14 * abs(float(J))
    for k in range(4):
        ph = rng.uniform(0, 2 * np.pi, 2)
        d = (torch.sin(5 * X + ph[0]) * torch.cos(4 * Y + ph[1]) * bubble)
        eps = 1e-6
        Jp = float(obj(run_with_dts(s, conserved_from_density(rho + eps * d, u, v, p, 1.4), dtv)))
        Jm = float(obj(run_with_dts(s, conserved_from_density(rho - eps * d, u, v, p, 1.4), dtv)))
        fd = (Jp - Jm) / (2 * eps)
        ad = float((grad * d).sum())
        assert ad == pytest.approx(fd, rel=1e-6), (k, ad, fd)


@pytest.mark.parametrize("riemann", ["hllc", "xie_am_hllc_p"])
def test_checkpointed_sweep_equals_chained_steps(ecamd, gpu, riemann):
    """The checkpointed sweep (recompute without diagnostics, each step's
    fold sharing a launch with the next step's prep) gives bit for bit the
    lambda_0 of chaining ec_adjoint_step over the recorded trajectory."""
    import ctypes as C
    from ecamd._lib import lib
    from ecamd.adjoint import adjoint_step
    nx, ny = 40, 24
    U = torch.from_numpy(smooth_state(nx, ny, 3)).to(gpu)
    grid = ecamd.Grid(nx, ny, 1.0 / nx, 1.0 / ny, 1.4, riemann, "minmod", 2,
                      ("outflow", "outflow", "periodic", "periodic"))
    s = ecamd.Solver(grid, gpu)
    nsteps, K = 11, 4
    n = nx * ny
    chk = torch.empty(((nsteps + K - 1) // K) * 4 * n, dtype=torch.float64, device=gpu)
    UN = torch.empty_like(U)
    dts = np.zeros(nsteps)
    done = C.c_int()
    assert lib().ec_forward_checkpointed(s.ctx, C.c_void_p(U.data_ptr()), C.c_void_p(UN.data_ptr()),
                                         C.byref(s.layout), nsteps, K, 0.4, 1e30,
                                         C.c_void_p(chk.data_ptr()),
                                         dts.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done)) == 0
    assert done.value == nsteps
    lamN = torch.from_numpy(np.random.default_rng(4).standard_normal((4, ny, nx))).to(gpu)
    lam0 = torch.empty_like(lamN)
    assert lib().ec_adjoint_checkpointed(s.ctx, C.c_void_p(chk.data_ptr()), C.byref(s.layout), nsteps, K,
                                         dts.ctypes.data_as(C.POINTER(C.c_double)),
                                         C.c_void_p(lamN.data_ptr()), C.c_void_p(lam0.data_ptr())) == 0
    states = [U.clone()]
    out = torch.empty_like(U)
    for k in range(nsteps):
        s.step(states[-1], out, float(dts[k]), diag=False)
        states.append(out.clone())
    assert torch.equal(states[-1], UN)
    lam = lamN.clone()
    for k in range(nsteps - 1, -1, -1):
        lam = adjoint_step(s, states[k], lam, float(dts[k]))
    assert torch.equal(lam, lam0)


@pytest.mark.parametrize("riemann", ["rusanov", "hllc", "lowmach_hllc", "xie_am_hllc_p"])
@pytest.mark.parametrize("bcs", [("periodic",) * 4, ("fixed", "outflow", "reflect", "outflow")])
def test_device_adjoint_step_equals_host_passes(ecamd, gpu, riemann, bcs):
    """The device passes (prep, faces, ratios, LDS-tiled cells, ring fold)
    reproduce the host run of the same passes bit for bit; the host run is
    pinned against oracle finite differences on the CPU."""
    import ctypes as C
    from ecamd._lib import lib
    from ecamd.adjoint import adjoint_step
    nx, ny = 70, 37   # several cell tiles, ragged edges
    U = smooth_state(nx, ny, 8)
    inflow = tuple(U[:, ny // 2, 0] * np.array([1.07, 0.9, 1.1, 1.05]))
    fixed = (inflow, None, None, None) if bcs[0] == "fixed" else (None,) * 4
    grid = ecamd.Grid(nx, ny, 1.0 / nx, 1.0 / ny, 1.4, riemann, "minmod", 2, bcs, fixed)
    s = ecamd.Solver(grid, gpu)
    Ud = torch.from_numpy(U).to(gpu)
    dt = 0.3 / s.diagnostics(Ud).max_rate
    w = np.random.default_rng(5).standard_normal(U.shape)
    lam_dev = adjoint_step(s, Ud, torch.from_numpy(w).to(gpu), dt).cpu().numpy()
    lam_host = np.zeros_like(U)
    dp = C.POINTER(C.c_double)
    Uc = np.ascontiguousarray(U)
    assert lib().ec_adjoint_step_host(C.byref(grid.desc()), Uc.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                      lam_host.ctypes.data_as(dp), dt) == 0
    assert np.array_equal(lam_dev, lam_host)
