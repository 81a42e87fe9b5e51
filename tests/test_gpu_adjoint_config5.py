"""BASELINE configuration 5 at its full size: the checkpointed hand-written
adjoint of the explicit O2 HLLC run (single-gamma Mach 1.22 shock into a
light bubble, 640 x 160 on [0,4] x [0,1], 768 steps, checkpoints every 32)
against central finite differences on 16 seeded smooth directions.

The adjoint differentiates the forward map with dt and every branch frozen
at the forward trajectory (limiter choices, wave-speed minima and maxima,
Riemann regions).  Central differences of the plain forward cross those
branches' kinks wherever a +-eps perturbation flips one, which bounds their
agreement (to 4.3e-6 on the benchmark's 16 directions, 2.7e-7 on these), so
the bar (<= 1e-6) is checked against central
differences of the frozen-branch map: tests/torch_step.py records every
branch decision of the forward trajectory and replays them for all 32
perturbed runs at once, as one batched FP64 run on the GPU.  The plain
finite differences are reported beside it."""
import ctypes as C

import numpy as np
import pytest
import torch

import torch_step

pytestmark = pytest.mark.gpu

NX, NY, LX, LY = 640, 160, 4.0, 1.0
NSTEPS, K, CFL, NDIRS = 768, 32, 0.4, 16
BCS = ("fixed", "outflow", "reflect", "outflow")


def _setup(gpu):
    import ecamd
    from ecamd import problems
    U0 = problems.initial_state("shock_bubble_mach", NX, NY, prob_hi=(LX, LY), gamma=1.4,
                                shock_mach=1.22, shock_x=0.5, density_outer=1.0, pressure=1.0,
                                density_inner=0.138, blob_center=(1.0, 0.5), blob_radius=0.25)
    post = problems.post_shock_state(1.4, 1.22)
    inflow = (post[0], post[0] * post[1], 0.0, post[3] / 0.4 + 0.5 * post[0] * post[1] ** 2)
    grid = ecamd.Grid(NX, NY, LX / NX, LY / NY, 1.4, "hllc", "minmod", 2, BCS,
                      (inflow, None, None, None))
    s = ecamd.Solver(grid, gpu)
    rho = torch.from_numpy(U0[0]).to(gpu)
    u = torch.from_numpy(U0[1] / U0[0]).to(gpu)
    v = torch.from_numpy(U0[2] / U0[0]).to(gpu)
    p = torch.from_numpy(0.4 * (U0[3] - 0.5 * (U0[1] ** 2 + U0[2] ** 2) / U0[0])).to(gpu)
    x = (torch.arange(NX, device=gpu, dtype=torch.float64) + 0.5) * LX / NX
    y = (torch.arange(NY, device=gpu, dtype=torch.float64) + 0.5) * LY / NY
    X, Y = torch.meshgrid(x, y, indexing="xy")
    bubble = (((X - 1.0) ** 2 + (Y - 0.5) ** 2) < 0.25 ** 2).to(torch.float64)
    region = (((X - 1.0) ** 2 + (Y - 0.5) ** 2) < 0.4 ** 2).to(torch.float64)
    rho = rho * (1.0 + 0.05 * torch.sin(3 * X) * torch.cos(2 * Y) * bubble)   # seeded perturbation
    return s, inflow, rho, u, v, p, X, Y, bubble, region


def test_config5_adjoint_against_frozen_branch_differences(gpu):
    from ecamd._lib import lib
    from ecamd.adjoint import ExplicitRun, conserved_from_density, kinetic_energy_in_region, run_with_dts
    s, inflow, rho, u, v, p, X, Y, bubble, region = _setup(gpu)
    area = LX / NX * LY / NY

    def obj(U):   # bubble-region kinetic energy of each state of a batch
        return (0.5 * (U[1] ** 2 + U[2] ** 2) / U[0] * region).sum((-2, -1)) * area

    # the product: checkpointed forward, its dts, and the adjoint gradient
    print("config5: adjoint", flush=True)
    rho_req = rho.clone().requires_grad_(True)
    J = ExplicitRun.apply(rho_req, u, v, p, s, NSTEPS, K, CFL,
                          lambda U: kinetic_energy_in_region(U, region, area))
    (grad,) = torch.autograd.grad(J, rho_req)
    Ub = conserved_from_density(rho, u, v, p, 1.4)
    n = NX * NY
    chk = torch.empty(((NSTEPS + K - 1) // K) * 4 * n, dtype=torch.float64, device=gpu)
    UN = torch.empty_like(Ub)
    dts = np.zeros(NSTEPS)
    done = C.c_int()
    assert lib().ec_forward_checkpointed(s.ctx, C.c_void_p(Ub.data_ptr()), C.c_void_p(UN.data_ptr()),
                                         C.byref(s.layout), NSTEPS, K, CFL, 1e30,
                                         C.c_void_p(chk.data_ptr()),
                                         dts.ctypes.data_as(C.POINTER(C.c_double)), C.byref(done)) == 0
    assert done.value == NSTEPS

    # the forward again in torch, recording every branch decision
    print("config5: torch forward, recording branches", flush=True)
    B = torch_step.Branches("record")
    fallback = []
    with torch.no_grad():
        V = Ub[:, None].clone()                        # a batch of one
        for dt in dts:
            V = torch_step.step(V, float(dt), LX / NX, LY / NY, riemann="hllc", bcs=BCS,
                                fixed=(inflow, None, None, None), B=B, check=fallback)
    assert not any(bool(f.any()) for f in fallback), "a face took an HLLC fallback branch"
    dev_final = (V[:, 0] - UN).abs().reshape(4, -1).sum(1) / UN.abs().reshape(4, -1).sum(1)
    print("torch vs device final state, rel L1:", dev_final.tolist())
    assert float(dev_final.max()) <= 1e-10

    # 16 seeded smooth directions supported in the bubble
    rng = np.random.default_rng(20261204)
    dirs = []
    for _ in range(NDIRS):
        ph = rng.uniform(0, 2 * np.pi, 2)
        kx, ky = rng.integers(2, 7, 2)
        dirs.append(torch.sin(float(kx) * X + ph[0]) * torch.cos(float(ky) * Y + ph[1]) * bubble)
    D = torch.stack(dirs)                                 # (NDIRS, ny, nx)
    ad = (grad[None] * D).sum((-2, -1))

    # central differences of the frozen-branch map, every direction at once
    errs = {}
    with torch.no_grad():
        for eps in (1e-5, 1e-6):
            print(f"config5: frozen-branch replay, eps {eps:g}", flush=True)
            R = torch.cat([rho[None] + eps * D, rho[None] - eps * D])          # (2 NDIRS, ny, nx)
            V = conserved_from_density(R, u[None], v[None], p[None], 1.4)     # (4, 2 NDIRS, ny, nx)
            B.replay()
            for dt in dts:
                V = torch_step.step(V, float(dt), LX / NX, LY / NY, riemann="hllc", bcs=BCS,
                                    fixed=(inflow, None, None, None), B=B)
            assert B.pos == len(B.kept)
            Jv = obj(V)
            fd = (Jv[:NDIRS] - Jv[NDIRS:]) / (2 * eps)
            errs[eps] = ((ad - fd).abs() / ad.abs()).cpu().numpy()
            print(f"frozen-branch FD eps {eps:g}: rel err median {np.median(errs[eps]):.2e} "
                  f"max {errs[eps].max():.2e}")

    # plain central differences of the product's frozen-dt forward
    print("config5: plain differences through the product forward", flush=True)
    eps = 1e-7
    plain = []
    for k in range(NDIRS):
        Jp = float(kinetic_energy_in_region(run_with_dts(s, conserved_from_density(rho + eps * D[k], u, v, p, 1.4), dts), region, area))
        Jm = float(kinetic_energy_in_region(run_with_dts(s, conserved_from_density(rho - eps * D[k], u, v, p, 1.4), dts), region, area))
        plain.append(abs(float(ad[k]) - (Jp - Jm) / (2 * eps)) / abs(float(ad[k])))
    plain = np.array(plain)
    print(f"plain FD eps 1e-7: rel err median {np.median(plain):.2e} max {plain.max():.2e}")
    best = min(errs, key=lambda e: errs[e].max())
    assert errs[best].max() <= 1e-6, (best, errs[best])
    assert plain.max() <= 1e-5
