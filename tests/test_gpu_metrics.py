"""Whole-field metrics on the device against the oracle: admissibility
(numerics/diagnostic_metrics_timestep.hpp:49-102), Gresho metrics (:147-225)
and shock-bubble metrics (io/output_cases.hpp:329-407).

Bar: minima, maxima and counts bit for bit (order-independent reductions);
sums (errors, kinetic energies, bubble area and centroid) to relative 1e-12,
their only difference from the oracle being the summation order.  The
device's Gresho exact state uses its own log(), within an ulp of the host's.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GAMMA = 1.4


@pytest.fixture(scope="module")
def ecamd():
    import ecamd as m
    return m


def cons(r, u, v, p):
    return np.stack([r, r * u, r * v, p / (GAMMA - 1) + 0.5 * r * (u * u + v * v)])


def noisy_state(nx, ny, seed):
    rng = np.random.default_rng(seed)
    r = 1.0 + 0.5 * rng.random((ny, nx))
    u = rng.standard_normal((ny, nx))
    v = rng.standard_normal((ny, nx))
    p = 0.2 + rng.random((ny, nx))
    return cons(r, u, v, p)


def test_admissibility_matches_oracle(ecor, ecamd, gpu):
    nx, ny = 301, 67
    U = noisy_state(nx, ny, 1)
    U[:, 5, 7] = [np.nan, 0, 0, 1]
    U[:, 9, 11] = [-1.0, 0, 0, 1]
    U[:, 20, 30] = [1.0, 3.0, 0, 1.0]
    s = ecamd.ExplicitSolver(ecamd.Grid(nx, ny, 1 / nx, 1 / ny))
    for state in (U, noisy_state(nx, ny, 2)):
        d = s.admissibility(torch.from_numpy(state).to(gpu))
        (r, p, e), c = ecor.admissibility(ecor.make_grid(nx, ny, 1 / nx, 1 / ny), state)
        assert (d.rho_min, d.pressure_min, d.internal_energy_min) == (r, p, e)
        assert (d.nonfinite_count, d.density_failure_count, d.pressure_failure_count,
                d.internal_energy_failure_count) == c


@pytest.mark.parametrize("n,steps", [(64, 0), (96, 20)])
def test_gresho_metrics_match_oracle(ecor, ecamd, gpu, n, steps):
    from ecamd import problems
    mach = 0.1
    params = dict(vortex_center=(0.5, 0.5), density_outer=1.0, mach=mach, pressure=1.0 / (GAMMA * mach * mach))
    pd = problems.problem_desc("gresho_vortex", n, n, **params)
    U0 = problems.initial_state("gresho_vortex", n, n, **params)
    s = ecamd.ExplicitSolver(ecamd.Grid(n, n, 1 / n, 1 / n, GAMMA, "xie_am_hllc_p", "minmod", 2))
    A = torch.from_numpy(U0).to(gpu)
    if steps:
        B = torch.empty_like(A)
        A = s.advance(A, B, steps, s.new_run(0.4, 1.0, steps))
    U = A.cpu().numpy()
    m = s.gresho_metrics(A, pd)
    ref = ecor.gresho_metrics(ecor.make_grid(n, n, 1 / n, 1 / n), U, 0.0, 0.0, 0.5, 0.5, 1.0, mach,
                              params["pressure"])
    for k, want in ref.items():
        got = getattr(m, k)
        if want == 0.0:
            assert abs(got) <= 1e-15, k
        else:
            assert got == pytest.approx(want, rel=1e-12), k
    if steps:
        assert 0.99 < m.kinetic_energy_ratio < 1.0


def test_shock_bubble_metrics_match_oracle(ecor, ecamd, gpu):
    from ecamd import problems
    nx, ny = 160, 40
    U = problems.initial_state("shock_density_bubble", nx, ny, prob_lo=(0.0, 0.0), prob_hi=(2.0, 0.5))
    fixed = cons(2.8181818181818183, 1.6064386578049976, 0.0, 5.0)
    dx = 2.0 / nx
    grid = ecamd.Grid(nx, ny, dx, 0.5 / ny, GAMMA, "hllc", "minmod", 2,
                      ("fixed", "outflow", "reflect", "outflow"), (tuple(fixed), None, None, None))
    s = ecamd.ExplicitSolver(grid)
    A = torch.from_numpy(U).to(gpu)
    B = torch.empty_like(A)
    A = s.advance(A, B, 30, s.new_run(0.5, 1.0, 30))
    Uh = A.cpu().numpy()
    m = s.shock_bubble_metrics(A, (0.0, 0.0), 0.55)
    g = ecor.make_grid(nx, ny, dx, 0.5 / ny, bc=(3, 1, 2, 1), fixed=(list(fixed), None, None, None))
    ref = ecor.shock_bubble_metrics(g, Uh, 0.0, 0.0, 0.55)
    for k in ("rho_min", "rho_max", "pressure_min", "pressure_max", "internal_energy_min",
              "schlieren_max"):
        assert getattr(m, k) == ref[k], k
    assert m.nonfinite_count == ref["nonfinite_count"] == 0
    for k in ("density_bubble_area", "density_bubble_centroid_x", "density_bubble_centroid_y"):
        assert getattr(m, k) == pytest.approx(ref[k], rel=1e-12), k
    assert m.density_bubble_area > 0.0 and not math.isnan(m.density_bubble_centroid_x)
