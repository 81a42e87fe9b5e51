"""Exact-Dirichlet members in a batched ensemble (ec_batch_set_exact_boundary):
Toro test-1 strips of several resolutions with exact x sides and periodic y
(the report's Riemann convergence rows), a Gresho box with exact sides on
both axes and an advected blob entering through exact sides.

Bar: bit for bit.  Every member equals the single-grid device loop on that
member alone (ec_explicit_advance with ec_set_exact_boundary), in both round
modes, and the oracle's driver loop with exact ghosts (no Toro ghost lies in
a rarefaction fan before t = 0.2, so no pow() of the device library enters).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GAMMA = 1.4
CFL = 0.8
STOP = 0.2
TORO1 = ([1.0, 0.0, 0.0, 1.0], [0.125, 0.0, 0.0, 0.1])

# (name, kind, nx, ny, prob_hi, periodic, params)
CASES = [
    ("toro50", "toro1", 50, 4, (1.0, 4 / 50), (0, 1), dict(toro_test=1, riemann_interface_x=0.5)),
    ("toro100", "toro1", 100, 4, (1.0, 4 / 100), (0, 1), dict(toro_test=1, riemann_interface_x=0.5)),
    ("toro200", "toro1", 200, 4, (1.0, 4 / 200), (0, 1), dict(toro_test=1, riemann_interface_x=0.5)),
    ("gresho", "gresho_vortex", 32, 28, (1.0, 0.875), (0, 0),
     dict(vortex_center=(0.3, 0.35), density_outer=1.0, mach=0.1, pressure=1.0 / (GAMMA * 0.01))),
    ("blob", "advection_blob", 24, 16, (1.0, 0.7), (0, 1),
     dict(blob_center=(-0.1, 0.2), blob_radius=0.15, density_inner=2.0, density_outer=1.0,
          pressure=1.0, velocity=(1.0, 0.5))),
]


@pytest.fixture(scope="module")
def ecamd():
    import ecamd as m
    return m


def oracle_exact(ecor, kind, periodic, hi, p):
    if kind == "toro1":
        L, R = TORO1
        return ecor.make_exact(ecor.EXACT_TORO, periodic, prob_hi=hi, ql=L, qr=R,
                               interface_x=p["riemann_interface_x"])
    if kind == "gresho_vortex":
        return ecor.make_exact(ecor.EXACT_GRESHO, periodic, prob_hi=hi, vortex_center=p["vortex_center"],
                               density_outer=p["density_outer"], pressure=p["pressure"], mach=p["mach"])
    return ecor.make_exact(ecor.EXACT_BLOB, periodic, prob_hi=hi, blob_center=p["blob_center"],
                           blob_radius=p["blob_radius"], density_inner=p["density_inner"],
                           density_outer=p["density_outer"], pressure=p["pressure"], velocity=p["velocity"])


def placeholder_bc(periodic):
    return tuple("periodic" if periodic[a // 2] else "outflow" for a in range(4))


@pytest.fixture(scope="module")
def setup_cases():
    from ecamd import problems
    out = []
    for name, kind, nx, ny, hi, periodic, p in CASES:
        pd = problems.problem_desc(kind, nx, ny, prob_hi=hi, gamma=GAMMA, **p)
        U0 = problems.initial_state(kind, nx, ny, prob_hi=hi, gamma=GAMMA, **p)
        out.append((name, kind, nx, ny, hi, periodic, p, pd, U0))
    return out


@pytest.fixture(scope="module")
def single_grid(ecamd, gpu, setup_cases):
    """Each member alone through the single-grid device loop."""
    res = {}
    for riemann in ("hllc", "xie_am_hllc_p"):
        for name, kind, nx, ny, hi, periodic, p, pd, U0 in setup_cases:
            s = ecamd.ExplicitSolver(ecamd.Grid(nx, ny, hi[0] / nx, hi[1] / ny, GAMMA, riemann, "minmod", 2,
                                                placeholder_bc(periodic)))
            s.set_exact_boundary(pd, periodic)
            A = torch.from_numpy(U0.copy()).to(gpu)
            B = torch.empty_like(A)
            run = s.new_run(CFL, STOP, 100000)
            A = s.advance(A, B, 100000, run)
            res[(riemann, name)] = (A.cpu().numpy(), run.step, run.time)
    return res


@pytest.mark.parametrize("riemann", ["hllc", "xie_am_hllc_p"])
@pytest.mark.parametrize("mode", ["graph", "persistent"])
def test_exact_members_match_single_grid_loop(gpu, setup_cases, single_grid, riemann, mode):
    from ecamd.batch import Ensemble, Member
    print(f"ensemble exact {riemann} {mode}", flush=True)
    members = [Member(nx, ny, hi[0] / nx, hi[1] / ny, bc=placeholder_bc(periodic), cfl=CFL, stop_time=STOP,
                      max_step=100000)
               for _, _, nx, ny, hi, periodic, _, _, _ in setup_cases]
    ens = Ensemble(members, gamma=GAMMA, riemann=riemann, device=gpu)
    for k, (_, _, _, _, _, periodic, _, pd, U0) in enumerate(setup_cases):
        ens.set_exact_boundary(k, pd, periodic)
        ens.set_state(k, U0)
    ens.configure(mode, 0, 4)
    res, _ = ens.run()
    for k, (name, *_rest) in enumerate(setup_cases):
        U, steps, t = single_grid[(riemann, name)]
        assert res[k].status == 0 and res[k].step == steps and res[k].time == t, name
        assert np.array_equal(ens.get_state(k), U), name


def test_exact_members_match_oracle(ecor, setup_cases, single_grid):
    for name, kind, nx, ny, hi, periodic, p, pd, U0 in setup_cases:
        bc = tuple(0 if periodic[a // 2] else 1 for a in range(4))
        for riemann, rid in (("hllc", ecor.HLLC), ("xie_am_hllc_p", ecor.XIE_HLLCP)):
            g = ecor.set_exact(ecor.make_grid(nx, ny, hi[0] / nx, hi[1] / ny, riemann=rid, bc=bc),
                               oracle_exact(ecor, kind, periodic, hi, p), 0.0)
            U, r, _ = ecor.run_explicit(g, U0, CFL, STOP, 100000)
            got, steps, t = single_grid[(riemann, name)]
            assert steps == r.steps and t == r.time, (name, riemann)
            assert np.array_equal(got, U), (name, riemann, np.abs(got - U).max())


def test_set_exact_boundary_rejections(gpu, setup_cases):
    from ecamd import problems
    from ecamd._lib import EcError
    from ecamd.batch import Ensemble, Member
    ens = Ensemble([Member(16, 16, 1 / 16, 1 / 16)], gamma=GAMMA, device=gpu)
    with pytest.raises(EcError):      # another grid
        ens.set_exact_boundary(0, problems.problem_desc("gresho_vortex", 8, 16), (0, 0))
    with pytest.raises(EcError):      # no exact state
        ens.set_exact_boundary(0, problems.problem_desc("isentropic_vortex", 16, 16), (0, 0))
