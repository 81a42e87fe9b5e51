"""The seeded ensemble fixture and what pins the ensemble's physics on the
CPU: the table regenerates from its seed, and the oracle's O2 HLLC and
HLLC-P runs of the members converge to the exact Riemann solution
(Toro ch. 4) as the grid is refined."""
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).parent / "golden"


def load_table():
    return json.loads((GOLDEN / "ensemble_64.json").read_text())


def test_table_regenerates_from_its_seed():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_ensemble", GOLDEN / "make_ensemble.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    t = load_table()
    assert t["seed"] == mod.SEED
    assert t["members"] == mod.draw()
    for m in t["members"]:
        assert m["nx"] in mod.NX_CHOICES
        for q in (m["left"], m["right"]):
            assert 0.1 <= q[0] <= 2.0 and -1.0 <= q[1] <= 1.0 and q[2] == 0.0 and 0.1 <= q[3] <= 10.0


def strip_run(ecor, m, nx, riemann, t_end):
    g = ecor.make_grid(nx, 4, 1.0 / nx, 1.0 / nx, riemann=riemann, bc=(1, 1, 0, 0))
    U0 = ecor.init_toro(nx, 4, m["left"], m["right"], dx=1.0 / nx)
    U, res, _ = ecor.run_explicit(g, U0, 0.8, t_end, 100000)
    return U, res


def exact_density(ecor, m, nx, t):
    x = (np.arange(nx) + 0.5) / nx
    return np.array([ecor.exact_riemann_sample(xi, t, m["left"], m["right"])[0] for xi in x])


@pytest.mark.parametrize("riemann", [1, 3])
@pytest.mark.parametrize("member", [0, 5, 17])
def test_members_converge_to_exact_riemann(ecor, riemann, member):
    m = load_table()["members"][member]
    t_end = 0.05
    errs = []
    for nx in (100, 200, 400):
        U, res = strip_run(ecor, m, nx, riemann, t_end)
        assert res.status == 0 and abs(res.time - t_end) < 1e-12
        # The strip is uniform in y.
        assert np.array_equal(U[:, 0, :], U[:, 3, :])
        errs.append(np.abs(U[0, 0] - exact_density(ecor, m, nx, t_end)).mean())
    # Discontinuous solutions: L1 error falls roughly like h^(1/2)..h^1.
    assert errs[1] < 0.85 * errs[0] and errs[2] < 0.85 * errs[1], errs
