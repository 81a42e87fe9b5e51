"""The drop-in driver's input checks (core/config_reader.hpp:118-322): each
rejected configuration aborts with the reference's message before any device
work, so these run on the CPU.  Skipped when the driver has not been built."""
import os
import subprocess
from pathlib import Path

import pytest

# EC_AMD_DRIVER: another build of the driver (the sanitizer build)
EXE = Path(os.environ.get("EC_AMD_DRIVER", Path(__file__).resolve().parent.parent
                          / "amrex-euler-allmach-explicit-imex_amd" / "euler_compare_amd"))

pytestmark = pytest.mark.skipif(not EXE.exists(), reason="driver not built")

IMEX = ("euler.method=imex", "euler.spatial_order=2", "euler.imex_acoustic_startup=0")
SB = ("euler.problem=shock_density_bubble_2d",)


@pytest.mark.parametrize("args,message", [
    (("euler.field_boundary=dirichlet",), "euler.field_boundary must be outflow or exact_dirichlet."),
    (("euler.imex_picard_extra_iterations=-1",), "euler.imex_picard_extra_iterations must be non-negative."),
    (IMEX + ("euler.imex_pressure_solver_path=eigen",),
     "IMEX rows require euler.imex_pressure_solver_path=host_single_level_gmres."),
    (SB + ("geometry.prob_hi=2.0 1.0",), "shock_density_bubble requires domain x=[0,2], y=[0,0.5]."),
    (SB + ("euler.gamma=1.3",), "shock_density_bubble is a same-gamma ideal-gas test with gamma=1.4."),
    (SB + ("amr.n_cell=8 40",), "shock_density_bubble requires at least 16x4 cells."),
    (SB + ("euler.riemann=rusanov",), "shock_density_bubble explicit rows support hllc or xie_am_hllc_p."),
    (("euler.imex_form=bdltv20_o1",), "Unknown euler.imex_form=bdltv20_o1"),
    (("euler.imex_predictor_dissipation=rusanov",), "Unknown euler.imex_predictor_dissipation=rusanov"),
    (IMEX + ("euler.bdltv20_paper_t1_s2=toro_xy_exact_dirichlet_2d", "euler.problem=toro1",
             "geometry.is_periodic=0 0", "euler.field_boundary=exact_dirichlet",
             "euler.bdltv20_paper_t1_s2_max_steps=0"),
     "euler.bdltv20_paper_t1_s2_max_steps must be at least one."),
    (("euler.problem=shock_density_bubble_cylindrical",) + IMEX,
     "cylindrical shock_density_bubble IMEX is not part of the cleaned Cartesian code pack."),
])
def test_rejected_configurations(tmp_path, args, message):
    inp = tmp_path / "inputs"
    inp.write_text("max_step = 8\nstop_time = 0.05\n")
    r = subprocess.run([str(EXE), str(inp), *args], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert r.returncode != 0
    assert message in r.stderr, r.stderr
