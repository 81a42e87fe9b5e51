"""MI355X-native euler_compare update path: host layer over libec_amd.so."""
from ._lib import EcError, lib  # noqa: F401
from ._abi_ext import IMEX_STATUS  # noqa: F401
from .explicit import ExplicitSolver, Grid, Solver, run_explicit  # noqa: F401
