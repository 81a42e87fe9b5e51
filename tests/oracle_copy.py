"""Builds a copy of the oracle with text edits planted in ec_oracle.c (in a
temporary directory, on the CPU), for tests that check what a change to the
oracle's arithmetic does.  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import shutil
import subprocess
from pathlib import Path

ORACLE = Path(__file__).resolve().parent.parent / "oracle"
SOURCES = ("ec_oracle.c", "ec_oracle_imex.c", "ec_oracle_direct.c")


def build_copy(root: Path, name: str, edits) -> Path:
    """edits: list of (text in ec_oracle.c, replacement); each text occurs once."""
    d = root / name
    d.mkdir()
    for s in SOURCES + ("ec_oracle.h",):
        shutil.copy(ORACLE / s, d / s)
    src = (d / "ec_oracle.c").read_text()
    for old, new in edits:
        assert src.count(old) == 1, (name, old)
        src = src.replace(old, new)
    (d / "ec_oracle.c").write_text(src)
    so = d / "libcopy.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-shared", "-o", str(so),
                    *[str(d / s) for s in SOURCES], "-lm"], check=True)
    return so


def load(so: Path, ecor):
    """The copy as a ctypes library with the oracle's argument types."""
    L = C.CDLL(str(so))
    dp = C.POINTER(C.c_double)
    G = C.POINTER(ecor.Grid)
    L.ecor_interface_flux.argtypes = [dp, dp, C.c_int, C.c_double, C.c_int, C.c_double, dp]
    L.ecor_face_flux.argtypes = [G, dp, C.c_int, C.c_int, C.c_int, dp]
    L.ecor_fv_step.argtypes = [G, dp, dp, C.c_double]
    L.ecor_fv_step.restype = C.c_int
    L.ecor_max_rate.argtypes = [G, dp]
    L.ecor_max_rate.restype = C.c_double
    return L
