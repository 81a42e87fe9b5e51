"""The product library loads on a CPU-only host and exports every entry point
that include/*.h declares (no compute calls: there is no GPU here)."""
import ctypes
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    from ecamd import _lib
    lib = _lib.lib()
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        names += _lib.exported_symbols(h)
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_only_entry_points():
    from ecamd import _lib
    lib = _lib.lib()
    assert lib.ec_status_string(0) == b"ok"
    assert lib.ec_euler_dt(4.0, 0.8, 0.0, 1.0) == 0.2
    assert lib.ec_explicit_tile_count() >= 1
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.ec_explicit_tile_shape(0, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
    assert c.value % 64 == 0 and (a.value + 1) * b.value <= c.value * 2
    assert lib.ec_explicit_tile_shape(99, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 1


def test_invalid_descriptor_is_rejected_without_gpu():
    from ecamd import _lib
    lib = _lib.lib()
    d = _lib.GridDesc()
    d.nx, d.ny, d.dx, d.dy, d.gamma, d.spatial_order = 0, 4, 0.1, 0.1, 1.4, 2
    ctx = ctypes.c_void_p()
    assert lib.ec_create(ctypes.byref(d), None, ctypes.byref(ctx)) == 1
    assert not ctx.value
