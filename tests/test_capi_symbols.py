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


def _desc(**kw):
    from ecamd import _lib
    d = _lib.GridDesc()
    d.nx, d.ny, d.dx, d.dy, d.gamma = 8, 8, 0.125, 0.125, 1.4
    d.riemann, d.limiter, d.spatial_order = 1, 0, 2
    for k, v in kw.items():
        if k == "bc":
            for s in range(4):
                d.bc[s] = v[s]
        else:
            setattr(d, k, v)
    return d


import pytest  # noqa: E402


@pytest.mark.parametrize("bad", [dict(nx=0), dict(ny=-3), dict(dx=0.0), dict(dy=float("nan")),
                                 dict(dx=float("inf")), dict(gamma=1.0), dict(riemann=4),
                                 dict(riemann=-1), dict(limiter=3), dict(spatial_order=3),
                                 dict(bc=(0, 1, 0, 0)), dict(bc=(1, 1, 2, 0)), dict(bc=(4, 4, 0, 0))])
def test_bad_descriptors_are_invalid_arguments(bad):
    """ec_create checks its descriptor before touching the device (status 1,
    EC_ERR_INVALID_ARGUMENT): sizes, cell sizes, gamma, selectors, and a
    periodic side paired with a non-periodic one."""
    from ecamd import _lib
    ctx = ctypes.c_void_p(12345)
    assert _lib.lib().ec_create(ctypes.byref(_desc(**bad)), None, ctypes.byref(ctx)) == 1
    assert not ctx.value


def test_bad_members_are_invalid_arguments():
    from ecamd import _lib
    from ecamd._abi_ext import MemberDesc
    L = _lib.lib()

    def member(**kw):
        m = MemberDesc()
        m.nx, m.ny, m.dx, m.dy, m.cfl, m.stop_time, m.max_step = 16, 4, 1 / 16, 1 / 16, 0.8, 0.1, 100
        for s, b in enumerate((1, 1, 0, 0)):
            m.bc[s] = b
        for k, v in kw.items():
            setattr(m, k, v)
        return m

    b = ctypes.c_void_p(7)
    for bad in (dict(nx=0), dict(cfl=0.0), dict(stop_time=float("nan")), dict(max_step=-1),
                dict(dx=-1.0)):
        arr = (MemberDesc * 2)(member(), member(**bad))
        assert L.ec_batch_create(arr, 2, 1.4, 1, 0, 2, None, ctypes.byref(b)) == 1
        assert not b.value
    arr = (MemberDesc * 1)(member())
    assert L.ec_batch_create(arr, 0, 1.4, 1, 0, 2, None, ctypes.byref(b)) == 1
    assert L.ec_batch_create(arr, 1, 1.4, 5, 0, 2, None, ctypes.byref(b)) == 1
    assert L.ec_batch_create(arr, 1, 1.0, 1, 0, 2, None, ctypes.byref(b)) == 1


def test_euler_dt_matches_the_oracle_on_edge_inputs(ecor):
    """compute_euler_dt's closing arithmetic (diagnostic_metrics_timestep.hpp:
    252-259) on zero, infinite, NaN and tiny rates and at the stop time."""
    from ecamd import _lib
    L = _lib.lib()
    inf, nan = float("inf"), float("nan")
    cases = [(4.0, 0.8, 0.0, 1.0), (0.0, 0.8, 0.0, 1.0), (inf, 0.8, 0.0, 1.0), (nan, 0.8, 0.3, 1.0),
             (1e-300, 0.8, 0.0, 1.0), (1e6, 0.4, 0.99999, 1.0), (2.0, 0.8, 1.0, 1.0),
             (-3.0, 0.8, 0.0, 1.0), (5.0, 0.0, 0.0, 1.0)]
    for c in cases:
        a, b = L.ec_euler_dt(*c), ecor.euler_dt(*c)
        assert (a == b) or (a != a and b != b), (c, a, b)
