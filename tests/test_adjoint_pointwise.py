"""The hand-written VJPs of the face fluxes against central finite
differences of the forward flux (host evaluation of the device arithmetic;
no GPU needed).  States are chosen away from branch switches so that the
frozen-branch derivative is the derivative."""
import ctypes as C

import numpy as np
import pytest

dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(dp)


def flux(lib, L, R, axis, kind, s):
    F = np.zeros(4)
    assert lib.ec_flux_host(_p(L), _p(R), axis, 1.4, kind, s, _p(F)) == 0
    return F


def vjp(lib, L, R, axis, kind, s, Fb):
    Lb, Rb, sb = np.zeros(4), np.zeros(4), C.c_double()
    assert lib.ec_flux_vjp_host(_p(L), _p(R), axis, 1.4, kind, s, _p(Fb), _p(Lb), _p(Rb),
                                C.byref(sb)) == 0
    return Lb, Rb, sb.value


STATES = [
    ([1.0, 0.3, -0.2, 1.0], [0.5, -0.1, 0.4, 0.6]),           # subsonic star, s_m > 0
    ([0.5, -0.3, 0.2, 0.6], [1.0, -0.2, -0.1, 1.1]),          # s_m < 0
    ([1.2, 0.05, 0.01, 71.43], [1.0, -0.02, 0.03, 71.40]),    # low Mach
    ([1.0, 1.5, 0.3, 1.0], [0.8, 0.2, -0.1, 0.9]),            # theta = 1
    ([1.0, 3.0, 0.1, 1.0], [0.5, 2.9, 0.02, 0.8]),            # supersonic: F = F(L)
]


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("axis", [0, 1])
@pytest.mark.parametrize("L,R", STATES)
def test_flux_vjp_matches_central_differences(kind, axis, L, R):
    from ecamd import _lib
    lib = _lib.lib()
    L, R = np.array(L), np.array(R)
    s = 0.37
    rng = np.random.default_rng(kind * 7 + axis)
    Fb = rng.standard_normal(4)
    Lb, Rb, sb = vjp(lib, L, R, axis, kind, s, Fb)
    x = np.concatenate([L, R, [s]])
    g = np.concatenate([Lb, Rb, [sb]])
    for _ in range(6):
        d = rng.standard_normal(9)
        if kind != 3:
            d[8] = 0.0
        h = 1e-6 * max(1.0, np.abs(x).max())
        xp, xm = x + h * d, x - h * d
        fp = flux(lib, xp[:4], xp[4:8], axis, kind, xp[8])
        fm = flux(lib, xm[:4], xm[4:8], axis, kind, xm[8])
        fd = Fb @ (fp - fm) / (2 * h)
        ad = g @ d
        assert ad == pytest.approx(fd, rel=2e-6, abs=2e-7 * max(1.0, np.abs(fp).max()))
