"""Independent checks of the low-Mach HLLC-P flux, its 5-face pressure sensor
and the low-Mach HLLC velocity blend (reference: numerics/
reconstruction_fluxes.hpp:212-252, 254-371, 390-440 of the euler_compare app).

Each check measures a flux implementation against what the mathematics or an
independent arrangement fixes, not against a retyped copy of the
implementation's own formula:

* roe_property_error -- the anti-diffusion term beta_p (1, u^, v^, |u^|^2/2)
  is recovered from F(f=0) - F(f=1) on an x-face and a y-face; that gives
  u^, v^ and, through the quadratic c^2 + |q^| c = K, the Roe sound speed and
  enthalpy.  Roe's property A(u^, v^, H^) (U_R - U_L) = F(U_R) - F(U_L) holds
  for the Roe-averaged state only (Roe, J. Comput. Phys. 43 (1981) 357;
  Toro, Riemann Solvers, 3rd ed., sec. 11.2), so arithmetic means, a dropped
  (1 + |M^|) factor, a wrong c^2 or a flipped beta_p sign all fail it.
* pressure_term_error -- the normal-momentum part of F(0) - F(1) minus
  beta_p u^ is (1 - theta)(p* - (p_l + p_r)/2), with p* taken from Toro's
  HLLC star pressure (an independent formula) and theta the clamped local
  Mach number built from the full speed.
* paper_form_error -- the flux arranged as Xie, Zhang, Lai and Li (Int. J.
  Numer. Meth. Fluids 89 (2019) 430-463) present it: Toro's HLLC flux, plus
  the sensor-weighted low-Mach pressure correction on the normal momentum,
  plus the pressure-difference dissipation term.
* sensor_errors -- the sensor value recovered at every face of a grid from
  the face flux (the flux is affine in f), against the cube of the minimum
  pressure ratio over the face and the four transverse faces of its two
  cells, from a reconstruction written here.
* low_mach_blend_error -- the low-Mach HLLC velocity blend at z = 1 (plain
  HLLC) and at an intermediate z.

``flux(qL, qR, axis, riemann, f)`` and ``face_flux(grid, U, fi, fj, axis)``
are the implementation under test (the oracle, or a deliberately broken copy
of it in the mutation test).
"""
from __future__ import annotations

import math

import numpy as np

G = 1.4
RUSANOV, HLLC, LOWMACH_HLLC, XIE_HLLCP = 0, 1, 2, 3


def conserved(q, g=G):
    r, u, v, p = q
    return np.array([r, r * u, r * v, p / (g - 1) + 0.5 * r * (u * u + v * v)])


def euler_flux(q, axis, g=G):
    r, u, v, p = q
    E = p / (g - 1) + 0.5 * r * (u * u + v * v)
    un = u if axis == 0 else v
    f = np.array([r * un, r * u * un, r * v * un, (E + p) * un])
    f[1 + axis] += p
    return f


def davis_speeds(L, R, axis, g=G):
    aL, aR = math.sqrt(g * L[3] / L[0]), math.sqrt(g * R[3] / R[0])
    uL, uR = L[1 + axis], R[1 + axis]
    return min(uL - aL, uR - aR), max(uL + aL, uR + aR)


def toro_star(L, R, axis, g=G):
    """Toro's HLLC contact speed S* and star pressure (eqs 10.37, 10.36)."""
    SL, SR = davis_speeds(L, R, axis, g)
    uL, uR = L[1 + axis], R[1 + axis]
    Ss = (R[3] - L[3] + L[0] * uL * (SL - uL) - R[0] * uR * (SR - uR)) / (
        L[0] * (SL - uL) - R[0] * (SR - uR))
    ps = L[3] + L[0] * (SL - uL) * (Ss - uL)
    return Ss, ps


def toro_hllc(L, R, axis, g=G):
    """HLLC in Toro's textbook form F_K + S_K (U*_K - U_K) (eqs 10.38-10.40,
    10.73) with the Davis wave-speed estimates."""
    SL, SR = davis_speeds(L, R, axis, g)
    uL, uR = L[1 + axis], R[1 + axis]
    if SL >= 0:
        return euler_flux(L, axis, g)
    if SR <= 0:
        return euler_flux(R, axis, g)
    Ss, _ = toro_star(L, R, axis, g)
    K, SK, uK = (L, SL, uL) if Ss >= 0 else (R, SR, uR)
    UK = conserved(K, g)
    fac = K[0] * (SK - uK) / (SK - Ss)
    Us = fac * np.array([1.0, K[1], K[2], UK[3] / K[0] + (Ss - uK) * (Ss + K[3] / (K[0] * (SK - uK)))])
    Us[1 + axis] = fac * Ss
    return euler_flux(K, axis, g) + SK * (Us - UK)


def flux_jacobian(u, v, H, axis, g=G):
    """Jacobian dF/dU of the 2-D Euler flux written in (u, v, H) (Toro,
    sec. 3.2.2; Hirsch, vol. 2, sec. 16.1)."""
    gm = g - 1.0
    phi = 0.5 * gm * (u * u + v * v)
    if axis == 0:
        return np.array([
            [0.0, 1.0, 0.0, 0.0],
            [phi - u * u, (3.0 - g) * u, -gm * v, gm],
            [-u * v, v, u, 0.0],
            [u * (phi - H), H - gm * u * u, -gm * u * v, g * u]])
    return np.array([
        [0.0, 0.0, 1.0, 0.0],
        [-u * v, v, u, 0.0],
        [phi - v * v, -gm * u, (3.0 - g) * v, gm],
        [v * (phi - H), -gm * u * v, H - gm * v * v, g * v]])


def random_subsonic_pairs(n, seed, g=G):
    """(L, R) primitive pairs with a pressure jump, a transverse velocity and
    s_l < 0 < s_r on both axes (the main branch of the flux)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        L = np.array([rng.uniform(0.2, 3.0), rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6),
                      rng.uniform(0.3, 3.0)])
        R = np.array([rng.uniform(0.2, 3.0), rng.uniform(-0.6, 0.6), rng.uniform(-0.6, 0.6),
                      rng.uniform(0.3, 3.0)])
        if abs(L[3] - R[3]) < 0.05:
            continue
        ok = True
        for ax in (0, 1):
            sl, sr = davis_speeds(L, R, ax, g)
            ok = ok and sl < -0.05 and sr > 0.05
        if ok:
            out.append((L, R))
    return out


def recover_roe_state(flux, L, R, g=G):
    """u^, v^, c^ and H^ recovered from the flux's beta_p term; also the
    y-face's c^ as a cross-check.  Returns a dict."""
    dx = flux(L, R, 0, XIE_HLLCP, 0.0) - flux(L, R, 0, XIE_HLLCP, 1.0)
    dy = flux(L, R, 1, XIE_HLLCP, 0.0) - flux(L, R, 1, XIE_HLLCP, 1.0)
    v_hat = dx[2] / dx[0]
    u_hat = dy[1] / dy[0]
    ke_x, ke_y = dx[3] / dx[0], dy[3] / dy[0]

    def c_from(d, q_hat, axis):
        sl, sr = davis_speeds(L, R, axis, g)
        K = -sl * sr * (R[3] - L[3]) / ((sr - sl) * d[0])   # c^2 (1 + |q|/c)
        disc = q_hat * q_hat + 4.0 * K
        if not (K > 0.0 and disc >= 0.0):
            return float("nan")
        return 0.5 * (-abs(q_hat) + math.sqrt(disc))

    c_x = c_from(dx, u_hat, 0)
    c_y = c_from(dy, v_hat, 1)
    H = c_x * c_x / (g - 1.0) + 0.5 * (u_hat * u_hat + v_hat * v_hat)
    return dict(u=u_hat, v=v_hat, c=c_x, c_y=c_y, H=H, ke_x=ke_x, ke_y=ke_y, dx=dx, dy=dy)


def roe_property_error(flux, L, R, g=G) -> float:
    """Largest relative violation of Roe's property on both axes, plus the
    consistency of the two faces' recovered sound speed and |u^|^2/2."""
    s = recover_roe_state(flux, L, R, g)
    if not all(math.isfinite(s[k]) for k in ("u", "v", "c", "c_y", "H")):
        return float("inf")
    dU = conserved(R, g) - conserved(L, g)
    err = 0.0
    for axis in (0, 1):
        dF = euler_flux(R, axis, g) - euler_flux(L, axis, g)
        A = flux_jacobian(s["u"], s["v"], s["H"], axis, g)
        err = max(err, float(np.max(np.abs(A @ dU - dF)) / np.max(np.abs(dF))))
    half_q2 = 0.5 * (s["u"] ** 2 + s["v"] ** 2)
    err = max(err, abs(s["c"] - s["c_y"]) / s["c"], abs(s["ke_x"] - half_q2) / half_q2,
              abs(s["ke_y"] - half_q2) / half_q2)
    return err


def theta_local_mach(L, R, g=G):
    """clamp01 of the larger local Mach number |u|/c of the two states."""
    m = max(math.hypot(L[1], L[2]) / math.sqrt(g * L[3] / L[0]),
            math.hypot(R[1], R[2]) / math.sqrt(g * R[3] / R[0]))
    return min(max(m, 0.0), 1.0)


def pressure_term_error(flux, L, R, g=G) -> float:
    """Normal-momentum part of F(0) - F(1), less beta_p u^ (resp. v^), against
    (1 - theta)(p* - (p_l + p_r)/2) with Toro's p*."""
    s = recover_roe_state(flux, L, R, g)
    th = theta_local_mach(L, R, g)
    err = 0.0
    for axis, d, qh in ((0, s["dx"], s["u"]), (1, s["dy"], s["v"])):
        _, ps = toro_star(L, R, axis, g)
        want = (1.0 - th) * (ps - 0.5 * (L[3] + R[3]))
        got = d[1 + axis] - d[0] * qh
        err = max(err, abs(got - want) / max(abs(want), 1e-3 * abs(ps)))
    return err


def paper_form(L, R, axis, f, g=G):
    """Xie et al.'s HLLC-P flux arranged as HLLC + pressure correction +
    pressure-difference dissipation, with textbook Roe averages."""
    sl, sr = davis_speeds(L, R, axis, g)
    _, ps = toro_star(L, R, axis, g)
    wl, wr = math.sqrt(L[0]), math.sqrt(R[0])
    HL = (conserved(L, g)[3] + L[3]) / L[0]
    HR = (conserved(R, g)[3] + R[3]) / R[0]
    u = (wl * L[1] + wr * R[1]) / (wl + wr)
    v = (wl * L[2] + wr * R[2]) / (wl + wr)
    H = (wl * HL + wr * HR) / (wl + wr)
    c2 = (g - 1.0) * (H - 0.5 * (u * u + v * v))
    c = math.sqrt(c2)
    qn = u if axis == 0 else v
    th = theta_local_mach(L, R, g)
    n = np.zeros(4)
    n[1 + axis] = 1.0
    correction = f * (1.0 - th) * (0.5 * (L[3] + R[3]) - ps) * n
    beta = (f - 1.0) * sl * sr * (R[3] - L[3]) / ((sr - sl) * (c2 + abs(qn) * c))
    dissipation = beta * np.array([1.0, u, v, 0.5 * (u * u + v * v)])
    return toro_hllc(L, R, axis, g) + correction + dissipation


def paper_form_error(flux, L, R, axis, f, g=G) -> float:
    want = paper_form(L, R, axis, f, g)
    got = flux(L, R, axis, XIE_HLLCP, f)
    return float(np.max(np.abs(got - want)) / np.max(np.abs(want)))


# ---------------------------------------------------------------- the sensor
def minmod(a, b):
    return np.where(a * b <= 0.0, 0.0, np.where(np.abs(a) < np.abs(b), a, b))


def periodic_face_states(Q, order):
    """Primitive face states of a periodic (4, ny, nx) field: for each axis,
    (minus-side state, plus-side state) of every cell, q -/+ slope/2."""
    out = {}
    for axis, ax in ((0, 2), (1, 1)):
        if order == 1:
            out[axis] = (Q, Q)
            continue
        back = Q - np.roll(Q, 1, axis=ax)
        fwd = np.roll(Q, -1, axis=ax) - Q
        s = minmod(back, fwd)
        out[axis] = (Q - 0.5 * s, Q + 0.5 * s)
    return out


def sensor_definition(Q, order):
    """Sensor of every x-face (fi, j) between cells fi-1 and fi and every
    y-face (i, fj) between cells fj-1 and fj of a periodic grid: the cube of
    the smallest pressure ratio over the face and the four faces of the other
    axis that bound its two cells."""
    fs = periodic_face_states(Q, order)
    # ratio of the x-face at the low side of cell (i, j): cells i-1 | i
    px_lo_side = np.roll(fs[0][1][3], 1, axis=1)   # plus-side state of cell i-1
    px_hi_side = fs[0][0][3]                        # minus-side state of cell i
    rx = np.minimum(px_lo_side, px_hi_side) / np.maximum(px_lo_side, px_hi_side)
    py_lo_side = np.roll(fs[1][1][3], 1, axis=0)
    py_hi_side = fs[1][0][3]
    ry = np.minimum(py_lo_side, py_hi_side) / np.maximum(py_lo_side, py_hi_side)
    # rx[j, i]: x-face at the low side of cell (i, j); ry[j, i]: y-face at the low side
    ny, nx = rx.shape
    fx = np.empty((ny, nx))
    fy = np.empty((ny, nx))
    for j in range(ny):
        for i in range(nx):
            cells = ((i - 1) % nx, i)               # the x-face's two cells
            m = rx[j, i]
            for c in cells:
                m = min(m, ry[j, c], ry[(j + 1) % ny, c])
            fx[j, i] = min(max(m, 0.0), 1.0) ** 3
            cells = ((j - 1) % ny, j)               # the y-face's two cells
            m = ry[j, i]
            for r in cells:
                m = min(m, rx[r, i], rx[r, (i + 1) % nx])
            fy[j, i] = min(max(m, 0.0), 1.0) ** 3
    return fx, fy, fs


def sensor_errors(flux, face_flux, grid, U, Q, order):
    """Largest |f_recovered - f_definition| over the x- and y-faces of a
    periodic grid, where f is recovered from the implementation's face flux
    through the flux's affine dependence on f."""
    fx, fy, fs = sensor_definition(Q, order)
    ny, nx = fx.shape
    worst = 0.0
    for j in range(ny):
        for i in range(nx):
            for axis, want in ((0, fx[j, i]), (1, fy[j, i])):
                if axis == 0:
                    L = fs[0][1][:, j, (i - 1) % nx]
                    R = fs[0][0][:, j, i]
                else:
                    L = fs[1][1][:, (j - 1) % ny, i]
                    R = fs[1][0][:, j, i]
                F = face_flux(grid, U, i, j, axis)
                F0 = flux(L, R, axis, XIE_HLLCP, 0.0)
                F1 = flux(L, R, axis, XIE_HLLCP, 1.0)
                d = F1 - F0
                got = float(np.dot(F - F0, d) / np.dot(d, d))
                worst = max(worst, abs(got - want))
    return worst


def sensor_field(nx, ny, seed, g=G):
    """A periodic primitive field with pressures spread over [1, 3] (so every
    face ratio differs) and small velocities; returns (Q, U) as (4, ny, nx)."""
    rng = np.random.default_rng(seed)
    Q = np.empty((4, ny, nx))
    Q[0] = rng.uniform(0.8, 1.2, (ny, nx))
    Q[1] = rng.uniform(-0.1, 0.1, (ny, nx))
    Q[2] = rng.uniform(-0.1, 0.1, (ny, nx))
    Q[3] = rng.uniform(1.0, 3.0, (ny, nx))
    U = np.stack([conserved(Q[:, j, i], g) for j in range(ny) for i in range(nx)], axis=1)
    return Q, np.ascontiguousarray(U.reshape(4, ny, nx))


# ---------------------------------------------------------- low-Mach HLLC
def low_mach_blend_error(flux, g=G) -> float:
    """Low-Mach HLLC: at z = 1 (one state locally supersonic) the velocity
    blend is the identity and the flux is plain HLLC; at z = 1/2 each state's
    velocity moves halfway toward the pair's mean."""
    err = 0.0
    # z = 1: |u_L| > c_L through the transverse velocity, face still subsonic
    L = np.array([1.0, 0.2, 1.5, 1.0])
    R = np.array([0.8, -0.1, 0.3, 0.9])
    for axis in (0,):
        want = toro_hllc(L, R, axis, g)
        got = flux(L, R, axis, LOWMACH_HLLC, 1.0)
        err = max(err, float(np.max(np.abs(got - want)) / np.max(np.abs(want))))
    # z = 1/2: local Mach numbers 0.5 (left) and 0.25 (right), exactly
    cL, cR = math.sqrt(g * 2.0 / 1.0), math.sqrt(g * 1.0 / 0.5)
    L = np.array([1.0, 0.3 * cL, 0.4 * cL, 2.0])       # |u| = 0.5 c
    R = np.array([0.5, -0.15 * cR, -0.2 * cR, 1.0])    # |u| = 0.25 c
    z = 0.5
    mean = 0.5 * (L + R)
    Lb, Rb = L.copy(), R.copy()
    for c in (1, 2):
        Lb[c] = mean[c] + 0.5 * z * (L[c] - R[c])
        Rb[c] = mean[c] + 0.5 * z * (R[c] - L[c])
    for axis in (0, 1):
        want = toro_hllc(Lb, Rb, axis, g)
        got = flux(L, R, axis, LOWMACH_HLLC, 1.0)
        err = max(err, float(np.max(np.abs(got - want)) / np.max(np.abs(want))))
    return err
