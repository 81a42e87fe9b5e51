"""A differentiable FP64 torch version of the explicit O2 step, for reference
gradients and frozen-branch finite differences (test infrastructure only).

torch.autograd differentiates it independently of the hand-written adjoint
(csrc/ec_adjoint.hpp).  It covers the periodic, outflow, reflecting and fixed
ghost rules (problems/boundary_conditions.hpp:2-73), the minmod MUSCL
reconstruction in primitive variables, and the Rusanov, HLLC and Xie AM-HLLC-P
fluxes with the 5-face pressure sensor, written from the formulas of
numerics/reconstruction_fluxes.hpp:66-210, 254-440 (main branches; the
degenerate fallbacks are not taken by the test states, and `check_main`
asserts that).

Every state-dependent branch (limiter choices, wave-speed minima and maxima,
Riemann regions, the sensor's minima, clamps and absolute values) goes
through a Branches object, which can record the decisions of a base
trajectory and replay them: the replayed map is the forward map with every
branch frozen, the map the hand-written adjoint differentiates, so finite
differences of it carry no kink-crossing error.  States may carry batch
dimensions between the component axis and the grid: (4, ..., ny, nx);
recorded decisions of an unbatched run broadcast over them.
"""
from __future__ import annotations

import torch


class Branches:
    """mode None: plain branches; "record": plain branches, decisions kept in
    order; "replay": the kept decisions, in the same order."""

    def __init__(self, mode: str | None = None):
        self.mode = mode
        self.kept: list[torch.Tensor] = []
        self.pos = 0

    def cond(self, c: torch.Tensor) -> torch.Tensor:
        if self.mode == "record":
            self.kept.append(c.detach())
            return c
        if self.mode == "replay":
            k = self.kept[self.pos]
            self.pos += 1
            return k
        return c

    def replay(self) -> "Branches":
        self.mode, self.pos = "replay", 0
        return self

    def where(self, c, a, b):
        return torch.where(self.cond(c), a, b)

    def min(self, a, b):      # std::min: b if b < a else a
        return self.where(b < a, b, a)

    def max(self, a, b):      # std::max: b if a < b else a
        return self.where(a < b, b, a)

    def abs(self, x):
        return self.where(x < 0.0, -x, x)

    def clamp01(self, x):
        return self.min(self.max(x, torch.zeros_like(x)), torch.ones_like(x))


PLAIN = Branches()


# ------------------------------------------------------------------ ghosts
def _axis_index(n: int, lo: str, hi: str):
    """Source index and reflection sign of padded positions -2..n+1 of one
    axis (the map_axis rule: wrap, mirror, else clamp)."""
    idx, sgn = [], []
    for i in range(-2, n + 2):
        s = 1.0
        if i < 0:
            if lo == "periodic":
                idx.append(i % n); sgn.append(1.0); continue
            if lo == "reflect":
                s, i = -1.0, -i - 1
        elif i >= n:
            if hi == "periodic":
                idx.append(i % n); sgn.append(1.0); continue
            if hi == "reflect":
                s, i = -1.0, 2 * n - 1 - i
        idx.append(min(max(i, 0), n - 1))
        sgn.append(s)
    return idx, sgn


def pad(U: torch.Tensor, bcs=("periodic",) * 4, fixed=(None,) * 4) -> torch.Tensor:
    """(4, ..., ny, nx) -> (4, ..., ny+4, nx+4) through the ghost rule; a FIXED
    side's ghost cells (corners included) hold its conserved state, xlo over
    xhi over ylo over yhi."""
    ny, nx = U.shape[-2:]
    ix, sx = _axis_index(nx, bcs[0], bcs[1])
    iy, sy = _axis_index(ny, bcs[2], bcs[3])
    dev = U.device
    P = U.index_select(-2, torch.tensor(iy, device=dev)).index_select(-1, torch.tensor(ix, device=dev))
    sxv = torch.tensor(sx, dtype=U.dtype, device=dev)
    syv = torch.tensor(sy, dtype=U.dtype, device=dev)
    P = torch.stack([P[0], P[1] * sxv, P[2] * syv[:, None], P[3]])
    for side in (3, 2, 1, 0):            # lowest priority first
        if bcs[side] != "fixed":
            continue
        st = torch.tensor(fixed[side], dtype=U.dtype, device=dev).reshape((4,) + (1,) * (P.dim() - 1))
        mask = torch.zeros(P.shape[-2:], dtype=torch.bool, device=dev)
        if side == 0:
            mask[:, :2] = True
        elif side == 1:
            mask[:, -2:] = True
        elif side == 2:
            mask[:2, :] = True
        else:
            mask[-2:, :] = True
        P = torch.where(mask, st, P)
    return P


# ----------------------------------------------------------- state algebra
def to_prim(U, gamma):
    rho = U[0]
    inv = 1.0 / rho
    u = U[1] * inv
    v = U[2] * inv
    p = (gamma - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v))
    return torch.stack([rho, u, v, p])


def cons(q, gamma):
    rho, u, v, p = q
    return torch.stack([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])


def minmod(b, f, B: Branches = PLAIN):
    zero = torch.zeros_like(b)
    return B.where(b * f <= 0.0, zero, B.where(b.abs() < f.abs(), b, f))


def flux_phys(q, axis, gamma):
    rho, u, v, p = q
    mx, my = rho * u, rho * v
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    if axis == 0:
        return torch.stack([mx, mx * u + p, my * u, (E + p) * u])
    return torch.stack([my, mx * v, my * v + p, (E + p) * v])


def sound(q, gamma):
    return torch.sqrt(gamma * q[3] / q[0])


def _supersonic(L, R, axis, gamma, s_l, s_r, star, B):
    out = B.where(s_r <= 0.0, flux_phys(R, axis, gamma), star)
    return B.where(s_l >= 0.0, flux_phys(L, axis, gamma), out)


def rusanov(L, R, axis, gamma, B: Branches = PLAIN, sensor=None):
    cl, cr = sound(L, gamma), sound(R, gamma)
    s = B.max(B.abs(L[1 + axis]) + cl, B.abs(R[1 + axis]) + cr)
    return 0.5 * (flux_phys(L, axis, gamma) + flux_phys(R, axis, gamma)) - 0.5 * s * (cons(R, gamma) - cons(L, gamma))


def hllc(L, R, axis, gamma, B: Branches = PLAIN, sensor=None, check=None):
    cl, cr = sound(L, gamma), sound(R, gamma)
    unl, unr = L[1 + axis], R[1 + axis]
    s_l = B.min(unl - cl, unr - cr)
    s_r = B.max(unl + cl, unr + cr)
    den = L[0] * (s_l - unl) - R[0] * (s_r - unr)
    s_m = (R[3] - L[3] + L[0] * unl * (s_l - unl) - R[0] * unr * (s_r - unr)) / den
    if check is not None:
        check.append(((den.abs() <= 1e-12) | ((s_l - s_m).abs() <= 1e-12) |
                      ((s_r - s_m).abs() <= 1e-12)) & (s_l < 0.0) & (s_r > 0.0))
    psl = L[3] + L[0] * (s_l - unl) * (s_m - unl)
    psr = R[3] + R[0] * (s_r - unr) * (s_m - unr)
    ps = 0.5 * (psl + psr)

    def star_flux(K, sk):
        unk = K[1 + axis]
        uk = cons(K, gamma)
        kd = sk - s_m
        rho_s = K[0] * (sk - unk) / kd
        e_s = ((sk - unk) * uk[3] - K[3] * unk + ps * s_m) / kd
        tang = K[2] if axis == 0 else K[1]
        smx = rho_s * s_m if axis == 0 else rho_s * tang
        smy = rho_s * tang if axis == 0 else rho_s * s_m
        us = torch.stack([rho_s, smx, smy, e_s])
        return flux_phys(K, axis, gamma) + sk * (us - uk)

    star = B.where(s_m >= 0.0, star_flux(L, s_l), star_flux(R, s_r))
    return _supersonic(L, R, axis, gamma, s_l, s_r, star, B)


def xie_hllcp(L, R, axis, gamma, B: Branches = PLAIN, sensor=None, check=None):
    """Xie AM-HLLC-P: the HLLC star flux from the alpha form of s* and p*,
    the low-Mach pressure p*** = f p** + (1 - f) p* in the normal momentum,
    p** = theta p* + (1 - theta)(p_l + p_r)/2, and the pressure-difference
    dissipation beta_p (1, u^, v^, |u^|^2/2) with Roe averages."""
    rho_l, rho_r = L[0], R[0]
    q_l, q_r = L[1 + axis], R[1 + axis]
    p_l, p_r = L[3], R[3]
    c_l, c_r = sound(L, gamma), sound(R, gamma)
    s_l = B.min(q_l - c_l, q_r - c_r)
    s_r = B.max(q_l + c_l, q_r + c_r)
    alpha_l = rho_l * (s_l - q_l)
    alpha_r = rho_r * (s_r - q_r)
    alpha_denom = alpha_r - alpha_l
    s_star = (alpha_r * q_r - alpha_l * q_l + p_l - p_r) / alpha_denom
    rld, rrd = s_l - s_star, s_r - s_star
    ul, ur = cons(L, gamma), cons(R, gamma)
    e_l, e_r = ul[3] / rho_l, ur[3] / rho_r
    e_star_l = e_l + (s_star - q_l) * (s_star + p_l / alpha_l)
    e_star_r = e_r + (s_star - q_r) * (s_star + p_r / alpha_r)
    p_star = (alpha_r * p_l - alpha_l * p_r - alpha_l * alpha_r * (q_l - q_r)) / alpha_denom
    w_l, w_r = torch.sqrt(rho_l), torch.sqrt(rho_r)
    roe = w_l + w_r
    u_hat = (w_l * L[1] + w_r * R[1]) / roe
    v_hat = (w_l * L[2] + w_r * R[2]) / roe
    q_hat = u_hat if axis == 0 else v_hat
    h_l, h_r = (ul[3] + p_l) / rho_l, (ur[3] + p_r) / rho_r
    h_hat = (w_l * h_l + w_r * h_r) / roe
    c_hat_sq = (gamma - 1.0) * (h_hat - 0.5 * (u_hat * u_hat + v_hat * v_hat))
    c_hat = torch.sqrt(c_hat_sq)
    m_hat = q_hat / c_hat                 # max(c^, 1e-12) is c^ for the test states
    sp_l = torch.sqrt(L[1] * L[1] + L[2] * L[2])
    sp_r = torch.sqrt(R[1] * R[1] + R[2] * R[2])
    theta = B.clamp01(B.max(sp_l / c_l, sp_r / c_r))
    f = sensor
    pss = theta * p_star + (1.0 - theta) * 0.5 * (p_l + p_r)
    psss = f * pss + (1.0 - f) * p_star
    beta = (f - 1.0) * s_l * s_r / (s_r - s_l) * (p_r - p_l) / (c_hat_sq * (1.0 + B.abs(m_hat)))
    if check is not None:
        small = lambda x: x.abs() <= 1e-12
        check.append((small(alpha_denom) | small(rld) | small(rrd) | small(alpha_l) | small(alpha_r) |
                      ~(c_hat_sq > 0.0) | (c_hat <= 1e-12) | (c_l <= 1e-12) | (c_r <= 1e-12)) &
                     (s_l < 0.0) & (s_r > 0.0))
    nx_ = 1.0 if axis == 0 else 0.0
    ny_ = 1.0 - nx_

    def star(K, qn, rs, es):
        return torch.stack([rs * s_star,
                            rs * s_star * (K[1] + nx_ * (s_star - qn)) + nx_ * psss,
                            rs * s_star * (K[2] + ny_ * (s_star - qn)) + ny_ * psss,
                            s_star * (rs * es + p_star)])

    phi = torch.stack([beta, beta * u_hat, beta * v_hat, 0.5 * beta * (u_hat * u_hat + v_hat * v_hat)])
    st = B.where(s_star >= 0.0, star(L, q_l, alpha_l / rld, e_star_l),
                 star(R, q_r, alpha_r / rrd, e_star_r)) + phi
    return _supersonic(L, R, axis, gamma, s_l, s_r, st, B)


def _ratio(pl, pr, B):
    return B.min(pl / pr, pr / pl)       # the floor 1e-14 and clamp01 are inactive here


FLUX = {"rusanov": rusanov, "hllc": hllc, "xie_am_hllc_p": xie_hllcp}


def step(U, dt, dx, dy, gamma=1.4, riemann="hllc", periodic=(True, True), bcs=None, fixed=(None,) * 4,
         B: Branches = PLAIN, check=None):
    """U^{n+1} = U^n - dt div F(U^n) (numerics/explicit_update.hpp:58-102).
    bcs (xlo, xhi, ylo, yhi) overrides `periodic`.  check: a list that
    receives, per flux call, the mask of faces that would take a fallback."""
    if bcs is None:
        bcs = tuple("periodic" if periodic[k // 2] else "outflow" for k in range(4))
    Q = to_prim(pad(U, bcs, fixed), gamma)                              # (4, ..., ny+4, nx+4)
    sx = minmod(Q[..., :, 1:-1] - Q[..., :, :-2], Q[..., :, 2:] - Q[..., :, 1:-1], B)   # cols 1..nx+2
    sy = minmod(Q[..., 1:-1, :] - Q[..., :-2, :], Q[..., 2:, :] - Q[..., 1:-1, :], B)   # rows 1..ny+2
    flux = FLUX[riemann]
    # x-faces between padded columns c and c+1 (c = 1..nx+1) on rows -1..ny
    # (the y-face sensor needs the x-faces of the rows above and below)
    qx = Q[..., 1:-1, 1:-1]
    sxx = sx[..., 1:-1, :]
    Lx = (qx + 0.5 * sxx)[..., :, :-1]
    Rx = (qx - 0.5 * sxx)[..., :, 1:]                                     # (4, ..., ny+2, nx+1)
    qy = Q[..., 1:-1, 1:-1]
    syy = sy[..., :, 1:-1]
    Ly = (qy + 0.5 * syy)[..., :-1, :]
    Ry = (qy - 0.5 * syy)[..., 1:, :]                                     # (4, ..., ny+1, nx+2)
    fx_sensor = fy_sensor = None
    if riemann == "xie_am_hllc_p":
        rx = _ratio(Lx[3], Rx[3], B)       # x-faces i-1/2, i = 0..nx, rows -1..ny
        ry = _ratio(Ly[3], Ry[3], B)       # y-faces j-1/2, j = 0..ny, columns -1..nx
        # x-face (i, j): itself, and the y-faces of cells (i-1, j) and (i, j)
        m = rx[..., 1:-1, :]
        for dj in (0, 1):
            for di in (0, 1):
                m = B.min(m, ry[..., dj:dj + ny_of(U), di:di + nx_of(U) + 1])
        fx_sensor = m * m * m
        # y-face (i, j): itself, and the x-faces of cells (i, j-1) and (i, j)
        m = ry[..., :, 1:-1]
        for dj in (0, 1):
            for di in (0, 1):
                m = B.min(m, rx[..., dj:dj + ny_of(U) + 1, di:di + nx_of(U)])
        fy_sensor = m * m * m
    Lx, Rx = Lx[..., 1:-1, :], Rx[..., 1:-1, :]
    Ly, Ry = Ly[..., :, 1:-1], Ry[..., :, 1:-1]
    kw = {} if riemann == "rusanov" else {"check": check}
    Fx = flux(Lx, Rx, 0, gamma, B, fx_sensor, **kw)                      # (4, ..., ny, nx+1)
    Fy = flux(Ly, Ry, 1, gamma, B, fy_sensor, **kw)                      # (4, ..., ny+1, nx)
    return U - dt * ((Fx[..., :, 1:] - Fx[..., :, :-1]) * (1.0 / dx) +
                     (Fy[..., 1:, :] - Fy[..., :-1, :]) * (1.0 / dy))


def nx_of(U):
    return U.shape[-1]


def ny_of(U):
    return U.shape[-2]
