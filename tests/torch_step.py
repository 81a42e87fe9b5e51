"""A differentiable FP64 torch version of the explicit O2 step, for a
reference gradient: torch.autograd differentiates it independently of the
hand-written adjoint (csrc/ec_adjoint.hpp).  Covers the periodic and outflow
ghost rules, the minmod MUSCL reconstruction in primitive variables and the
Rusanov and HLLC fluxes with the branch structure of
numerics/reconstruction_fluxes.hpp:66-210 (main branches; the degenerate
fallbacks are not reached by the smooth test states).  Test infrastructure
only."""
from __future__ import annotations

import torch


def _ghost_index(n: int, periodic: bool) -> torch.Tensor:
    k = torch.arange(-2, n + 2)
    return torch.remainder(k, n) if periodic else k.clamp(0, n - 1)


def pad(U: torch.Tensor, periodic_x: bool, periodic_y: bool) -> torch.Tensor:
    """(4, ny, nx) -> (4, ny+4, nx+4) through the ghost rule (a gather, whose
    transpose is the ghost fold)."""
    ny, nx = U.shape[1:]
    iy = _ghost_index(ny, periodic_y)
    ix = _ghost_index(nx, periodic_x)
    return U[:, iy][:, :, ix]


def to_prim(U, gamma):
    rho = U[0]
    inv = 1.0 / rho
    u = U[1] * inv
    v = U[2] * inv
    p = (gamma - 1.0) * (U[3] - 0.5 * rho * (u * u + v * v))
    return torch.stack([rho, u, v, p])


def minmod(b, f):
    return torch.where(b * f <= 0.0, torch.zeros_like(b), torch.where(b.abs() < f.abs(), b, f))


def flux_phys(q, axis, gamma):
    rho, u, v, p = q
    mx, my = rho * u, rho * v
    E = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    un = u if axis == 0 else v
    if axis == 0:
        return torch.stack([mx, mx * u + p, my * u, (E + p) * u])
    return torch.stack([my, mx * v, my * v + p, (E + p) * v])


def cons(q, gamma):
    rho, u, v, p = q
    return torch.stack([rho, rho * u, rho * v, p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)])


def rusanov(L, R, axis, gamma):
    cl = torch.sqrt(gamma * L[3] / L[0])
    cr = torch.sqrt(gamma * R[3] / R[0])
    unl, unr = L[1 + axis], R[1 + axis]
    s = torch.maximum(unl.abs() + cl, unr.abs() + cr)
    return 0.5 * (flux_phys(L, axis, gamma) + flux_phys(R, axis, gamma)) - 0.5 * s * (cons(R, gamma) - cons(L, gamma))


def hllc(L, R, axis, gamma):
    cl = torch.sqrt(gamma * L[3] / L[0])
    cr = torch.sqrt(gamma * R[3] / R[0])
    unl, unr = L[1 + axis], R[1 + axis]
    s_l = torch.minimum(unl - cl, unr - cr)
    s_r = torch.maximum(unl + cl, unr + cr)
    den = L[0] * (s_l - unl) - R[0] * (s_r - unr)
    s_m = (R[3] - L[3] + L[0] * unl * (s_l - unl) - R[0] * unr * (s_r - unr)) / den
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

    star = torch.where(s_m >= 0.0, star_flux(L, s_l), star_flux(R, s_r))
    out = torch.where(s_r <= 0.0, flux_phys(R, axis, gamma), star)
    return torch.where(s_l >= 0.0, flux_phys(L, axis, gamma), out)


def step(U, dt, dx, dy, gamma=1.4, riemann="hllc", periodic=(True, True)):
    """U^{n+1} = U^n - dt div F(U^n) (numerics/explicit_update.hpp:58-102)."""
    ny, nx = U.shape[1:]
    Q = to_prim(pad(U, *periodic), gamma)            # (4, ny+4, nx+4)
    sx = minmod(Q[:, :, 1:-1] - Q[:, :, :-2], Q[:, :, 2:] - Q[:, :, 1:-1])   # cols 1..nx+2
    sy = minmod(Q[:, 1:-1, :] - Q[:, :-2, :], Q[:, 2:, :] - Q[:, 1:-1, :])   # rows 1..ny+2
    flux = rusanov if riemann == "rusanov" else hllc
    # x-faces between padded columns c and c+1 for c = 1..nx+1 (valid rows)
    qx = Q[:, 2:-2, 1:-1]
    sxx = sx[:, 2:-2, :]
    Lx = (qx + 0.5 * sxx)[:, :, :-1]
    Rx = (qx - 0.5 * sxx)[:, :, 1:]
    Fx = flux(Lx, Rx, 0, gamma)                      # (4, ny, nx+1)
    qy = Q[:, 1:-1, 2:-2]
    syy = sy[:, :, 2:-2]
    Ly = (qy + 0.5 * syy)[:, :-1, :]
    Ry = (qy - 0.5 * syy)[:, 1:, :]
    Fy = flux(Ly, Ry, 1, gamma)                      # (4, ny+1, nx)
    return U - dt * ((Fx[:, :, 1:] - Fx[:, :, :-1]) * (1.0 / dx) + (Fy[:, 1:, :] - Fy[:, :-1, :]) * (1.0 / dy))
