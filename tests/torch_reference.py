"""Plain FP64 torch restatement of the Godunov step (HLLC, perfect gas) used as
the autograd reference for the hand-written adjoint: the same expression trees as
hllc.hpp:53-104 / godunov.hpp:31-90 / cons_to_prim.hpp:28-31 / prim_to_cons.hpp:28-31,
branches as torch.where (gradient flows through the selected branch, as in the
hand-written VJPs), dt from the detached primitives (frozen per step).

Fields are tensors of shape (n2, n1, n0); a state is a list of five of them.
"""
from __future__ import annotations

import torch

AXIS = {0: 2, 1: 1, 2: 0}  # direction -> tensor dimension


def kmin(a, b):
    return torch.where(b < a, b, a)


def kmax(a, b):
    return torch.where(a < b, b, a)


def hllc(qL, qR, d, gamma):
    gm1 = gamma - 1
    dL, pL, uL = qL[0], qL[1], qL[2:]
    dR, pR, uR = qR[0], qR[1], qR[2:]
    unL, unR = uL[d], uR[d]
    cL = torch.sqrt(gamma * pL / dL)
    cR = torch.sqrt(gamma * pR / dR)
    cmax = kmax(cL, cR)
    SL = kmin(unL, unR) - cmax
    SR = kmax(unL, unR) + cmax
    rcL = dL * (SL - unL)
    rcR = dR * (SR - unR)
    us = (pR - pL + rcL * unL - rcR * unR) / (rcL - rcR)
    ps = 0.5 * (pL + pR + rcL * (us - unL) + rcR * (us - unR))
    left = us > 0
    S = torch.where(left, SL, SR)
    qd = torch.where(left, dL, dR)
    qp = torch.where(left, pL, pR)
    qu = [torch.where(left, a, b) for a, b in zip(uL, uR)]
    un = qu[d]
    etot = qp / gm1 + qd * (qu[0] * qu[0] + qu[1] * qu[1] + qu[2] * qu[2]) / 2
    sup = SL * SR > 0
    uo = torch.where(sup, un, us)
    po = torch.where(sup, qp, ps)
    ratio = (S - un) / (S - uo)
    d_o = ratio * qd
    eo = ratio * etot + (po * uo - qp * un) / (S - us)
    mass = d_o * uo
    f = [mass, (eo + po) * uo, mass * qu[0], mass * qu[1], mass * qu[2]]
    f[2 + d] = mass * uo + po
    return f


def prim_to_cons(P, gamma):
    d, p, u = P[0], P[1], P[2:]
    e = d * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) / 2 + p / (gamma - 1)
    return [d, e, d * u[0], d * u[1], d * u[2]]


def cons_to_prim(U, gamma):
    d, e, m = U[0], U[1], U[2:]
    ke = (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / d / 2
    p = (gamma - 1) * (e - ke)
    inv = 1 / d
    return [d, p, m[0] * inv, m[1] * inv, m[2] * inv]


def time_step(P, gamma, dx):
    d, p, u = P[0], P[1], P[2:]
    c = torch.sqrt(gamma * p / d)
    inv = (c + u[0].abs()) * (1 / dx[0]) + (c + u[1].abs()) * (1 / dx[1]) + (c + u[2].abs()) * (1 / dx[2])
    return float((1 / inv).min())


def _slice(t, dim, start, stop):
    idx = [slice(None)] * 3
    idx[dim] = slice(start, stop)
    return t[tuple(idx)]


def godunov(P, U, gamma, dx, dt):
    """cons - dt/dv * sum_d ds_d (F_hi - F_lo) on interior cells (new tensors)."""
    ds = (dx[1] * dx[2], dx[2] * dx[0], dx[0] * dx[1])
    dtodv = dt / (dx[0] * dx[1] * dx[2])
    inner = tuple(slice(1, -1) for _ in range(3))
    flux = None
    for d in range(3):
        ax = AXIS[d]
        n = P[0].shape[ax]
        qL = [_slice(q, ax, 0, n - 1) for q in P]
        qR = [_slice(q, ax, 1, n) for q in P]
        F = hllc(qL, qR, d, gamma)  # faces between consecutive cells along ax
        contrib = []
        for f in F:
            hi = _slice(f, ax, 1, n - 1)
            lo = _slice(f, ax, 0, n - 2)
            diff = ds[d] * (hi - lo)
            # restrict the two other axes to the interior
            idx = [slice(1, -1)] * 3
            idx[ax] = slice(None)
            contrib.append(diff[tuple(idx)])
        flux = contrib if flux is None else [a + b for a, b in zip(flux, contrib)]
    out = []
    for u, f in zip(U, flux):
        v = u.clone()
        v[inner] = u[inner] - dtodv * f
        out.append(v)
    return out


def rollout(P0, gamma, dx, cfl, nsteps):
    U = prim_to_cons(P0, gamma)
    P = list(P0)
    for _ in range(nsteps):
        dt = cfl * time_step([q.detach() for q in P], gamma, dx)
        U = godunov(P, U, gamma, dx, dt)
        P = cons_to_prim(U, gamma)
    return P


def switch_margins(P, gamma, box=None):
    """Smallest distance to each HLLC branch switch over the faces of state P
    between cells of `box` ((lo, hi) per tensor dimension, default all):
    |cL - cR| (max of sound speeds), |unL - unR| (min/max of normal speeds),
    |u*| (upwind side) and |S_L * S_R| (supersonic test), as floats."""
    if box is not None:
        P = [q[tuple(slice(lo, hi) for lo, hi in box)] for q in P]
    out = {"c": float("inf"), "un": float("inf"), "us": float("inf"), "sup": float("inf")}
    for d in range(3):
        ax = AXIS[d]
        n = P[0].shape[ax]
        qL = [_slice(q.detach(), ax, 0, n - 1) for q in P]
        qR = [_slice(q.detach(), ax, 1, n) for q in P]
        cL = torch.sqrt(gamma * qL[1] / qL[0])
        cR = torch.sqrt(gamma * qR[1] / qR[0])
        unL, unR = qL[2 + d], qR[2 + d]
        cmax = kmax(cL, cR)
        SL = kmin(unL, unR) - cmax
        SR = kmax(unL, unR) + cmax
        rcL, rcR = qL[0] * (SL - unL), qR[0] * (SR - unR)
        us = (qR[1] - qL[1] + rcL * unL - rcR * unR) / (rcL - rcR)
        for key, v in (("c", (cL - cR).abs()), ("un", (unL - unR).abs()), ("us", us.abs()),
                       ("sup", (SL * SR).abs())):
            out[key] = min(out[key], float(v.min()))
    return out
