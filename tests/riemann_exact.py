"""Exact solution of the 1-D Riemann problem for an ideal gas (textbook: E. F.
Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4:
pressure function f_K, Newton iteration for p*, then wave sampling). Written
independently of the scheme under test; used only as the textbook check of the
Godunov/HLLC oracle."""
from __future__ import annotations

import numpy as np


def _f_and_df(p, rho, pk, ck, g):
    if p > pk:  # shock
        a = 2.0 / ((g + 1.0) * rho)
        b = (g - 1.0) / (g + 1.0) * pk
        q = np.sqrt(a / (p + b))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (b + p))
    # rarefaction
    pr = p / pk
    f = 2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
    df = 1.0 / (rho * ck) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    if 2.0 * (cl + cr) / (g - 1.0) <= ur - ul:
        raise ValueError("vacuum is generated")
    p = max(1e-8, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))  # PVRS guess
    for _ in range(100):
        fl, dfl = _f_and_df(p, rl, pl, cl, g)
        fr, dfr = _f_and_df(p, rr, pr, cr, g)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-10)
        if abs(pn - p) <= 1e-14 * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f_and_df(p, rl, pl, cl, g)
    fr, _ = _f_and_df(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x_over_t, rl, ul, pl, rr, ur, pr, g):
    """(rho, u, p) at similarity coordinate x/t for each entry of x_over_t."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    out = np.empty((3, len(x_over_t)))
    gm = (g - 1.0) / (g + 1.0)
    for n, s in enumerate(x_over_t):
        if s <= us:  # left of contact
            if ps > pl:  # left shock
                sl = ul - cl * np.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
                if s <= sl:
                    st = (rl, ul, pl)
                else:
                    st = (rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps)
            else:  # left rarefaction
                shl = ul - cl
                cml = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
                stl = us - cml
                if s <= shl:
                    st = (rl, ul, pl)
                elif s >= stl:
                    st = (rl * (ps / pl) ** (1.0 / g), us, ps)
                else:
                    u = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * ul + s)
                    c = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * (ul - s))
                    rho = rl * (c / cl) ** (2.0 / (g - 1.0))
                    st = (rho, u, pl * (c / cl) ** (2.0 * g / (g - 1.0)))
        else:  # right of contact
            if ps > pr:  # right shock
                sr = ur + cr * np.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
                if s >= sr:
                    st = (rr, ur, pr)
                else:
                    st = (rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps)
            else:  # right rarefaction
                shr = ur + cr
                cmr = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
                str_ = us + cmr
                if s >= shr:
                    st = (rr, ur, pr)
                elif s <= str_:
                    st = (rr * (ps / pr) ** (1.0 / g), us, ps)
                else:
                    u = 2.0 / (g + 1.0) * (-cr + (g - 1.0) / 2.0 * ur + s)
                    c = 2.0 / (g + 1.0) * (cr - (g - 1.0) / 2.0 * (ur - s))
                    rho = rr * (c / cr) ** (2.0 / (g - 1.0))
                    st = (rho, u, pr * (c / cr) ** (2.0 * g / (g - 1.0)))
        out[:, n] = st
    return out
