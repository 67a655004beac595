"""Exact solution of the Riemann problem for the 1D Euler equations of an
ideal gas (Toro, Riemann Solvers and Numerical Methods for Fluid Dynamics,
ch. 4): Newton iteration on the pressure function for p*, then sampling of
the self-similar solution.  Used to check the schemes against the closed
form, independently of how they are written."""
from __future__ import annotations

import math

import numpy as np


def _fK(p, rho, pK, cK, g):
    if p > pK:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pK
        f = (p - pK) * math.sqrt(A / (p + B))
        df = math.sqrt(A / (B + p)) * (1.0 - (p - pK) / (2.0 * (B + p)))
    else:  # rarefaction
        f = 2.0 * cK / (g - 1.0) * ((p / pK) ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rho * cK) * (p / pK) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rL, uL, pL, rR, uR, pR, g):
    cL = math.sqrt(g * pL / rL)
    cR = math.sqrt(g * pR / rR)
    p = max(1e-8, 0.5 * (pL + pR) - 0.125 * (uR - uL) * (rL + rR) * (cL + cR))
    for _ in range(100):
        fL, dL = _fK(p, rL, pL, cL, g)
        fR, dR = _fK(p, rR, pR, cR, g)
        pn = p - (fL + fR + uR - uL) / (dL + dR)
        pn = max(pn, 1e-10)
        if abs(pn - p) / (0.5 * (pn + p)) < 1e-14:
            p = pn
            break
        p = pn
    fL, _ = _fK(p, rL, pL, cL, g)
    fR, _ = _fK(p, rR, pR, cR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(rL, uL, pL, rR, uR, pR, g, xi):
    """rho, u, p at similarity coordinates xi = (x - x0)/t (array)."""
    ps, us = star_state(rL, uL, pL, rR, uR, pR, g)
    cL = math.sqrt(g * pL / rL)
    cR = math.sqrt(g * pR / rR)
    out = np.zeros((len(xi), 3))
    gm = (g - 1.0) / (g + 1.0)
    for n, s in enumerate(xi):
        if s <= us:  # left of contact
            if ps > pL:
                SL = uL - cL * math.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                if s <= SL:
                    out[n] = (rL, uL, pL)
                else:
                    out[n] = (rL * (ps / pL + gm) / (gm * ps / pL + 1.0), us, ps)
            else:
                SHL = uL - cL
                cs = cL * (ps / pL) ** ((g - 1) / (2 * g))
                STL = us - cs
                if s <= SHL:
                    out[n] = (rL, uL, pL)
                elif s >= STL:
                    out[n] = (rL * (ps / pL) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1) * (cL + (g - 1) / 2 * (uL - s))
                    u = 2.0 / (g + 1) * (cL + (g - 1) / 2 * uL + s)
                    r = rL * (c / cL) ** (2.0 / (g - 1))
                    out[n] = (r, u, pL * (c / cL) ** (2 * g / (g - 1)))
        else:
            if ps > pR:
                SR = uR + cR * math.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                if s >= SR:
                    out[n] = (rR, uR, pR)
                else:
                    out[n] = (rR * (ps / pR + gm) / (gm * ps / pR + 1.0), us, ps)
            else:
                SHR = uR + cR
                cs = cR * (ps / pR) ** ((g - 1) / (2 * g))
                STR = us + cs
                if s >= SHR:
                    out[n] = (rR, uR, pR)
                elif s <= STR:
                    out[n] = (rR * (ps / pR) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1) * (cR - (g - 1) / 2 * (uR - s))
                    u = 2.0 / (g + 1) * (-cR + (g - 1) / 2 * uR + s)
                    r = rR * (c / cR) ** (2.0 / (g - 1))
                    out[n] = (r, u, pR * (c / cR) ** (2 * g / (g - 1)))
    return out
