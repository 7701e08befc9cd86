"""Exact solution of the Riemann problem for two ideal gases with different
ratios of specific heats (Toro, Riemann Solvers and Numerical Methods for
Fluid Dynamics, 3rd ed., ch. 4: pressure function f_K, Newton iteration for
p*, sampling along x/t).  Each wave family uses the gamma of the gas it
travels through; the contact separates the two gases."""
import math

import numpy as np


def _f(p, rho, pK, gam):
    """f_K(p) and its derivative (Toro eqs. 4.6-4.7, 4.37)."""
    c = math.sqrt(gam * pK / rho)
    if p > pK:  # shock
        A = 2.0 / ((gam + 1.0) * rho)
        B = (gam - 1.0) / (gam + 1.0) * pK
        s = math.sqrt(A / (p + B))
        return (p - pK) * s, s * (1.0 - 0.5 * (p - pK) / (p + B))
    # rarefaction
    e = (gam - 1.0) / (2.0 * gam)
    return (2.0 * c / (gam - 1.0)) * ((p / pK) ** e - 1.0), (p / pK) ** (-(gam + 1.0) / (2.0 * gam)) / (rho * c)


def star_state(L, R):
    """p*, u* for primitive states L = (rho, u, p, gamma), R likewise."""
    rL, uL, pL, gL = L
    rR, uR, pR, gR = R
    p = max(1e-12, 0.5 * (pL + pR))
    for _ in range(100):
        fL, dL = _f(p, rL, pL, gL)
        fR, dR = _f(p, rR, pR, gR)
        dp = (fL + fR + (uR - uL)) / (dL + dR)
        pn = max(1e-14, p - dp)
        if abs(pn - p) <= 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fL, _ = _f(p, rL, pL, gL)
    fR, _ = _f(p, rR, pR, gR)
    return p, 0.5 * (uL + uR) + 0.5 * (fR - fL)


def sample(L, R, x, t, x0):
    """rho, u, p and the left-gas indicator at positions x and time t."""
    rL, uL, pL, gL = L
    rR, uR, pR, gR = R
    ps, us = star_state(L, R)
    s = (np.asarray(x, dtype=float) - x0) / t
    rho, u, p, left = (np.empty_like(s) for _ in range(4))
    for i, si in enumerate(s):
        if si <= us:  # left of the contact
            r0, u0, p0, g = rL, uL, pL, gL
            c0 = math.sqrt(g * p0 / r0)
            if ps > p0:  # left shock
                rs = r0 * ((ps / p0 + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / p0 + 1))
                S = u0 - c0 * math.sqrt((g + 1) / (2 * g) * ps / p0 + (g - 1) / (2 * g))
                val = (r0, u0, p0) if si <= S else (rs, us, ps)
            else:  # left rarefaction
                rs = r0 * (ps / p0) ** (1 / g)
                cs = c0 * (ps / p0) ** ((g - 1) / (2 * g))
                if si <= u0 - c0:
                    val = (r0, u0, p0)
                elif si >= us - cs:
                    val = (rs, us, ps)
                else:
                    uf = 2 / (g + 1) * (c0 + (g - 1) / 2 * u0 + si)
                    cf = 2 / (g + 1) * (c0 + (g - 1) / 2 * (u0 - si))
                    val = (r0 * (cf / c0) ** (2 / (g - 1)), uf, p0 * (cf / c0) ** (2 * g / (g - 1)))
            left[i] = 1.0
        else:  # right of the contact
            r0, u0, p0, g = rR, uR, pR, gR
            c0 = math.sqrt(g * p0 / r0)
            if ps > p0:  # right shock
                rs = r0 * ((ps / p0 + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / p0 + 1))
                S = u0 + c0 * math.sqrt((g + 1) / (2 * g) * ps / p0 + (g - 1) / (2 * g))
                val = (r0, u0, p0) if si >= S else (rs, us, ps)
            else:  # right rarefaction
                rs = r0 * (ps / p0) ** (1 / g)
                cs = c0 * (ps / p0) ** ((g - 1) / (2 * g))
                if si >= u0 + c0:
                    val = (r0, u0, p0)
                elif si <= us + cs:
                    val = (rs, us, ps)
                else:
                    uf = 2 / (g + 1) * (-c0 + (g - 1) / 2 * u0 + si)
                    cf = 2 / (g + 1) * (c0 - (g - 1) / 2 * (u0 - si))
                    val = (r0 * (cf / c0) ** (2 / (g - 1)), uf, p0 * (cf / c0) ** (2 * g / (g - 1)))
            left[i] = 0.0
        rho[i], u[i], p[i] = val
    return rho, u, p, left
