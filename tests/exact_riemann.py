"""Exact solution of the Riemann problem of the 1-D Euler equations (ideal gas).

Follows the published algorithm of E. F. Toro, "Riemann Solvers and Numerical Methods
for Fluid Dynamics", 3rd ed., chapter 4: pressure function f_K (eq. 4.6-4.8), Newton-
Raphson for p* (eq. 4.44-4.46), sampling of the self-similar solution (section 4.5).
The reference's own check of the Sod test uses the same textbook solution
(test/HD/sod/python/testidefix.py:38-77 through pytools/sod.py).
"""
import numpy as np


def _fK(p, rho, pk, gamma):
    a = np.sqrt(gamma * pk / rho)
    if p > pk:  # shock
        A = 2.0 / ((gamma + 1.0) * rho)
        B = (gamma - 1.0) / (gamma + 1.0) * pk
        f = (p - pk) * np.sqrt(A / (p + B))
        df = np.sqrt(A / (B + p)) * (1.0 - (p - pk) / (2.0 * (B + p)))
    else:  # rarefaction
        f = 2.0 * a / (gamma - 1.0) * ((p / pk) ** ((gamma - 1.0) / (2.0 * gamma)) - 1.0)
        df = 1.0 / (rho * a) * (p / pk) ** (-(gamma + 1.0) / (2.0 * gamma))
    return f, df


def star_state(rL, uL, pL, rR, uR, pR, gamma=1.4, tol=1e-14):
    """(p*, u*) of the exact solution."""
    aL = np.sqrt(gamma * pL / rL)
    aR = np.sqrt(gamma * pR / rR)
    du = uR - uL
    # two-rarefaction guess (Toro eq. 4.46), floored
    z = (gamma - 1.0) / (2.0 * gamma)
    p = ((aL + aR - 0.5 * (gamma - 1.0) * du) /
         (aL / pL ** z + aR / pR ** z)) ** (1.0 / z)
    p = max(p, 1e-8)
    for _ in range(200):
        fL, dfL = _fK(p, rL, pL, gamma)
        fR, dfR = _fK(p, rR, pR, gamma)
        pn = p - (fL + fR + du) / (dfL + dfR)
        pn = max(pn, 1e-12)
        if abs(pn - p) / (0.5 * (pn + p)) < tol:
            p = pn
            break
        p = pn
    fL, _ = _fK(p, rL, pL, gamma)
    fR, _ = _fK(p, rR, pR, gamma)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(x, t, x0, left, right, gamma=1.4):
    """rho, u, p at positions x and time t for the states left=(rho,u,p), right=(...)."""
    rL, uL, pL = left
    rR, uR, pR = right
    ps, us = star_state(rL, uL, pL, rR, uR, pR, gamma)
    aL = np.sqrt(gamma * pL / rL)
    aR = np.sqrt(gamma * pR / rR)
    g1 = (gamma - 1.0) / (gamma + 1.0)
    out = np.zeros((3, len(x)))
    for n, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if ps > pL:  # left shock
                SL = uL - aL * np.sqrt((gamma + 1) / (2 * gamma) * ps / pL + (gamma - 1) / (2 * gamma))
                if s <= SL:
                    st = (rL, uL, pL)
                else:
                    st = (rL * (ps / pL + g1) / (g1 * ps / pL + 1.0), us, ps)
            else:  # left rarefaction
                aLs = aL * (ps / pL) ** ((gamma - 1) / (2 * gamma))
                if s <= uL - aL:
                    st = (rL, uL, pL)
                elif s >= us - aLs:
                    st = (rL * (ps / pL) ** (1.0 / gamma), us, ps)
                else:
                    c = 2.0 / (gamma + 1) + (gamma - 1) / ((gamma + 1) * aL) * (uL - s)
                    st = (rL * c ** (2.0 / (gamma - 1)),
                          2.0 / (gamma + 1) * (aL + (gamma - 1) / 2.0 * uL + s),
                          pL * c ** (2.0 * gamma / (gamma - 1)))
        else:  # right of contact
            if ps > pR:  # right shock
                SR = uR + aR * np.sqrt((gamma + 1) / (2 * gamma) * ps / pR + (gamma - 1) / (2 * gamma))
                if s >= SR:
                    st = (rR, uR, pR)
                else:
                    st = (rR * (ps / pR + g1) / (g1 * ps / pR + 1.0), us, ps)
            else:  # right rarefaction
                aRs = aR * (ps / pR) ** ((gamma - 1) / (2 * gamma))
                if s >= uR + aR:
                    st = (rR, uR, pR)
                elif s <= us + aRs:
                    st = (rR * (ps / pR) ** (1.0 / gamma), us, ps)
                else:
                    c = 2.0 / (gamma + 1) - (gamma - 1) / ((gamma + 1) * aR) * (uR - s)
                    st = (rR * c ** (2.0 / (gamma - 1)),
                          2.0 / (gamma + 1) * (-aR + (gamma - 1) / 2.0 * uR + s),
                          pR * c ** (2.0 * gamma / (gamma - 1)))
        out[:, n] = st
    return out
