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


# ---------------------------------------------------------------------------------------
# Isothermal gas (p = cs^2 rho): the same construction with the isothermal wave curves
# (e.g. LeVeque, "Finite Volume Methods for Hyperbolic Problems", sec. 14.6 and 13.7):
# across a shock into density rho_K, u changes by -cs (rho - rho_K) / sqrt(rho rho_K);
# across a rarefaction by -cs ln(rho / rho_K).
# ---------------------------------------------------------------------------------------
def _fK_iso(rho, rk, cs):
    if rho > rk:
        f = cs * (rho - rk) / np.sqrt(rho * rk)
        df = cs * (0.5 / np.sqrt(rho * rk) + 0.5 * rk / (rho * np.sqrt(rho * rk)))
    else:
        f = cs * np.log(rho / rk)
        df = cs / rho
    return f, df


def star_state_iso(rL, uL, rR, uR, cs=1.0, tol=1e-14):
    """(rho*, u*) of the isothermal Riemann problem."""
    rho = np.sqrt(rL * rR)
    for _ in range(200):
        fL, dfL = _fK_iso(rho, rL, cs)
        fR, dfR = _fK_iso(rho, rR, cs)
        rn = rho - (fL + fR + uR - uL) / (dfL + dfR)
        rn = max(rn, 1e-12)
        if abs(rn - rho) / (0.5 * (rn + rho)) < tol:
            rho = rn
            break
        rho = rn
    fL, _ = _fK_iso(rho, rL, cs)
    fR, _ = _fK_iso(rho, rR, cs)
    return rho, 0.5 * (uL + uR) + 0.5 * (fR - fL)


def sample_iso(x, t, x0, left, right, cs=1.0):
    """(rho, u) at positions x and time t; left/right = (rho, u)."""
    rL, uL = left
    rR, uR = right
    rs, us = star_state_iso(rL, uL, rR, uR, cs)
    out = np.zeros((2, len(x)))
    for n, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:  # left of the contact (no contact in isothermal flow: u* splits)
            if rs > rL:  # left shock
                S = uL - cs * np.sqrt(rs / rL)
                st = (rL, uL) if s < S else (rs, us)
            else:  # left rarefaction: head uL - cs, tail us - cs
                if s < uL - cs:
                    st = (rL, uL)
                elif s > us - cs:
                    st = (rs, us)
                else:
                    u = s + cs
                    st = (rL * np.exp(-(u - uL) / cs), u)
        else:
            if rs > rR:  # right shock
                S = uR + cs * np.sqrt(rs / rR)
                st = (rR, uR) if s > S else (rs, us)
            else:
                if s > uR + cs:
                    st = (rR, uR)
                elif s < us + cs:
                    st = (rs, us)
                else:
                    u = s - cs
                    st = (rR * np.exp((u - uR) / cs), u)
        out[:, n] = st
    return out
