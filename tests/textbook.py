"""Independent textbook formulas used to pin the oracle (numpy / pure Python).

E. F. Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd
ed., Springer 2009:
  - exact Riemann solver for the Euler equations, chapter 4 (pressure function
    f_K of (4.6)-(4.7), sampling of section 4.5);
  - HLL flux, section 10.3 (eqs. 10.20, 10.21);
  - HLLC flux, section 10.4 (eqs. 10.37-10.39, S* of 10.37);
  - primitive-variable Euler equations in quasi-linear form (section 3.2.3),
    W_t + A(W) W_x + B(W) W_y = 0 with W = (rho, p, u, v) in the oracle's
    variable order (rho, p, u, v) = (ID, IP, IU, IV).
Nothing here is derived from the reference's code.
"""
from __future__ import annotations

import math

import numpy as np


def cons(gamma, w):
    """(rho, p, u, v) -> (rho, E, rho u, rho v)."""
    r, p, u, v = w
    return np.array([r, p / (gamma - 1) + 0.5 * r * (u * u + v * v), r * u, r * v])


def xflux(gamma, w):
    """Physical x-flux in the (rho, E, mx, my) order."""
    r, p, u, v = w
    E = p / (gamma - 1) + 0.5 * r * (u * u + v * v)
    return np.array([r * u, u * (E + p), r * u * u + p, r * u * v])


def wave_speeds(gamma, wl, wr):
    """Davis-type estimates SL = min(uL,uR) - max(cL,cR), SR = max(uL,uR) + max(cL,cR)."""
    cl = math.sqrt(gamma * wl[1] / wl[0])
    cr = math.sqrt(gamma * wr[1] / wr[0])
    cm = max(cl, cr)
    return min(wl[2], wr[2]) - cm, max(wl[2], wr[2]) + cm


def hll_flux(gamma, wl, wr, form="10.21"):
    SL, SR = wave_speeds(gamma, wl, wr)
    UL, UR = cons(gamma, wl), cons(gamma, wr)
    FL, FR = xflux(gamma, wl), xflux(gamma, wr)
    if SL >= 0:
        return FL
    if SR <= 0:
        return FR
    if form == "10.21":
        return (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL)
    # eq. (10.20): HLL state, then F = FL + SL (U_hll - UL)
    Uhll = (SR * UR - SL * UL - (FR - FL)) / (SR - SL)
    return FL + SL * (Uhll - UL)


def hllc_flux(gamma, wl, wr):
    """Toro (10.37)-(10.39) with the same wave-speed estimates."""
    SL, SR = wave_speeds(gamma, wl, wr)
    rl, pl, ul, vl = wl
    rr, pr, ur, vr = wr
    Ss = (pr - pl + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))

    def star(w, S):
        r, p, u, v = w
        E = cons(gamma, w)[1]
        fac = r * (S - u) / (S - Ss)
        return fac * np.array([1.0, E / r + (Ss - u) * (Ss + p / (r * (S - u))), Ss, v])

    if SL >= 0:
        return xflux(gamma, wl)
    if Ss >= 0:
        return xflux(gamma, wl) + SL * (star(wl, SL) - cons(gamma, wl))
    if SR >= 0:
        return xflux(gamma, wr) + SR * (star(wr, SR) - cons(gamma, wr))
    return xflux(gamma, wr)


# ------------------------------------------------------- exact Riemann solver
def _fK(gamma, p, rK, pK):
    cK = math.sqrt(gamma * pK / rK)
    if p > pK:  # shock
        A = 2.0 / ((gamma + 1) * rK)
        B = (gamma - 1) / (gamma + 1) * pK
        f = (p - pK) * math.sqrt(A / (p + B))
        df = math.sqrt(A / (B + p)) * (1 - (p - pK) / (2 * (B + p)))
    else:  # rarefaction
        f = 2 * cK / (gamma - 1) * ((p / pK) ** ((gamma - 1) / (2 * gamma)) - 1)
        df = 1.0 / (rK * cK) * (p / pK) ** (-(gamma + 1) / (2 * gamma))
    return f, df


def exact_star(gamma, wl, wr, tol=1e-15):
    """(p*, u*) of the exact Riemann problem (Toro 4.3, Newton iteration)."""
    rl, pl, ul, _ = wl
    rr, pr, ur, _ = wr
    p = max(0.5 * (pl + pr), 1e-12)
    for _ in range(200):
        fl, dfl = _fK(gamma, p, rl, pl)
        fr, dfr = _fK(gamma, p, rr, pr)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-14)
        if abs(pn - p) <= tol * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fK(gamma, p, rl, pl)
    fr, _ = _fK(gamma, p, rr, pr)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def exact_sample(gamma, wl, wr, s=0.0):
    """(rho, p, u) of the exact solution at x/t = s (Toro 4.5)."""
    ps, us = exact_star(gamma, wl, wr)
    g = gamma
    if s <= us:
        r, p, u = wl[0], wl[1], wl[2]
        c = math.sqrt(g * p / r)
        if ps > p:
            S = u - c * math.sqrt((g + 1) / (2 * g) * ps / p + (g - 1) / (2 * g))
            if s <= S:
                return r, p, u
            rs = r * ((ps / p + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / p + 1))
            return rs, ps, us
        cs = c * (ps / p) ** ((g - 1) / (2 * g))
        if s <= u - c:
            return r, p, u
        if s >= us - cs:
            return r * (ps / p) ** (1 / g), ps, us
        uf = 2 / (g + 1) * (c + (g - 1) / 2 * u + s)
        cf = 2 / (g + 1) * (c + (g - 1) / 2 * (u - s))
        return r * (cf / c) ** (2 / (g - 1)), p * (cf / c) ** (2 * g / (g - 1)), uf
    r, p, u = wr[0], wr[1], wr[2]
    c = math.sqrt(g * p / r)
    if ps > p:
        S = u + c * math.sqrt((g + 1) / (2 * g) * ps / p + (g - 1) / (2 * g))
        if s >= S:
            return r, p, u
        rs = r * ((ps / p + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / p + 1))
        return rs, ps, us
    cs = c * (ps / p) ** ((g - 1) / (2 * g))
    if s >= u + c:
        return r, p, u
    if s <= us + cs:
        return r * (ps / p) ** (1 / g), ps, us
    uf = 2 / (g + 1) * (-c + (g - 1) / 2 * u + s)
    cf = 2 / (g + 1) * (c - (g - 1) / 2 * (u - s))
    return r * (cf / c) ** (2 / (g - 1)), p * (cf / c) ** (2 * g / (g - 1)), uf


# ------------------------------------------------ primitive quasi-linear form
def jacobians(gamma, w):
    """A(W), B(W) of W_t + A W_x + B W_y = 0 for W = (rho, p, u, v)."""
    r, p, u, v = w
    A = np.array([[u, 0, r, 0],
                  [0, u, gamma * p, 0],
                  [0, 1 / r, u, 0],
                  [0, 0, 0, u]])
    B = np.array([[v, 0, 0, r],
                  [0, v, 0, gamma * p],
                  [0, 0, v, 0],
                  [0, 1 / r, 0, v]])
    return A, B


def hancock_face(gamma, w, dwx, dwy, dtdx, dtdy, face):
    """MUSCL-Hancock evolved boundary value (Toro 14.4.2, unsplit 2-D form):
    W_face = W -/+ dW_n/2 + (dt/(2 dx_n)) * (-A dWx dx_n/dx - B dWy dx_n/dy)
    where the slopes dW are per-cell differences, so dW/dx = dWx/dx."""
    A, B = jacobians(gamma, w)
    dwdt_dx = -(A @ dwx) - (B @ dwy)  # = dW/dt * dx assuming dx == dy scaling below
    if face in (0, 1):
        sgn = -1 if face == 0 else 1
        return np.asarray(w) + sgn * 0.5 * np.asarray(dwx) + 0.5 * dtdx * dwdt_dx
    sgn = -1 if face == 2 else 1
    return np.asarray(w) + sgn * 0.5 * np.asarray(dwy) + 0.5 * dtdy * dwdt_dx
