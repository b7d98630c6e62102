"""Independent textbook references used to pin the oracle (not derived from
the reference code): Toro's HLLC and Roe solvers and the exact Riemann
solver for the ideal-gas Euler equations (E. F. Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4, 10, 11), and the
Legendre-product basis evaluated from its definition."""
from __future__ import annotations

import math

import numpy as np


# ---------------------------------------------------------------------------
# Euler flux in direction d (0,1,2) for U = (rho, mx, my, mz, E)
# ---------------------------------------------------------------------------
def euler_flux(U, d, gamma):
    rho = U[0]
    v = U[1:4] / rho
    p = (gamma - 1) * (U[4] - 0.5 * rho * np.dot(v, v))
    F = np.empty(5)
    F[0] = rho * v[d]
    F[1:4] = rho * v[d] * v
    F[1 + d] += p
    F[4] = (U[4] + p) * v[d]
    return F


def hllc_toro(UL, UR, d, gamma):
    """Toro eq. (10.71)-(10.73) with Davis wave-speed estimates (10.48)."""
    def prim(U):
        rho = U[0]; v = U[1:4] / rho
        p = (gamma - 1) * (U[4] - 0.5 * rho * np.dot(v, v))
        return rho, v, p, math.sqrt(gamma * p / rho)
    rL, vL, pL, cL = prim(UL)
    rR, vR, pR, cR = prim(UR)
    uL, uR = vL[d], vR[d]
    SL = min(uL - cL, uR - cR)
    SR = max(uL + cL, uR + cR)
    Ss = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR))
    FL, FR = euler_flux(UL, d, gamma), euler_flux(UR, d, gamma)

    def ustar(U, r, v, p, S):
        u = v[d]
        fac = r * (S - u) / (S - Ss)
        Us = np.empty(5)
        Us[0] = fac
        Us[1:4] = fac * v
        Us[1 + d] = fac * Ss
        Us[4] = fac * (U[4] / r + (Ss - u) * (Ss + p / (r * (S - u))))
        return Us
    if SL >= 0:
        return FL
    if SR <= 0:
        return FR
    if Ss >= 0:
        return FL + SL * (ustar(UL, rL, vL, pL, SL) - UL)
    return FR + SR * (ustar(UR, rR, vR, pR, SR) - UR)


def roe_toro(UL, UR, d, gamma, h_thermal=1.0):
    """Roe flux, Toro sec. 11.2: F = (FL+FR)/2 - 1/2 sum |lambda_i| alpha_i K_i.
    h_thermal scales the thermal part of the averaged enthalpy,
    H = v^2/2 + h_thermal * gamma p / ((gamma-1) rho); the textbook value is 1."""
    def prim(U):
        rho = U[0]; v = U[1:4] / rho
        p = (gamma - 1) * (U[4] - 0.5 * rho * np.dot(v, v))
        H = 0.5 * np.dot(v, v) + h_thermal * gamma / (gamma - 1) * p / rho
        return rho, v, p, H
    rL, vL, pL, HL = prim(UL)
    rR, vR, pR, HR = prim(UR)
    sL, sR = math.sqrt(rL), math.sqrt(rR)
    v = (sL * vL + sR * vR) / (sL + sR)
    H = (sL * HL + sR * HR) / (sL + sR)
    a2 = (gamma - 1) * (H - 0.5 * np.dot(v, v))
    a = math.sqrt(a2)
    t = [i for i in range(3) if i != d]
    u = v[d]
    dU = UR - UL
    # Toro (11.68)-(11.70) generalised to 3D with the normal direction d
    du5 = dU[4] - (dU[1 + t[0]] - v[t[0]] * dU[0]) * v[t[0]] - (dU[1 + t[1]] - v[t[1]] * dU[0]) * v[t[1]]
    al2 = (gamma - 1) / a2 * (dU[0] * (H - u * u) + u * dU[1 + d] - du5)
    al1 = 1 / (2 * a) * (dU[0] * (u + a) - dU[1 + d] - a * al2)
    al5 = dU[0] - (al1 + al2)
    alt0 = dU[1 + t[0]] - v[t[0]] * dU[0]
    alt1 = dU[1 + t[1]] - v[t[1]] * dU[0]
    K1 = np.empty(5); K1[0] = 1; K1[1:4] = v; K1[1 + d] = u - a; K1[4] = H - u * a
    K2 = np.empty(5); K2[0] = 1; K2[1:4] = v; K2[4] = 0.5 * np.dot(v, v)
    Kt0 = np.zeros(5); Kt0[1 + t[0]] = 1; Kt0[4] = v[t[0]]
    Kt1 = np.zeros(5); Kt1[1 + t[1]] = 1; Kt1[4] = v[t[1]]
    K5 = np.empty(5); K5[0] = 1; K5[1:4] = v; K5[1 + d] = u + a; K5[4] = H + u * a
    diss = abs(u - a) * al1 * K1 + abs(u) * (al2 * K2 + alt0 * Kt0 + alt1 * Kt1) + abs(u + a) * al5 * K5
    return 0.5 * (euler_flux(UL, d, gamma) + euler_flux(UR, d, gamma)) - 0.5 * diss


# ---------------------------------------------------------------------------
# exact Riemann solver (Toro ch. 4), sampled at x/t
# ---------------------------------------------------------------------------
def exact_riemann(rL, uL, pL, rR, uR, pR, gamma, xi):
    g = gamma
    cL, cR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)

    def fK(p, rK, pK, cK):
        if p > pK:
            A, B = 2 / ((g + 1) * rK), (g - 1) / (g + 1) * pK
            return (p - pK) * math.sqrt(A / (p + B))
        return 2 * cK / (g - 1) * ((p / pK) ** ((g - 1) / (2 * g)) - 1)

    def f(p):
        return fK(p, rL, pL, cL) + fK(p, rR, pR, cR) + uR - uL
    lo, hi = 1e-12, 10 * max(pL, pR)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (uL + uR) + 0.5 * (fK(ps, rR, pR, cR) - fK(ps, rL, pL, cL))
    out = []
    for s in np.atleast_1d(xi):
        if s <= us:
            if ps > pL:
                SL = uL - cL * math.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                r = rL * ((ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1)) if s >= SL else rL
            else:
                csL = cL * (ps / pL) ** ((g - 1) / (2 * g))
                if s <= uL - cL:
                    r = rL
                elif s >= us - csL:
                    r = rL * (ps / pL) ** (1 / g)
                else:
                    r = rL * (2 / (g + 1) + (g - 1) / ((g + 1) * cL) * (uL - s)) ** (2 / (g - 1))
        else:
            if ps > pR:
                SR = uR + cR * math.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                r = rR * ((ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1)) if s <= SR else rR
            else:
                csR = cR * (ps / pR) ** ((g - 1) / (2 * g))
                if s >= uR + cR:
                    r = rR
                elif s <= us + csR:
                    r = rR * (ps / pR) ** (1 / g)
                else:
                    r = rR * (2 / (g + 1) - (g - 1) / ((g + 1) * cR) * (uR - s)) ** (2 / (g - 1))
        out.append(r)
    return np.array(out), ps, us


# ---------------------------------------------------------------------------
# basis from its definition
# ---------------------------------------------------------------------------
def basis_functions(nb):
    P0 = lambda s: 1.0
    P1 = lambda s: s
    P2 = lambda s: 1.5 * s * s - 0.5
    fs = [(P0, P0, P0), (P1, P0, P0), (P0, P1, P0), (P0, P0, P1), (P0, P1, P1), (P1, P0, P1), (P1, P1, P0)]
    if nb == 4:
        fs = fs[:4]
    if nb == 8:
        fs.append((P1, P1, P1))
    if nb == 10:
        fs += [(P2, P0, P0), (P0, P2, P0), (P0, P0, P2)]
    return [lambda x, y, z, a=a, b=b, c=c: a(x) * b(y) * c(z) for (a, b, c) in fs]


def gauss_legendre(q):
    return np.polynomial.legendre.leggauss(q)
