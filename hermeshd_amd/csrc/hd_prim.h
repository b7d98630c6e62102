// hd_prim.h -- point physics written around shared primitive variables, for
// the fused kernels: each state's 1/rho, velocity and pressures are computed
// once and used by both the physical flux (flux_calc_pnts_r,
// src/flux.f90:23-193) and the Riemann solver (flux_hllc, src/flux.f90:466-647),
// which in the reference recompute them separately.  Reciprocals use the
// hardware estimate refined by two Newton steps (faithful in FP64), which
// agrees with IEEE division to an ulp; the reference's formulas are otherwise
// kept term by term.
#pragma once
#include "hd_physics.h"

namespace hd {
namespace pr {

__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    r = fma(fma(-x, r, 1.0), r, r);
    return r;
}
// a/b with one residual correction on top of the refined reciprocal
__device__ __forceinline__ double divide(double a, double b) {
    const double r = rcp(b);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}

struct Prim {
    double ri;       // 1/rho
    double u[3];     // velocity
    double pf;       // flux pressure: max(gm1*(E - 0.5/rho*|m|^2), P_floor)
};

template <int NQ>
__device__ __forceinline__ Prim prim(const Phys& ph, const double* Q) {
    Prim s;
    s.ri = rcp(Q[RH]);
    s.u[0] = Q[MX] * s.ri;
    s.u[1] = Q[MY] * s.ri;
    s.u[2] = Q[MZ] * s.ri;
    const double smsq = Q[MX] * Q[MX] + Q[MY] * Q[MY] + Q[MZ] * Q[MZ];
    double P = ph.gm1 * (Q[EN] - 0.5 * s.ri * smsq);
    s.pf = P < ph.P_floor ? ph.P_floor : P;
    return s;
}

// physical flux in direction DIR from primitives (flux_calc_pnts_r, ivis 0/1)
template <int NQ, int DIR>
__device__ __forceinline__ void flux(const Phys& ph, const double* Q, const Prim& s, double* F, const double* S) {
    constexpr int ND = MX + DIR;
    const double un = s.u[DIR];
    const double P = s.pf;
    double Pn[3] = {0., 0., 0.};  // stress row P_{DIR,*}
    if constexpr (NQ >= 11) {
        if (DIR == 0) { Pn[0] = Q[PXX]; Pn[1] = Q[PXY]; Pn[2] = Q[PXZ]; }
        if (DIR == 1) { Pn[0] = Q[PXY]; Pn[1] = Q[PYY]; Pn[2] = Q[PYZ]; }
        if (DIR == 2) { Pn[0] = Q[PXZ]; Pn[1] = Q[PYZ]; Pn[2] = Q[PZZ]; }
    }
    if (S) {
        if (DIR == 0) { Pn[0] -= S[0]; Pn[1] -= S[3]; Pn[2] -= S[4]; }
        if (DIR == 1) { Pn[0] -= S[3]; Pn[1] -= S[1]; Pn[2] -= S[5]; }
        if (DIR == 2) { Pn[0] -= S[4]; Pn[1] -= S[5]; Pn[2] -= S[2]; }
    }
    F[RH] = Q[ND];
    F[MX] = Q[MX] * un + Pn[0];
    F[MY] = Q[MY] * un + Pn[1];
    F[MZ] = Q[MZ] * un + Pn[2];
    F[ND] += P;
    F[EN] = (Q[EN] + P) * un + Pn[0] * s.u[0] + Pn[1] * s.u[1] + Pn[2] * s.u[2];
    if constexpr (NQ >= 11) {
        const double c4 = ph.c4d3cv * un, c2 = -ph.c2d3cv * un;
        F[PXX] = DIR == 0 ? c4 : c2;
        F[PYY] = DIR == 1 ? c4 : c2;
        F[PZZ] = DIR == 2 ? c4 : c2;
        F[PXY] = DIR == 0 ? ph.colvis * s.u[1] : (DIR == 1 ? ph.colvis * s.u[0] : 0.);
        F[PXZ] = DIR == 0 ? ph.colvis * s.u[2] : (DIR == 2 ? ph.colvis * s.u[0] : 0.);
        F[PYZ] = DIR == 1 ? ph.colvis * s.u[2] : (DIR == 2 ? ph.colvis * s.u[1] : 0.);
    }
}

// Numerator of the HLLC contact speed, aR*uR - aL*uL + pL - pR, rounded
// product by product: at a mirror state (the x-high outflow ghost reverses
// the normal momentum) it is exactly zero, and a fused multiply-add would
// leave a signed rounding residue that picks a different star state than
// the reference's arithmetic does.
__device__ __forceinline__ double contact_num(double aR, double uR, double aL, double uL, double pL, double pR) {
#pragma clang fp contract(off)
    return aR * uR - aL * uL + pL - pR;
}

// HLLC (Batten 1997 with Davis speeds) at one face point from primitives;
// out[] holds the central flux on entry and keeps it when no branch applies.
template <int DIR>
__device__ __forceinline__ void hllc(const Phys& ph, const double* QL, const Prim& L, const double* QR, const Prim& R,
                                     const double* FL, const double* FR, double* out) {
    constexpr int IP = MX + DIR, I1 = MX + (DIR + 1) % 3, I2 = MX + (DIR + 2) % 3;
    const double rL = QL[RH], rR = QR[RH];
    const double uL = L.u[DIR], uR = R.u[DIR];
    // HLLC pressure: gm1*(E - 0.5*rho*|v|^2), no floor (src/flux.f90:510-520)
    const double pL = ph.gm1 * (QL[EN] - 0.5 * rL * (L.u[0] * L.u[0] + L.u[1] * L.u[1] + L.u[2] * L.u[2]));
    const double pR = ph.gm1 * (QR[EN] - 0.5 * rR * (R.u[0] * R.u[0] + R.u[1] * R.u[1] + R.u[2] * R.u[2]));
    const double cL = sqrt(ph.gamma * pL * L.ri), cR = sqrt(ph.gamma * pR * R.ri);
    const double SL = fmin(uL - cL, uR - cR);
    const double SR = fmax(uR + cR, uL + cL);
    const double aR = rR * (SR - uR), aL = rL * (SL - uL);
    double sm_den = aR - aL;
    if (sm_den == 0.0) sm_den = ph.rh_floor;
    const double SM = divide(contact_num(aR, uR, aL, uL, pL, pR), sm_den);
    int sel = -1;
    if (SL > 0.0) sel = 0;
    if (SL <= 0.0 && 0.0 < SM) sel = 1;
    if (SM <= 0.0 && 0.0 <= SR) sel = 2;
    if (SR < 0.0) sel = 3;
    if (sel == 0) {
        out[RH] = FL[RH]; out[MX] = FL[MX]; out[MY] = FL[MY]; out[MZ] = FL[MZ]; out[EN] = FL[EN];
    } else if (sel == 3) {
        out[RH] = FR[RH]; out[MX] = FR[MX]; out[MY] = FR[MY]; out[MZ] = FR[MZ]; out[EN] = FR[EN];
    } else if (sel > 0) {
        const bool left = (sel == 1);
        const double* Qk = left ? QL : QR;
        const double* Fk = left ? FL : FR;
        const double Sk = left ? SL : SR;
        const double uk = left ? uL : uR, pk = left ? pL : pR;
        const double pstar = rL * (uL - SL) * (uL - SM) + pL;  // Eq. (36)
        double den = Sk - SM;
        if (den == 0.0) den = ph.rh_floor;
        const double slrm_i = rcp(den);
        const double sq = Sk - uk;
        const double s0 = Qk[RH] * sq * slrm_i;
        const double s1 = (sq * Qk[IP] + pstar - pk) * slrm_i;
        const double s2 = sq * Qk[I1] * slrm_i;
        const double s3 = sq * Qk[I2] * slrm_i;
        const double s4 = (sq * Qk[EN] - pk * uk + pstar * SM) * slrm_i;
        out[RH] = Fk[RH] + Sk * (s0 - Qk[RH]);
        out[IP] = Fk[IP] + Sk * (s1 - Qk[IP]);
        out[I1] = Fk[I1] + Sk * (s2 - Qk[I1]);
        out[I2] = Fk[I2] + Sk * (s3 - Qk[I2]);
        out[EN] = Fk[EN] + Sk * (s4 - Qk[EN]);
    }
}

}  // namespace pr
}  // namespace hd
