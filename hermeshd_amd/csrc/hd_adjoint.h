// hd_adjoint.h -- hand-written reverse mode (vector-Jacobian products) of the
// RK stage, for gradients of a run with respect to its initial state and the
// relaxation rate.
//
// For a stage  out = a*Q^n + b*v,  v = F_ivis(m + dt*R(m)),  R = -glflux + source
// (src/integrator.f90:62-252; F_ivis scales the stress fields by
// 1/(1 + dt*coll) when ivis = 1), the kernel k_stage_vjp computes, for one
// cell c,
//     lam_in[c] = sum_k w_k acc_k[c] + L'(raw)^T ( lp[c] + dt * sum_j (dR_j/dm_c)^T lp[j] )
// with lp = F_ivis'^T (b*lam_out): the transpose of the stage map applied to
// the cotangent of its output.  (dR/dm)^T is gathered per cell: the cell's
// own volume term, and for each of its six faces the face flux's Jacobian
// with respect to this cell's trace, applied to the cotangent that the face
// flux receives from both cells that use it (volume: innerintegral2,
// src/flux.f90:751-928; faces: calc_flux_x/y/z and flux_hllc,
// src/flux.f90:197-647; lift: glflux, src/flux.f90:932-1022; sources:
// src/sources.f90:12-84).  L' is the limiter (src/flux.f90:1077-1164) with
// its active set and factors frozen at the forward values: the factors that
// scaled the modes scale the cotangent.
//
// The point functions are differentiated by hand: pr::flux (physical flux
// from primitives) and pr::hllc (Batten HLLC with Davis speeds), each
// recomputing its forward intermediates and sweeping them backwards.  Branch
// decisions (wave-speed selection, min/max, floors, boundary-ghost choices)
// are those of the forward evaluation at the same state.  The stochastic
// stress and the Roe solver have no adjoint here.
#pragma once
#include "hd_fused.h"

namespace hd {
namespace adj {

// ---------------------------------------------------------------- flux ----
// reverse of flux_point13<DIR> (the 13-moment flux) from primitives:
// bF (cotangent of F) -> bQ (+=)
template <int DIR>
__device__ __forceinline__ void flux13_vjp(const Phys& ph, const double* Q, const pr::Prim& s, const double* bF,
                                           double* bQ) {
    constexpr int ND = MX + DIR;
    const double un = s.u[DIR];
    const double P = s.pf;
    // stress row DIR: indices of T_{DIR,a} (-1: t_zz = -t_xx - t_yy)
    const int trow[3][3] = {{T_PXX, T_PXY, T_PXZ}, {T_PXY, T_PYY, T_PYZ}, {T_PXZ, T_PYZ, -1}};
    double Tn[3];
#pragma unroll
    for (int a = 0; a < 3; a++) Tn[a] = trow[DIR][a] >= 0 ? Q[trow[DIR][a]] : -(Q[T_PXX] + Q[T_PYY]);
    double bun = bF[MX] * Q[MX] + bF[MY] * Q[MY] + bF[MZ] * Q[MZ] + bF[EN] * (Q[EN] + P);
    bun += (DIR == 0 ? ph.c4d3cv : -ph.c2d3cv) * bF[T_PXX] + (DIR == 1 ? ph.c4d3cv : -ph.c2d3cv) * bF[T_PYY];
    bQ[MX] += bF[MX] * un;
    bQ[MY] += bF[MY] * un;
    bQ[MZ] += bF[MZ] * un;
    bQ[EN] += bF[EN] * un;
    bQ[ND] += bF[RH];
    bQ[T_QX + DIR] += bF[EN];
    // heat-flux row: colq * P / rho
    const double bq = bF[T_QX + DIR] * ph.colq;
    const double bP = bF[ND] + bF[EN] * un + bq * s.ri;
    double bri = bq * P;
    double bu[3] = {bF[EN] * Tn[0], bF[EN] * Tn[1], bF[EN] * Tn[2]};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double bt = bF[MX + a] + bF[EN] * s.u[a];
        if (trow[DIR][a] >= 0) {
            bQ[trow[DIR][a]] += bt;
        } else {
            bQ[T_PXX] -= bt;
            bQ[T_PYY] -= bt;
        }
    }
    if (DIR == 0) { bu[1] += ph.colvis * bF[T_PXY]; bu[2] += ph.colvis * bF[T_PXZ]; }
    if (DIR == 1) { bu[0] += ph.colvis * bF[T_PXY]; bu[2] += ph.colvis * bF[T_PYZ]; }
    if (DIR == 2) { bu[0] += ph.colvis * bF[T_PXZ]; bu[1] += ph.colvis * bF[T_PYZ]; }
    bu[DIR] += bun;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        bQ[MX + a] += bu[a] * s.ri;
        bri += bu[a] * Q[MX + a];
    }
    const double smsq = Q[MX] * Q[MX] + Q[MY] * Q[MY] + Q[MZ] * Q[MZ];
    const double Praw = ph.gm1 * (Q[EN] - 0.5 * s.ri * smsq);
    if (!(Praw < ph.P_floor)) {
        bQ[EN] += ph.gm1 * bP;
        bri += -0.5 * ph.gm1 * smsq * bP;
#pragma unroll
        for (int a = 0; a < 3; a++) bQ[MX + a] += -ph.gm1 * s.ri * Q[MX + a] * bP;
    }
    bQ[RH] += -s.ri * s.ri * bri;
}

// reverse of pr::prim + pr::flux<NQ,DIR>: bF (cotangent of F) -> bQ (+=)
template <int NQ, int DIR>
__device__ __forceinline__ void flux_vjp(const Phys& ph, const double* Q, const pr::Prim& s, const double* bF,
                                         double* bQ) {
    if constexpr (NQ == 13) {
        flux13_vjp<DIR>(ph, Q, s, bF, bQ);
        return;
    }
    constexpr int ND = MX + DIR;
    const double un = s.u[DIR];
    double Pn[3] = {0., 0., 0.};
    int pidx[3] = {0, 0, 0};
    if constexpr (NQ == 11) {
        if (DIR == 0) { pidx[0] = PXX; pidx[1] = PXY; pidx[2] = PXZ; }
        if (DIR == 1) { pidx[0] = PXY; pidx[1] = PYY; pidx[2] = PYZ; }
        if (DIR == 2) { pidx[0] = PXZ; pidx[1] = PYZ; pidx[2] = PZZ; }
        Pn[0] = Q[pidx[0]]; Pn[1] = Q[pidx[1]]; Pn[2] = Q[pidx[2]];
    }
    const double P = s.pf;
    double bun = bF[MX] * Q[MX] + bF[MY] * Q[MY] + bF[MZ] * Q[MZ] + bF[EN] * (Q[EN] + P);
    bQ[MX] += bF[MX] * un;
    bQ[MY] += bF[MY] * un;
    bQ[MZ] += bF[MZ] * un;
    bQ[EN] += bF[EN] * un;
    bQ[ND] += bF[RH];
    const double bP = bF[ND] + bF[EN] * un;
    double bu[3] = {bF[EN] * Pn[0], bF[EN] * Pn[1], bF[EN] * Pn[2]};
    if constexpr (NQ == 11) {
#pragma unroll
        for (int a = 0; a < 3; a++) bQ[pidx[a]] += bF[MX + a] + bF[EN] * s.u[a];
        bun += (DIR == 0 ? ph.c4d3cv : -ph.c2d3cv) * bF[PXX] + (DIR == 1 ? ph.c4d3cv : -ph.c2d3cv) * bF[PYY] +
               (DIR == 2 ? ph.c4d3cv : -ph.c2d3cv) * bF[PZZ];
        if (DIR == 0) { bu[1] += ph.colvis * bF[PXY]; bu[2] += ph.colvis * bF[PXZ]; }
        if (DIR == 1) { bu[0] += ph.colvis * bF[PXY]; bu[2] += ph.colvis * bF[PYZ]; }
        if (DIR == 2) { bu[0] += ph.colvis * bF[PXZ]; bu[1] += ph.colvis * bF[PYZ]; }
    }
    bu[DIR] += bun;
    // primitives: u_a = m_a/rho, P = gm1*(E - 0.5/rho*|m|^2) unless floored
    double bri = 0.;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        bQ[MX + a] += bu[a] * s.ri;
        bri += bu[a] * Q[MX + a];
    }
    const double smsq = Q[MX] * Q[MX] + Q[MY] * Q[MY] + Q[MZ] * Q[MZ];
    const double Praw = ph.gm1 * (Q[EN] - 0.5 * s.ri * smsq);
    if (!(Praw < ph.P_floor)) {
        bQ[EN] += ph.gm1 * bP;
        bri += -0.5 * ph.gm1 * smsq * bP;
#pragma unroll
        for (int a = 0; a < 3; a++) bQ[MX + a] += -ph.gm1 * s.ri * Q[MX + a] * bP;
    }
    bQ[RH] += -s.ri * s.ri * bri;
}

// ---------------------------------------------------------------- HLLC ----
// reverse of pr::hllc<DIR> for the Euler fields: mu (cotangent of out[0..4])
// -> bQL, bQR (+=, Euler fields), bFL, bFR (+=, Euler fields).  Returns false
// when no branch applied (out kept the central flux).
template <int DIR>
__device__ __forceinline__ bool hllc_vjp(const Phys& ph, const double* QL, const pr::Prim& L, const double* QR,
                                         const pr::Prim& R, const double* mu, double* bQL, double* bQR, double* bFL,
                                         double* bFR) {
    constexpr int IP = MX + DIR, I1 = MX + (DIR + 1) % 3, I2 = MX + (DIR + 2) % 3;
    // ---- forward, as pr::hllc
    const double rL = QL[RH], rR = QR[RH];
    const double uL = L.u[DIR], uR = R.u[DIR];
    const double v2L = L.u[0] * L.u[0] + L.u[1] * L.u[1] + L.u[2] * L.u[2];
    const double v2R = R.u[0] * R.u[0] + R.u[1] * R.u[1] + R.u[2] * R.u[2];
    const double pL = ph.gm1 * (QL[EN] - 0.5 * rL * v2L);
    const double pR = ph.gm1 * (QR[EN] - 0.5 * rR * v2R);
    const double cL = sqrt(ph.gamma * pL * L.ri), cR = sqrt(ph.gamma * pR * R.ri);
    const bool sl_left = (uL - cL) <= (uR - cR);   // fmin picks uL - cL
    const bool sr_right = (uR + cR) >= (uL + cL);  // fmax picks uR + cR
    const double SL = fmin(uL - cL, uR - cR);
    const double SR = fmax(uR + cR, uL + cL);
    const double aR = rR * (SR - uR), aL = rL * (SL - uL);
    double sm_den = aR - aL;
    const bool den_floor = (sm_den == 0.0);
    if (den_floor) sm_den = ph.rh_floor;
    const double SM = pr::divide(pr::contact_num(aR, uR, aL, uL, pL, pR), sm_den);
    int sel = -1;
    if (SL > 0.0) sel = 0;
    if (SL <= 0.0 && 0.0 < SM) sel = 1;
    if (SM <= 0.0 && 0.0 <= SR) sel = 2;
    if (SR < 0.0) sel = 3;
    if (sel < 0) return false;
    if (sel == 0) {
#pragma unroll
        for (int f = 0; f < 5; f++) bFL[f] += mu[f];
        return true;
    }
    if (sel == 3) {
#pragma unroll
        for (int f = 0; f < 5; f++) bFR[f] += mu[f];
        return true;
    }
    const bool left = (sel == 1);
    const double* Qk = left ? QL : QR;
    const double Sk = left ? SL : SR;
    const double uk = left ? uL : uR, pk = left ? pL : pR;
    const double pstar = rL * (uL - SL) * (uL - SM) + pL;
    double den = Sk - SM;
    const bool kden_floor = (den == 0.0);
    if (kden_floor) den = ph.rh_floor;
    const double rid = pr::rcp(den);
    const double sq = Sk - uk;
    const double t1 = sq * Qk[IP] + pstar - pk;
    const double t4 = sq * Qk[EN] - pk * uk + pstar * SM;
    const double s[5] = {Qk[RH] * sq * rid, t1 * rid, sq * Qk[I1] * rid, sq * Qk[I2] * rid, t4 * rid};
    const int fld[5] = {RH, IP, I1, I2, EN};
    // ---- reverse: out[f] = Fk[f] + Sk*(s_f - Qk[f])
    double bQk[5] = {0., 0., 0., 0., 0.};  // indexed like fld
    double bSk = 0., bs[5];
#pragma unroll
    for (int e = 0; e < 5; e++) {
        const double m = mu[fld[e]];
        bSk += m * (s[e] - Qk[fld[e]]);
        bs[e] = m * Sk;
        bQk[e] -= m * Sk;
    }
    double* bFk = left ? bFL : bFR;
#pragma unroll
    for (int f = 0; f < 5; f++) bFk[f] += mu[f];
    double bsq = 0., brid = 0., bpstar = 0., bpk = 0., buk = 0., bSM = 0.;
    // s0 = Qk[RH]*sq*rid
    bQk[0] += bs[0] * sq * rid; bsq += bs[0] * Qk[RH] * rid; brid += bs[0] * Qk[RH] * sq;
    // s1 = t1*rid
    bsq += bs[1] * Qk[IP] * rid; bQk[1] += bs[1] * sq * rid; bpstar += bs[1] * rid; bpk -= bs[1] * rid;
    brid += bs[1] * t1;
    // s2, s3
    bsq += bs[2] * Qk[I1] * rid; bQk[2] += bs[2] * sq * rid; brid += bs[2] * sq * Qk[I1];
    bsq += bs[3] * Qk[I2] * rid; bQk[3] += bs[3] * sq * rid; brid += bs[3] * sq * Qk[I2];
    // s4 = t4*rid
    bsq += bs[4] * Qk[EN] * rid; bQk[4] += bs[4] * sq * rid; bpk -= bs[4] * uk * rid; buk -= bs[4] * pk * rid;
    bpstar += bs[4] * SM * rid; bSM += bs[4] * pstar * rid; brid += bs[4] * t4;
    // rid = 1/den, den = Sk - SM
    if (!kden_floor) {
        const double bden = -brid * rid * rid;
        bSk += bden;
        bSM -= bden;
    }
    // sq = Sk - uk
    bSk += bsq;
    buk -= bsq;
    // pstar = rL*(uL-SL)*(uL-SM) + pL
    double brL = bpstar * (uL - SL) * (uL - SM);
    double buL = bpstar * rL * ((uL - SM) + (uL - SL));
    double bSL = -bpstar * rL * (uL - SM);
    bSM -= bpstar * rL * (uL - SL);
    double bpL = bpstar, bpR = 0., brR = 0., buR = 0., bSR = 0.;
    if (left) { bpL += bpk; buL += buk; bSL += bSk; } else { bpR += bpk; buR += buk; bSR += bSk; }
    // SM = sm_num/sm_den
    const double bnum = bSM / sm_den;
    const double bdenm = den_floor ? 0. : -bSM * SM / sm_den;
    double baR = bnum * uR + bdenm, baL = -bnum * uL - bdenm;
    buR += bnum * aR;
    buL -= bnum * aL;
    bpL += bnum;
    bpR -= bnum;
    // aR = rR*(SR-uR), aL = rL*(SL-uL)
    brR += baR * (SR - uR); bSR += baR * rR; buR -= baR * rR;
    brL += baL * (SL - uL); bSL += baL * rL; buL -= baL * rL;
    // SL = min(uL-cL, uR-cR), SR = max(uR+cR, uL+cL)
    double bcL = 0., bcR = 0.;
    if (sl_left) { buL += bSL; bcL -= bSL; } else { buR += bSL; bcR -= bSL; }
    if (sr_right) { buR += bSR; bcR += bSR; } else { buL += bSR; bcL += bSR; }
    // c = sqrt(gamma*p*ri)
    double briL = 0., briR = 0.;
    if (cL > 0.) { bpL += bcL * ph.gamma * L.ri / (2. * cL); briL += bcL * ph.gamma * pL / (2. * cL); }
    if (cR > 0.) { bpR += bcR * ph.gamma * R.ri / (2. * cR); briR += bcR * ph.gamma * pR / (2. * cR); }
    // scatter the star-state field cotangents onto side k
    double* bQkf = left ? bQL : bQR;
    bQkf[RH] += bQk[0]; bQkf[IP] += bQk[1]; bQkf[I1] += bQk[2]; bQkf[I2] += bQk[3]; bQkf[EN] += bQk[4];
    // p = gm1*(E - 0.5*rho*|u|^2), u_a = m_a*ri, ri = 1/rho; u_dir also enters directly
    {
        double bu[3] = {0., 0., 0.};
        bu[DIR] += buL;
        bQL[EN] += ph.gm1 * bpL;
        brL += -0.5 * ph.gm1 * v2L * bpL;
#pragma unroll
        for (int a = 0; a < 3; a++) bu[a] += -ph.gm1 * rL * L.u[a] * bpL;
#pragma unroll
        for (int a = 0; a < 3; a++) { bQL[MX + a] += bu[a] * L.ri; briL += bu[a] * QL[MX + a]; }
        bQL[RH] += brL - L.ri * L.ri * briL;
    }
    {
        double bu[3] = {0., 0., 0.};
        bu[DIR] += buR;
        bQR[EN] += ph.gm1 * bpR;
        brR += -0.5 * ph.gm1 * v2R * bpR;
#pragma unroll
        for (int a = 0; a < 3; a++) bu[a] += -ph.gm1 * rR * R.u[a] * bpR;
#pragma unroll
        for (int a = 0; a < 3; a++) { bQR[MX + a] += bu[a] * R.ri; briR += bu[a] * QR[MX + a]; }
        bQR[RH] += brR - R.ri * R.ri * briR;
    }
    return true;
}

// face flux reverse: mu (cotangent of the face flux, all fields) -> bQL, bQR (+=)
template <int NQ, int RM, int DIR>
__device__ __forceinline__ void face_vjp(const Phys& ph, const double* QL, const double* QR, const double* mu,
                                         double* bQL, double* bQR) {
    const pr::Prim sL = pr::prim<NQ>(ph, QL), sR = pr::prim<NQ>(ph, QR);
    double bFL[NQ], bFR[NQ];
#pragma unroll
    for (int f = 0; f < NQ; f++) { bFL[f] = 0.5 * mu[f]; bFR[f] = 0.5 * mu[f]; }
    if constexpr (RM == RIEMANN_HLLC) {
        double zL[5] = {0., 0., 0., 0., 0.}, zR[5] = {0., 0., 0., 0., 0.};
        if (hllc_vjp<DIR>(ph, QL, sL, QR, sR, mu, bQL, bQR, zL, zR)) {
            // HLLC replaced the central flux of rh..en
#pragma unroll
            for (int f = 0; f < 5; f++) { bFL[f] = zL[f]; bFR[f] = zR[f]; }
        }
    }
    if constexpr (RM == RIEMANN_ROE) {
        // Roe: 0.5 (FL + FR) minus the dissipation of rh..en (or the
        // Lax-Friedrichs fallback); the reverse adds both parts
#pragma unroll
        for (int f = 0; f < 5; f++) { bFL[f] = 0.; bFR[f] = 0.; }
        roe_point_vjp<DIR>(ph, QL, QR, mu, bQL, bQR, bFL, bFR);
    }
    flux_vjp<NQ, DIR>(ph, QL, sL, bFL, bQL);
    flux_vjp<NQ, DIR>(ph, QR, sR, bFR, bQR);
}

// ------------------------------------------------------------ limiter ----
// The limiter's factors for one cell (the same decisions and values as
// fz::limit_cell): modes >= 2 of rh are scaled by th_rho and then, with the
// other Euler fields, by th_p; `zero` zeroes modes >= 2 of rh..en; `floor`
// also replaces the density average by rh_floor.
struct LimFactors {
    bool floor, zero;
    double th_rho, th_p;
};

template <int NB, int QP, int MQ>
__device__ __forceinline__ LimFactors limit_factors(const double (&m0)[MQ][NB], const LimParams& lp) {
    LimFactors F{false, false, 1.0, 1.0};
    if (m0[RH][0] < lp.rh_floor) {
        F.floor = true;
        F.zero = true;
        return F;
    }
    if (fz::limiter_inactive<NB, MQ>(m0, lp)) return F;
    const auto& T = fz::face_tab<NB>();
    double rho[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) rho[r] = m0[RH][r];
    double rmin = 0.;
#pragma unroll 1
    for (int p = 0; p < 6 * QP * QP; p++) {
        const double* b = T.bf[p / (2 * QP * QP)][(p / (QP * QP)) % 2][p % (QP * QP)];
        double s = 0.;
#pragma unroll
        for (int r = 0; r < NB; r++) s += b[r] * rho[r];
        rmin = (p == 0) ? s : (s < rmin ? s : rmin);
    }
    if (rmin < lp.rh_floor) {
        double theta = (lp.rh_floor - rho[0]) / (rmin - rho[0]);
        if (theta > 1.) theta = 1.;
        if (theta < 0.) theta = 0.;
        F.th_rho = fabs(theta);
#pragma unroll
        for (int r = 1; r < NB; r++) rho[r] = F.th_rho * rho[r];
    }
    const double Pave =
        lp.gm1 * (m0[EN][0] - 0.5 * (m0[MX][0] * m0[MX][0] + m0[MY][0] * m0[MY][0] + m0[MZ][0] * m0[MZ][0]) / rho[0]);
    if (Pave < lp.epsP) {
        F.zero = true;
        return F;
    }
    double pmin = 0.;
#pragma unroll 1
    for (int p = 0; p < 6 * QP * QP; p++) {
        const double* b = T.bf[p / (2 * QP * QP)][(p / (QP * QP)) % 2][p % (QP * QP)];
        double e[5];
#pragma unroll
        for (int f = 0; f < 5; f++) {
            double s = 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) s += b[r] * (f == RH ? rho[r] : m0[f][r]);
            e[f] = s;
        }
        const double dni = pr::rcp(e[RH]);
        const double P = lp.gm1 * (e[EN] - 0.5 * dni * (e[MX] * e[MX] + e[MY] * e[MY] + e[MZ] * e[MZ]));
        pmin = (p == 0) ? P : (P < pmin ? P : pmin);
    }
    if (pmin < lp.epsP && Pave != pmin) {
        double theta = (Pave - lp.epsP) / (Pave - pmin);
        if (theta > 1.) theta = 1.;
        if (theta < 0.) theta = 0.;
        F.th_p = theta;
    }
    return F;
}

// cotangent through the frozen limiter, in place
template <int NB, int MQ>
__device__ __forceinline__ void limit_vjp(double (&lam)[MQ][NB], const LimFactors& F) {
    if (F.zero) {
#pragma unroll
        for (int r = 1; r < NB; r++)
#pragma unroll
            for (int f = 0; f < 5; f++) lam[f][r] = 0.0;
        if (F.floor) lam[RH][0] = 0.0;
        return;
    }
#pragma unroll
    for (int r = 1; r < NB; r++) {
        lam[RH][r] *= F.th_rho * F.th_p;
#pragma unroll
        for (int f = 1; f < 5; f++) lam[f][r] *= F.th_p;
    }
}

// ------------------------------------------------------------- stage -----
// grids up to this many cells use the four-wave split of the transpose (when
// its three 64-cell partial sums fit the 160 KB of LDS)
constexpr long kSplitCells = 1L << 17;
template <int NB, int NQ>
constexpr bool split_fits() { return 3L * NQ * NB * 64 * 8 <= 150L * 1024; }

struct AdjParams {
    const double* dt;          // device time step
    double coll;
    double coll_q;             // 13 moments: relaxation rate of the heat flux
    int ivis;
    double a, b;               // out = a*Q^n + b*v
    double bf[3];              // body force
    const double* acc[2];      // lam_in += w[k]*acc[k] (nullptr: none)
    double w[2];
    const double* q_raw;       // stage input before the limiter (nullptr: no limiter)
    LimParams lp;
    const double* out;         // forward stage output and Q^n, for d/dcoll (ivis 1)
    const double* qn;
    double* partial;           // per-block partial sums of d(lam.out)/dcoll
    int rhs_only;              // 1: lam_in = (dR/dm)^T lam (no stage map, no limiter, no sums)
    const double* inflow;      // x-low inflow exterior state (ny,nz,nface,nQ), or nullptr
};

// scale * lam through the transposed relaxation: the stress fields by faci,
// the 13-moment heat flux by faciq
template <int NB, int NQ>
__device__ __forceinline__ void load_lp(const double* __restrict__ lam, const Grid& g, long c, double scale,
                                        double faci, double faciq, double (&l)[NQ][NB]) {
#pragma unroll
    for (int r = 0; r < NB; r++)
#pragma unroll
        for (int f = 0; f < NQ; f++) {
            double v = scale * lam[((long)r * NQ + f) * g.N + c];
            if (NQ == 13 && f >= T_QX) v *= faciq;
            else if (f >= PXX) v *= faci;
            l[f][r] = v;
        }
}

// Lifted cotangent of a face side in face-polynomial form: the transposed
// lift sum_r cf[side][p][r] l_r equals sum_k wb_k(p) L_k with
// L_k = sum over the modes r of pair k of (+-0.25 cb_r Lface_r) l_r.
template <int NB, int NQ, int DIR, int SIDE, int NK>
__device__ __forceinline__ void lift_coeffs(const double (&l)[NQ][NB], double (&L)[NQ][NK]) {
#pragma unroll
    for (int f = 0; f < NQ; f++)
#pragma unroll
        for (int k = 0; k < NK; k++) {
            double s = 0.;
            bool first = true;
#pragma unroll
            for (int r = 0; r < NB; r++) {
                if (fz::pair_index(fz::deg<NB>(r, fz::row_axis(DIR)), fz::deg<NB>(r, fz::col_axis(DIR))) != k) continue;
                constexpr double q = SIDE ? -0.25 : 0.25;
                const double c = q * fz::cbasis<NB>(r) * fz::Lface(fz::deg<NB>(r, DIR), SIDE);
                s = first ? c * l[f][r] : s + c * l[f][r];
                first = false;
            }
            L[f][k] = s;
        }
}

// one face of the cell: adds this cell's share of (dR/dm)^T lp to rt
template <int NB, int QP, int NQ, int RM, int DIR, int SIDE>
__device__ __forceinline__ void face_vjp_cell(const double* __restrict__ q, const double* __restrict__ lam,
                                              const Grid& g, const fz::CellPos& cp, const double (&m)[NQ][NB],
                                              const double (&lo)[NQ][NB], double (&rt)[NQ][NB], const Phys& ph,
                                              double dxi_d, double scale, double faci, double faciq,
                                              const double* __restrict__ inflow) {
    constexpr int NF = QP * QP;
    const auto& T = fz::face_tab<NB>();
    const int n_ax = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
    const int a = DIR == 0 ? cp.i : DIR == 1 ? cp.j : cp.k;
    const bool at_edge = SIDE ? (a == n_ax - 1) : (a == 0);
    const int bc = g.bc[2 * DIR + SIDE];
    const bool ext = at_edge && bc != BC_PERIODIC;
    // the neighbour enters through its face polynomial: its trace
    // sum_k kn_k b_k(p) and its lifted cotangent sum_k ln_k wb_k(p), so its
    // modes and cotangents are collapsed as they are loaded
    constexpr int NK = fz::face_nk<NB>();
    const auto& FP = fz::face_poly<NB>();
    double ko[NQ][NK], kn[NQ][NK], ln[NQ][NK];
    // both sides' traces with the forward kernel's arithmetic, so that every
    // Riemann branch is taken as in the forward stage
    fz::face_coeffs<NB, NQ, DIR, SIDE, NK>(m, ko);
    if (!ext) {
        int an = SIDE ? a + 1 : a - 1;
        if (an < 0) an += n_ax;
        if (an >= n_ax) an -= n_ax;
        const long cn = DIR == 0 ? (cp.c - cp.i + an)
                      : DIR == 1 ? (cp.c + (long)(an - cp.j) * g.nx)
                                 : (cp.c + (long)(an - cp.k) * g.nx * g.ny);
        double nm[NQ][NB], nl[NQ][NB];
        fz::load_modes<NB, NQ>(q, g, cn, nm);
        fz::face_coeffs<NB, NQ, DIR, 1 - SIDE, NK>(nm, kn);
        load_lp<NB, NQ>(lam, g, cn, scale, faci, faciq, nl);
        lift_coeffs<NB, NQ, DIR, 1 - SIDE, NK>(nl, ln);
    }
    bool zero_normal = false;
    if (ext && bc == BC_OUTFLOW && !(DIR == 0 && SIDE == 1)) {
        double v = 0.;
        for (int p = 0; p < NF; p++) {
            double s = 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) s += T.bf[DIR][SIDE][p][r] * m[MX + DIR][r];
            v = (p == 0) ? s : (SIDE ? fmin(v, s) : fmax(v, s));
        }
        zero_normal = SIDE ? (v < 0.0) : (v > 0.0);
    }
#pragma unroll 1
    for (int p = 0; p < NF; p++) {
        double own[NQ], oth[NQ], mu[NQ];
        fz::face_eval<NQ, NK>(ko, FP.b[p], own);
#pragma unroll
        for (int f = 0; f < NQ; f++) {
            double t = 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) t += T.cf[DIR][SIDE][p][r] * lo[f][r];
            mu[f] = t;
        }
        bool flip_mx = false;
        if (!ext) {
            fz::face_eval<NQ, NK>(kn, FP.b[p], oth);
#pragma unroll
            for (int f = 0; f < NQ; f++) {
                double t = ln[f][0] * FP.wb[p][0];
#pragma unroll
                for (int k = 1; k < NK; k++) t += ln[f][k] * FP.wb[p][k];
                mu[f] += t;
            }
        } else {
#pragma unroll
            for (int f = 0; f < NQ; f++) oth[f] = own[f];
            if (bc == BC_NOSLIP) {
                oth[MX] = oth[MY] = oth[MZ] = 0.0;
            } else if (bc == BC_WALL) {
                oth[MX + DIR] = 0.0;
            } else if (bc == BC_OUTFLOW) {
                if (DIR == 0 && SIDE == 1) {
                    if (oth[MX] > 0.0) { oth[MX] = -own[MX]; flip_mx = true; }
                } else if (zero_normal) {
                    oth[MX + DIR] = 0.0;
                }
            }
            if constexpr (DIR == 0 && SIDE == 0) {
                if (bc == BC_INFLOW) fz::inflow_ghost<NQ>(inflow, g, cp, p, NF, oth);
            }
        }
#pragma unroll
        for (int f = 0; f < NQ; f++) mu[f] *= dxi_d;
        double bL[NQ], bR[NQ];
#pragma unroll
        for (int f = 0; f < NQ; f++) { bL[f] = 0.; bR[f] = 0.; }
        const double* QL = SIDE ? own : oth;
        const double* QR = SIDE ? oth : own;
        face_vjp<NQ, RM, DIR>(ph, QL, QR, mu, bL, bR);
        double* bo = SIDE ? bL : bR;
        const double* bt = SIDE ? bR : bL;
        if (ext) {
            // ghost = G(own): add G^T of the ghost's cotangent
#pragma unroll
            for (int f = 0; f < NQ; f++) {
                double gcoef = 1.0;
                if (bc == BC_NOSLIP && (f == MX || f == MY || f == MZ)) gcoef = 0.0;
                if (bc == BC_WALL && f == MX + DIR) gcoef = 0.0;
                if (bc == BC_INFLOW && f <= EN) gcoef = 0.0;   // prescribed: no dependence on the interior
                if (bc == BC_OUTFLOW) {
                    if (DIR == 0 && SIDE == 1) {
                        if (f == MX && flip_mx) gcoef = -1.0;
                    } else if (zero_normal && f == MX + DIR) {
                        gcoef = 0.0;
                    }
                }
                bo[f] += gcoef * bt[f];
            }
        }
#pragma unroll
        for (int f = 0; f < NQ; f++) {
            const double v = bo[f];
#pragma unroll
            for (int r = 0; r < NB; r++) rt[f][r] += T.bf[DIR][SIDE][p][r] * v;
        }
    }
}

template <int NB, int QP, int NQ>
__device__ __forceinline__ void volume_vjp(const double (&m)[NQ][NB], const double (&lo)[NQ][NB], double (&rt)[NQ][NB],
                                           const Phys& ph, double dxi, double dyi, double dzi) {
#pragma unroll
    for (int i = 0; i < QP; i++)
#pragma unroll
        for (int j = 0; j < QP; j++)
#pragma unroll
            for (int k = 0; k < QP; k++) {
                double Qp[NQ], mx[NQ], my[NQ], mz[NQ];
#pragma unroll
                for (int f = 0; f < NQ; f++) {
                    double s = 0., ax = 0., ay = 0., az = 0.;
#pragma unroll
                    for (int r = 0; r < NB; r++) {
                        const double b = fz::phi_int<NB, QP>(r, i, j, k);
                        if (b != 0.0) s += b * m[f][r];
                        const double cx = fz::cvol<NB, QP>(0, r, i, j, k);
                        const double cy = fz::cvol<NB, QP>(1, r, i, j, k);
                        const double cz = fz::cvol<NB, QP>(2, r, i, j, k);
                        if (r > 0 && cx != 0.0) ax += cx * lo[f][r];
                        if (r > 0 && cy != 0.0) ay += cy * lo[f][r];
                        if (r > 0 && cz != 0.0) az += cz * lo[f][r];
                    }
                    Qp[f] = s;
                    mx[f] = dxi * ax;
                    my[f] = dyi * ay;
                    mz[f] = dzi * az;
                }
                const pr::Prim st = pr::prim<NQ>(ph, Qp);
                double bQ[NQ];
#pragma unroll
                for (int f = 0; f < NQ; f++) bQ[f] = 0.;
                flux_vjp<NQ, 0>(ph, Qp, st, mx, bQ);
                flux_vjp<NQ, 1>(ph, Qp, st, my, bQ);
                flux_vjp<NQ, 2>(ph, Qp, st, mz, bQ);
#pragma unroll
                for (int f = 0; f < NQ; f++)
#pragma unroll
                    for (int r = 0; r < NB; r++) {
                        const double b = fz::phi_int<NB, QP>(r, i, j, k);
                        if (b != 0.0) rt[f][r] += b * bQ[f];
                    }
                fz::point_barrier();
            }
}

// Parts of the transpose of one cell: bit 0 volume + sources, bits 1..3 the
// x, y, z face pairs.  Adds to rt; returns the cell's share of d/dcoll.
template <int NB, int QP, int NQ, int RM>
__device__ __forceinline__ double cell_vjp_parts(int parts, const double* __restrict__ q,
                                                 const double* __restrict__ lam, const Grid& g, const Phys& ph,
                                                 const AdjParams& ap, const fz::CellPos& cp, double dt, double faci,
                                                 double faciq, const double (&m)[NQ][NB], const double (&lo)[NQ][NB],
                                                 double (&rt)[NQ][NB]) {
    double dcoll = 0.;
    const long c = cp.c;
    if (parts & 1) {
        volume_vjp<NB, QP, NQ>(m, lo, rt, ph, g.dxi, g.dyi, g.dzi);
        fz::fence();
        // sources: ivis 0 explicit relaxation -coll*P; body force rho*g, m.g
        if constexpr (NQ >= 11) {
            // the stress fields (and, with 13 moments, the heat flux at its own rate)
            constexpr int SLAST = NQ == 13 ? T_PYZ : PYZ;
            if (ap.ivis == 0 && ap.coll != 0.) {
#pragma unroll
                for (int f = PXX; f <= SLAST; f++)
#pragma unroll
                    for (int r = 0; r < NB; r++) {
                        rt[f][r] += -ap.coll * lo[f][r];
                        dcoll += -dt * lo[f][r] * m[f][r];
                    }
            }
            if constexpr (NQ == 13) {
                if (ap.ivis == 0 && ap.coll_q != 0.) {
#pragma unroll
                    for (int f = T_QX; f <= T_QZ; f++)
#pragma unroll
                        for (int r = 0; r < NB; r++) rt[f][r] += -ap.coll_q * lo[f][r];
                }
            }
            if (ap.ivis == 1 && ap.coll != 0. && ap.out) {
                // v = faci*(m + dt*R): dv/dcoll = -dt*faci*v, b*v = out - a*Q^n;
                // the branch on a stays outside the element loop so that the
                // loads are issued together
                if (ap.a != 0.) {
#pragma unroll
                    for (int f = PXX; f <= SLAST; f++)
#pragma unroll
                        for (int r = 0; r < NB; r++) {
                            const long o = ((long)r * NQ + f) * g.N + c;
                            const double bv = ap.out[o] - ap.a * ap.qn[o];
                            dcoll += -dt * faci * lam[o] * bv;
                        }
                } else {
#pragma unroll
                    for (int f = PXX; f <= SLAST; f++)
#pragma unroll
                        for (int r = 0; r < NB; r++) {
                            const long o = ((long)r * NQ + f) * g.N + c;
                            const double bv = ap.out[o] - 0.;
                            dcoll += -dt * faci * lam[o] * bv;
                        }
                }
            }
        }
        if (ap.bf[0] != 0. || ap.bf[1] != 0. || ap.bf[2] != 0.) {
#pragma unroll
            for (int r = 0; r < NB; r++) {
                rt[RH][r] += ap.bf[0] * lo[MX][r] + ap.bf[1] * lo[MY][r] + ap.bf[2] * lo[MZ][r];
                rt[MX][r] += ap.bf[0] * lo[EN][r];
                rt[MY][r] += ap.bf[1] * lo[EN][r];
                rt[MZ][r] += ap.bf[2] * lo[EN][r];
            }
        }
    }
    if (parts & 2) {
        face_vjp_cell<NB, QP, NQ, RM, 0, 0>(q, lam, g, cp, m, lo, rt, ph, g.dxi, ap.b, faci, faciq, ap.inflow);
        fz::fence();
        face_vjp_cell<NB, QP, NQ, RM, 0, 1>(q, lam, g, cp, m, lo, rt, ph, g.dxi, ap.b, faci, faciq, ap.inflow);
        fz::fence();
    }
    if (parts & 4) {
        face_vjp_cell<NB, QP, NQ, RM, 1, 0>(q, lam, g, cp, m, lo, rt, ph, g.dyi, ap.b, faci, faciq, ap.inflow);
        fz::fence();
        face_vjp_cell<NB, QP, NQ, RM, 1, 1>(q, lam, g, cp, m, lo, rt, ph, g.dyi, ap.b, faci, faciq, ap.inflow);
        fz::fence();
    }
    if (parts & 8) {
        face_vjp_cell<NB, QP, NQ, RM, 2, 0>(q, lam, g, cp, m, lo, rt, ph, g.dzi, ap.b, faci, faciq, ap.inflow);
        fz::fence();
        face_vjp_cell<NB, QP, NQ, RM, 2, 1>(q, lam, g, cp, m, lo, rt, ph, g.dzi, ap.b, faci, faciq, ap.inflow);
        fz::fence();
    }
    return dcoll;
}

// lam_in = sum_k w_k acc_k + L'^T (lo + dt*rt), or rt alone in rhs mode
template <int NB, int QP, int NQ>
__device__ __forceinline__ void cell_vjp_finish(double* __restrict__ lam_in, const Grid& g, const AdjParams& ap,
                                                long c, double dt, const double (&lo)[NQ][NB], double (&rt)[NQ][NB]) {
    if (ap.rhs_only) {
#pragma unroll
        for (int f = 0; f < NQ; f++)
#pragma unroll
            for (int r = 0; r < NB; r++) lam_in[((long)r * NQ + f) * g.N + c] = rt[f][r];
        return;
    }
#pragma unroll
    for (int f = 0; f < NQ; f++)
#pragma unroll
        for (int r = 0; r < NB; r++) rt[f][r] = lo[f][r] + dt * rt[f][r];
    // each accumulated term under one branch, its values loaded together
#pragma unroll
    for (int t = 0; t < 2; t++) {
        if (ap.acc[t]) {
            double av[NQ][NB];
            fz::load_modes<NB, NQ>(ap.acc[t], g, c, av);
#pragma unroll
            for (int f = 0; f < NQ; f++)
#pragma unroll
                for (int r = 0; r < NB; r++) rt[f][r] += ap.w[t] * av[f][r];
        }
    }
    if (ap.q_raw) {
        double raw[NQ][NB];
        fz::load_modes<NB, NQ>(ap.q_raw, g, c, raw);
        const LimFactors F = limit_factors<NB, QP, NQ>(raw, ap.lp);
        limit_vjp<NB, NQ>(rt, F);
    }
#pragma unroll
    for (int f = 0; f < NQ; f++)
#pragma unroll
        for (int r = 0; r < NB; r++) lam_in[((long)r * NQ + f) * g.N + c] = rt[f][r];
}

// per-block partial of d(lam.out)/dcoll (every thread of the block calls it)
__device__ __forceinline__ void block_partial(double v, double* partial) {
    __shared__ double red[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
        partial[blockIdx.x] = s;
    }
}

__device__ __forceinline__ fz::CellPos cell_pos(const Grid& g, long c) {
    fz::CellPos cp;
    cp.c = c;
    cp.i = (int)(c % g.nx);
    cp.j = (int)((c / g.nx) % g.ny);
    cp.k = (int)(c / ((long)g.nx * g.ny));
    return cp;
}

// One thread per cell.
template <int NB, int QP, int NQ, int RM>
__global__ __launch_bounds__(256) void k_stage_vjp(const double* __restrict__ q, const double* __restrict__ lam,
                                                   double* __restrict__ lam_in, Grid g, Phys ph, AdjParams ap) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double dcoll = 0.;
    if (c < g.N) {
        const fz::CellPos cp = cell_pos(g, c);
        const double dt = *ap.dt;
        const double faci = (ap.ivis == 1 && !ap.rhs_only) ? 1. / (1. + dt * ap.coll) : 1.0;
        const double faciq = (ap.ivis == 1 && !ap.rhs_only) ? 1. / (1. + dt * ap.coll_q) : 1.0;
        double m[NQ][NB], lo[NQ][NB], rt[NQ][NB];
        fz::load_modes<NB, NQ>(q, g, c, m);
        load_lp<NB, NQ>(lam, g, c, ap.b, faci, faciq, lo);
#pragma unroll
        for (int f = 0; f < NQ; f++)
#pragma unroll
            for (int r = 0; r < NB; r++) rt[f][r] = 0.;
        dcoll = cell_vjp_parts<NB, QP, NQ, RM>(15, q, lam, g, ph, ap, cp, dt, faci, faciq, m, lo, rt);
        cell_vjp_finish<NB, QP, NQ>(lam_in, g, ap, c, dt, lo, rt);
    }
    if (ap.partial) block_partial(dcoll, ap.partial);
}

// For small grids (latency-bound): a block of 4 waves takes 64 cells and each
// wave one part (volume + sources, x, y, z faces); the face waves pass their
// sums through LDS to the volume wave, which finishes the cells.  Four times
// the parallelism of one thread per cell, at the same total work.
template <int NB, int QP, int NQ, int RM>
__global__ __launch_bounds__(256) void k_stage_vjp_split(const double* __restrict__ q,
                                                         const double* __restrict__ lam, double* __restrict__ lam_in,
                                                         Grid g, Phys ph, AdjParams ap) {
    static_assert(split_fits<NB, NQ>(), "partial sums exceed LDS");
    __shared__ double part_sum[3][NQ * NB][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long c = (long)blockIdx.x * 64 + lane;
    double dcoll = 0.;
    const double dt = *ap.dt;
    const double faci = (ap.ivis == 1 && !ap.rhs_only) ? 1. / (1. + dt * ap.coll) : 1.0;
    const double faciq = (ap.ivis == 1 && !ap.rhs_only) ? 1. / (1. + dt * ap.coll_q) : 1.0;
    double m[NQ][NB], lo[NQ][NB], rt[NQ][NB];
#pragma unroll
    for (int f = 0; f < NQ; f++)
#pragma unroll
        for (int r = 0; r < NB; r++) rt[f][r] = 0.;
    const bool valid = c < g.N;
    fz::CellPos cp;
    if (valid) {
        cp = cell_pos(g, c);
        fz::load_modes<NB, NQ>(q, g, c, m);
        load_lp<NB, NQ>(lam, g, c, ap.b, faci, faciq, lo);
        dcoll = cell_vjp_parts<NB, QP, NQ, RM>(1 << wave, q, lam, g, ph, ap, cp, dt, faci, faciq, m, lo, rt);
    }
    if (wave > 0) {
#pragma unroll
        for (int f = 0; f < NQ; f++)
#pragma unroll
            for (int r = 0; r < NB; r++) part_sum[wave - 1][f * NB + r][lane] = rt[f][r];
    }
    __syncthreads();
    if (wave == 0 && valid) {
#pragma unroll
        for (int f = 0; f < NQ; f++)
#pragma unroll
            for (int r = 0; r < NB; r++)
                rt[f][r] += part_sum[0][f * NB + r][lane] + part_sum[1][f * NB + r][lane] +
                            part_sum[2][f * NB + r][lane];
        cell_vjp_finish<NB, QP, NQ>(lam_in, g, ap, c, dt, lo, rt);
    }
    if (ap.partial) block_partial(dcoll, ap.partial);
}

// ------------------------------------------------- 13-moment transpose ----
// The transpose of a 13-moment stage on a grid periodic in every direction,
// split as the forward stage is (fz::k_face13 + k_cell13), so that no lane
// holds a whole cell: the face kernel writes, per face point, the cotangents
// that the face flux sends to the traces of its two cells; the cell kernel,
// one field per wave, adds the volume transpose and gathers those face
// cotangents.  The sum over the terms is the fused k_stage_vjp's, evaluated
// in another order.

// One wave per face point over 64 consecutive faces of direction DIR (as
// fz::k_face13): phase 1 stages both sides' trace coefficients (face_coeffs
// of the limited modes q) and lifted-cotangent coefficients (lift_coeffs of
// b*lam through the relaxation factors) in LDS; phase 2 forms the face
// flux's cotangent mu at point p, reverses the Riemann solver and the
// physical fluxes (face_vjp), and writes the left cell's trace cotangent to
// bql and the right cell's to bqr, laid out as the forward face fluxes.
template <int NB, int QP, int RM, int DIR>
__global__ __launch_bounds__(QP * QP * 64) void k_face13_vjp(const double* __restrict__ q,
                                                            const double* __restrict__ lam, double* __restrict__ bql,
                                                            double* __restrict__ bqr, Grid g, Phys ph, AdjParams ap) {
    constexpr int NQ = 13, NF = QP * QP, NK = fz::face_nk<NB>();
    __shared__ double ks[2][NQ][NK][64];
    __shared__ double ls[2][NQ][NK][64];
    const double dt = *ap.dt;
    const double faci = (ap.ivis == 1 && !ap.rhs_only) ? 1. / (1. + dt * ap.coll) : 1.0;
    const double faciq = (ap.ivis == 1 && !ap.rhs_only) ? 1. / (1. + dt * ap.coll_q) : 1.0;
    const int ni = g.nx + (DIR == 0), nj = g.ny + (DIR == 1);
    const long nfaces = (long)ni * nj * (g.nz + (DIR == 2));
    const long f0 = (long)xcd_block() * 64;
    for (int it = (int)threadIdx.x; it < 2 * NQ * 64; it += NF * 64) {
        const int l = it & 63, side = (it >> 6) / NQ, fld = (it >> 6) % NQ;
        const long fid = f0 + l < nfaces ? f0 + l : nfaces - 1;
        int ci = (int)(fid % ni), cj = (int)((fid / ni) % nj), ck = (int)(fid / ((long)ni * nj));
        int& a = DIR == 0 ? ci : DIR == 1 ? cj : ck;
        const int n_ax = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
        if (side == 0) a = a == 0 ? n_ax - 1 : a - 1;
        else if (a == n_ax) a = 0;
        const long cell = (long)ci + (long)g.nx * ((long)cj + (long)g.ny * ck);
        const double fac = fld >= T_QX ? faciq : (fld >= PXX ? faci : 1.0);
        double m1[1][NB], l1[1][NB], K1[1][NK], L1[1][NK];
#pragma unroll
        for (int r = 0; r < NB; r++) {
            const long o = ((long)r * NQ + fld) * g.N + cell;
            m1[0][r] = q[o];
            l1[0][r] = ap.b * lam[o] * fac;
        }
        if (side == 0) {
            fz::face_coeffs<NB, 1, DIR, 1, NK>(m1, K1);
            lift_coeffs<NB, 1, DIR, 1, NK>(l1, L1);
        } else {
            fz::face_coeffs<NB, 1, DIR, 0, NK>(m1, K1);
            lift_coeffs<NB, 1, DIR, 0, NK>(l1, L1);
        }
#pragma unroll
        for (int k = 0; k < NK; k++) {
            ks[side][fld][k][l] = K1[0][k];
            ls[side][fld][k][l] = L1[0][k];
        }
    }
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int l = (int)(threadIdx.x & 63);
    const long fid = f0 + l;
    if (fid >= nfaces) return;
    const auto& FP = fz::face_poly<NB>();
    const double dxi_d = DIR == 0 ? g.dxi : DIR == 1 ? g.dyi : g.dzi;
    double QL[NQ], QR[NQ], mu[NQ];
#pragma unroll
    for (int f = 0; f < NQ; f++) {
        double sl = ks[0][f][0][l], sr = ks[1][f][0][l];
        double tl = ls[0][f][0][l] * FP.wb[p][0], tr = ls[1][f][0][l] * FP.wb[p][0];
#pragma unroll
        for (int k = 1; k < NK; k++) {
            sl += ks[0][f][k][l] * FP.b[p][k];
            sr += ks[1][f][k][l] * FP.b[p][k];
            tl += ls[0][f][k][l] * FP.wb[p][k];
            tr += ls[1][f][k][l] * FP.wb[p][k];
        }
        QL[f] = sl;
        QR[f] = sr;
        mu[f] = dxi_d * (tl + tr);
    }
    double bL[NQ], bR[NQ];
#pragma unroll
    for (int f = 0; f < NQ; f++) { bL[f] = 0.; bR[f] = 0.; }
    face_vjp<NQ, RM, DIR>(ph, QL, QR, mu, bL, bR);
#pragma unroll
    for (int f = 0; f < NQ; f++) {
        const long o = ((long)f * NF + p) * nfaces + fid;
        bql[o] = bL[f];
        bqr[o] = bR[f];
    }
}

// per-block partial of d(lam.out)/dcoll for workgroups of 13 waves
__device__ __forceinline__ void block_partial13(double v, double* partial) {
    __shared__ double red[CELL13_WAVES];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.;
        for (int w = 0; w < CELL13_WAVES; w++) s += red[w];
        partial[blockIdx.x] = s;
    }
}

// One field per wave over 64 consecutive cells (as k_cell13).  The faces
// come first (their loads beside the modes'): bfc^T of the cotangents written
// by k_face13_vjp.  The volume transpose then takes one row of the rule (QP
// points) per exchange: each wave writes its field's point value and the
// cotangents of its three point fluxes (the transposes of k_cell's volume
// sums); one wave per point reverses the point's physical fluxes (flux_vjp)
// and shares the state cotangent, which each wave projects onto its field's
// modes.  Then the sources, the stage map (lo + dt*rt, the accumulated terms)
// and the frozen limiter, as cell_vjp_finish; d/dcoll as cell_vjp_parts.
template <int NB, int QP, int RM>
__global__ __launch_bounds__(CELL13_WAVES * 64) void k_cell13_vjp(
    const double* __restrict__ q, const double* __restrict__ lam, double* __restrict__ lam_in,
    const double* __restrict__ bql0, const double* __restrict__ bql1, const double* __restrict__ bql2,
    const double* __restrict__ bqr0, const double* __restrict__ bqr1, const double* __restrict__ bqr2,
    const Tables* __restrict__ T, Grid g, Phys ph, AdjParams ap) {
    constexpr int NQ = 13, NF = QP * QP, NPG = QP * QP * QP, NCH = QP;
    __shared__ double qp_s[NQ][NCH][64];
    __shared__ double bf_s[NQ][3][NCH][64];
    __shared__ double bq_s[NQ][NCH][64];
    const int f = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const long c = (long)xcd_block() * 64 + lane;
    const bool live = c < g.N;
    const long cc = live ? c : g.N - 1;  // lanes past the grid compute a copy of the last cell and store nothing
    const int i = (int)(cc % g.nx), j = (int)((cc / g.nx) % g.ny), k = (int)(cc / ((long)g.nx * g.ny));
    const double dt = *ap.dt;
    const double faci = (ap.ivis == 1 && !ap.rhs_only) ? 1. / (1. + dt * ap.coll) : 1.0;
    const double faciq = (ap.ivis == 1 && !ap.rhs_only) ? 1. / (1. + dt * ap.coll_q) : 1.0;
    const double fac = f >= T_QX ? faciq : (f >= PXX ? faci : 1.0);
    const double dxi = g.dxi, dyi = g.dyi, dzi = g.dzi;
    double mf[NB], lof[NB], rt[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) {
        const long o = qoff(g, cc, f, r);
        mf[r] = q[o];
        lof[r] = ap.b * lam[o] * fac;
        rt[r] = 0.;
    }
    // ---- faces first, their loads beside those of the modes: the cell is
    // the right side of its low faces and the left side of its high faces ----
    {
        const long nfl[3] = {(long)(g.nx + 1) * g.ny * g.nz, (long)g.nx * (g.ny + 1) * g.nz,
                             (long)g.nx * g.ny * (g.nz + 1)};
        const long fm[3] = {(long)i + (long)(g.nx + 1) * ((long)j + (long)g.ny * k),
                            (long)i + (long)g.nx * ((long)j + (long)(g.ny + 1) * k),
                            (long)i + (long)g.nx * ((long)j + (long)g.ny * k)};
        const long fp[3] = {fm[0] + 1, fm[1] + g.nx, fm[2] + (long)g.nx * g.ny};
        const double* const bqr[3] = {bqr0, bqr1, bqr2};
        const double* const bql[3] = {bql0, bql1, bql2};
#pragma unroll
        for (int d = 0; d < 3; d++) {
#pragma unroll 1
            for (int p = 0; p < NF; p++) {
                const long o = (long)f * NF + p;
                const double vR = bqr[d][o * nfl[d] + fm[d]];  // this cell is the right side of its low face
                const double vL = bql[d][o * nfl[d] + fp[d]];  // and the left side of its high face
#pragma unroll
                for (int r = 0; r < NB; r++) rt[r] += T->bfc[2 * d][p][r] * vR + T->bfc[2 * d + 1][p][r] * vL;
            }
        }
    }
    // ---- volume transpose; every wave reaches every barrier ----
#pragma unroll 1
    for (int p0 = 0; p0 < NPG; p0 += NCH) {
#pragma unroll
        for (int qq = 0; qq < NCH; qq++) {
            const int p = p0 + qq;
            const double* b = T->bint[p];
            double v = 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) v += b[r] * mf[r];
            qp_s[f][qq][lane] = v;
            // cotangents of F_x, F_y, F_z of field f at point p: the
            // transposes of k_cell's volume sums acc[*] and intr[*]
            double bx = T->cb[KX] * lof[KX], by = T->cb[KY] * lof[KY], bz = T->cb[KZ] * lof[KZ];
            if constexpr (NB > 4) {
                by += T->cb[KYZ] * b[KZ] * lof[KYZ];
                bz += T->cb[KYZ] * b[KY] * lof[KYZ];
                bx += T->cb[KZX] * b[KZ] * lof[KZX];
                bz += T->cb[KZX] * b[KX] * lof[KZX];
                bx += T->cb[KXY] * b[KY] * lof[KXY];
                by += T->cb[KXY] * b[KX] * lof[KXY];
                if constexpr (NB == 10) {
                    bx += T->cb[KXX] * 3. * b[KX] * lof[KXX];
                    by += T->cb[KYY] * 3. * b[KY] * lof[KYY];
                    bz += T->cb[KZZ] * 3. * b[KZ] * lof[KZZ];
                } else {
                    bx += T->cb[KXYZ] * b[KYZ] * lof[KXYZ];
                    by += T->cb[KXYZ] * b[KZX] * lof[KXYZ];
                    bz += T->cb[KXYZ] * b[KXY] * lof[KXYZ];
                }
            }
            const double w = 0.25 * T->w3[p];
            bf_s[f][0][qq][lane] = dxi * w * bx;
            bf_s[f][1][qq][lane] = dyi * w * by;
            bf_s[f][2][qq][lane] = dzi * w * bz;
        }
        __syncthreads();
        if (f < NCH) {
            double Qp[NQ], bx[NQ], by[NQ], bz[NQ], bQ[NQ];
#pragma unroll
            for (int e = 0; e < NQ; e++) {
                Qp[e] = qp_s[e][f][lane];
                bx[e] = bf_s[e][0][f][lane];
                by[e] = bf_s[e][1][f][lane];
                bz[e] = bf_s[e][2][f][lane];
                bQ[e] = 0.;
            }
            const pr::Prim st = pr::prim<NQ>(ph, Qp);
            flux_vjp<NQ, 0>(ph, Qp, st, bx, bQ);
            flux_vjp<NQ, 1>(ph, Qp, st, by, bQ);
            flux_vjp<NQ, 2>(ph, Qp, st, bz, bQ);
#pragma unroll
            for (int e = 0; e < NQ; e++) bq_s[e][f][lane] = bQ[e];
        }
        __syncthreads();
#pragma unroll
        for (int qq = 0; qq < NCH; qq++) {
            const double bQf = bq_s[f][qq][lane];
            const double* b = T->bint[p0 + qq];
#pragma unroll
            for (int r = 0; r < NB; r++) rt[r] += b[r] * bQf;
        }
        __syncthreads();  // the next row overwrites the exchange arrays
    }
    double dcoll = 0.;
    if (live) {
        // ---- sources (as cell_vjp_parts) ----
        if (ap.ivis == 0) {
            if (f >= PXX && f <= T_PYZ && ap.coll != 0.)
#pragma unroll
                for (int r = 0; r < NB; r++) {
                    rt[r] += -ap.coll * lof[r];
                    dcoll += -dt * lof[r] * mf[r];
                }
            if (f >= T_QX && ap.coll_q != 0.)
#pragma unroll
                for (int r = 0; r < NB; r++) rt[r] += -ap.coll_q * lof[r];
        }
        if (ap.ivis == 1 && ap.coll != 0. && ap.out && f >= PXX && f <= T_PYZ) {
            // qn is read only when its weight a is nonzero (callers pass no
            // qn for the first stage)
            const bool with_qn = ap.a != 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) {
                const long o = qoff(g, c, f, r);
                const double bv = with_qn ? ap.out[o] - ap.a * ap.qn[o] : ap.out[o];
                dcoll += -dt * faci * lam[o] * bv;
            }
        }
        if (ap.bf[0] != 0. || ap.bf[1] != 0. || ap.bf[2] != 0.) {
            // lo of the Euler fields (relaxation factor 1)
            if (f == RH)
#pragma unroll
                for (int r = 0; r < NB; r++)
                    rt[r] += ap.bf[0] * ap.b * lam[qoff(g, c, MX, r)] + ap.bf[1] * ap.b * lam[qoff(g, c, MY, r)] +
                             ap.bf[2] * ap.b * lam[qoff(g, c, MZ, r)];
            if (f >= MX && f <= MZ)
#pragma unroll
                for (int r = 0; r < NB; r++) rt[r] += ap.bf[f - MX] * ap.b * lam[qoff(g, c, EN, r)];
        }
        // ---- finish (as cell_vjp_finish) ----
        if (ap.rhs_only) {
#pragma unroll
            for (int r = 0; r < NB; r++) lam_in[qoff(g, c, f, r)] = rt[r];
        } else {
#pragma unroll
            for (int r = 0; r < NB; r++) rt[r] = lof[r] + dt * rt[r];
#pragma unroll
            for (int t = 0; t < 2; t++)
                if (ap.acc[t])
#pragma unroll
                    for (int r = 0; r < NB; r++) rt[r] += ap.w[t] * ap.acc[t][qoff(g, c, f, r)];
            if (ap.q_raw && f <= EN) {
                double raw[5][NB];
#pragma unroll
                for (int e = 0; e < 5; e++)
#pragma unroll
                    for (int r = 0; r < NB; r++) raw[e][r] = ap.q_raw[qoff(g, c, e, r)];
                const LimFactors F = limit_factors<NB, QP, 5>(raw, ap.lp);
                if (F.zero) {
#pragma unroll
                    for (int r = 1; r < NB; r++) rt[r] = 0.0;
                    if (F.floor && f == RH) rt[0] = 0.0;
                } else {
                    const double th = f == RH ? F.th_rho * F.th_p : F.th_p;
#pragma unroll
                    for (int r = 1; r < NB; r++) rt[r] *= th;
                }
            }
#pragma unroll
            for (int r = 0; r < NB; r++) lam_in[qoff(g, c, f, r)] = rt[r];
        }
    }
    if (ap.partial) block_partial13(dcoll, ap.partial);
}

// adds the block partials to *acc (one block, fixed order)
static __global__ __launch_bounds__(256) void k_sum_partials(const double* __restrict__ partial, int n,
                                                             double* __restrict__ acc) {
    __shared__ double red[4];
    double s = 0.;
    for (int b = threadIdx.x; b < n; b += blockDim.x) s += partial[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
        *acc += t;
    }
}

}  // namespace adj
}  // namespace hd
