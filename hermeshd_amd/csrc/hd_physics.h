// hd_physics.h -- point-wise physics of HERMESHD's DG operator, written for
// one quadrature point per call so that device kernels can keep everything in
// registers.  The arithmetic follows the reference routines cited below; the
// per-face loops of the reference are independent across points, so a
// per-point function computes the same values.
//
//   flux_point  : flux_calc_pnts_r  (src/flux.f90:23-193), ivis in {0,1}, ieos 1 (2 with TAIT)
//   hllc_point  : flux_hllc         (src/flux.f90:466-647), ibatten=0 (Davis speeds), ieos 1/2
//   roe_point   : flux_roe          (src/ces/10_moment/dg_3D_hydro_lqc_SEAN.f90:1614-1789)
//   cfcal       : cfcal             (src/flux.f90:1029-1072)
//   flux_point10, cfcal10 : the ivis = 2 branches of flux_calc_pnts_r and
//                 cfcal (src/flux.f90:64-66,89-117,134-155,172-191,1038-1044)
#pragma once
#include <hip/hip_runtime.h>

#define HD_DEV __host__ __device__ __forceinline__

namespace hd {

enum : int { RH = 0, MX = 1, MY = 2, MZ = 3, EN = 4, PXX = 5, PYY = 6, PZZ = 7, PXY = 8, PXZ = 9, PYZ = 10 };
// 13-moment model (nQ 13; no counterpart in the reference): the traceless
// stress without pzz (= -pxx - pyy), then the heat flux
enum : int { T_PXX = 5, T_PYY = 6, T_PXY = 7, T_PXZ = 8, T_PYZ = 9, T_QX = 10, T_QY = 11, T_QZ = 12 };
enum : int { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_WALL = 2, BC_NOSLIP = 3, BC_INFLOW = 4 };
enum : int { RIEMANN_CENTRAL = 0, RIEMANN_HLLC = 1, RIEMANN_ROE = 2 };

struct Phys {
    double gamma, gm1;       // aindex, aindm1
    double P_floor, rh_floor;
    double colvis, c2d3cv, c4d3cv;
    int ivis;                // 2: full 10-moment equations (table-driven kernels only)
    int ieos;                // 2: Murnaghan-Tait water (table-driven kernels only)
    double P1, Pbase;        // P = P1*(rho^7.2 - 1) + Pbase + P_ideal when ieos = 2
    double colq;             // 13 moments: coll_q * conductivity (heat-flux flux col_q*T)
};

// Murnaghan-Tait pressure on top of the ideal-gas one (src/flux.f90:64,526,
// src/params.f90:97-105)
HD_DEV double tait_pressure(const Phys& ph, double dn, double P) {
    return ph.P1 * (pow(dn, 7.2) - 1.) + ph.Pbase + P;
}

// DIR: 0=x, 1=y, 2=z.  S: stochastic stress xx,yy,zz,xy,xz,yz or nullptr.
// TAIT: Murnaghan-Tait water pressure (ieos 2; table-driven kernels only).
// 13-moment flux (nQ 13; no counterpart in the reference): the stress rows of
// the relaxation model (src/flux.f90:101-106,138-143,175-180) for the five
// stored components, pzz = -pxx - pyy, and the Maxwell-Cattaneo heat flux,
// d_t q_i + d_i(colq T) = -coll_q q_i with T = P/rho; q_DIR enters the energy
// flux.
template <int DIR>
HD_DEV void flux_point13(const Phys& ph, const double* Q, double* F) {
    const double dn = Q[RH], M_x = Q[MX], M_y = Q[MY], M_z = Q[MZ], Ener = Q[EN];
    const double T_xx = Q[T_PXX], T_yy = Q[T_PYY], T_zz = -(T_xx + T_yy);
    const double T_xy = Q[T_PXY], T_xz = Q[T_PXZ], T_yz = Q[T_PYZ];
    const double dni = 1. / dn;
    const double vx = M_x * dni, vy = M_y * dni, vz = M_z * dni;
    const double smsq = M_x * M_x + M_y * M_y + M_z * M_z;
    double P = ph.gm1 * (Ener - 0.5 * dni * smsq);
    if (P < ph.P_floor) P = ph.P_floor;
    const double ct = ph.colq * (P * dni);
    if constexpr (DIR == 0) {
        F[RH] = M_x;
        F[MX] = M_x * vx + P + T_xx;
        F[MY] = M_y * vx + T_xy;
        F[MZ] = M_z * vx + T_xz;
        F[EN] = (Ener + P) * vx + T_xx * vx + T_xy * vy + T_xz * vz + Q[T_QX];
        F[T_PXX] = ph.c4d3cv * vx; F[T_PYY] = -ph.c2d3cv * vx;
        F[T_PXY] = ph.colvis * vy; F[T_PXZ] = ph.colvis * vz; F[T_PYZ] = 0.;
        F[T_QX] = ct; F[T_QY] = 0.; F[T_QZ] = 0.;
    } else if constexpr (DIR == 1) {
        F[RH] = M_y;
        F[MX] = M_x * vy + T_xy;
        F[MY] = M_y * vy + P + T_yy;
        F[MZ] = M_z * vy + T_yz;
        F[EN] = (Ener + P) * vy + T_yy * vy + T_xy * vx + T_yz * vz + Q[T_QY];
        F[T_PXX] = -ph.c2d3cv * vy; F[T_PYY] = ph.c4d3cv * vy;
        F[T_PXY] = ph.colvis * vx; F[T_PXZ] = 0.; F[T_PYZ] = ph.colvis * vz;
        F[T_QX] = 0.; F[T_QY] = ct; F[T_QZ] = 0.;
    } else {
        F[RH] = M_z;
        F[MX] = M_x * vz + T_xz;
        F[MY] = M_y * vz + T_yz;
        F[MZ] = M_z * vz + P + T_zz;
        F[EN] = (Ener + P) * vz + T_zz * vz + T_xz * vx + T_yz * vy + Q[T_QZ];
        F[T_PXX] = -ph.c2d3cv * vz; F[T_PYY] = -ph.c2d3cv * vz;
        F[T_PXY] = 0.; F[T_PXZ] = ph.colvis * vx; F[T_PYZ] = ph.colvis * vy;
        F[T_QX] = 0.; F[T_QY] = 0.; F[T_QZ] = ct;
    }
}

template <int NQ, int DIR, bool TAIT = false>
HD_DEV void flux_point(const Phys& ph, const double* Q, double* F, const double* S) {
    if constexpr (NQ == 13) {
        flux_point13<DIR>(ph, Q, F);
        return;
    }
    const double dn = Q[RH], M_x = Q[MX], M_y = Q[MY], M_z = Q[MZ], Ener = Q[EN];
    double P_xx = 0., P_yy = 0., P_zz = 0., P_xy = 0., P_xz = 0., P_yz = 0.;
    if constexpr (NQ == 11) {
        P_xx = Q[PXX]; P_yy = Q[PYY]; P_zz = Q[PZZ]; P_xy = Q[PXY]; P_xz = Q[PXZ]; P_yz = Q[PYZ];
    }
    double Sxx = 0., Syy = 0., Szz = 0., Sxy = 0., Sxz = 0., Syz = 0.;
    if (S) { Sxx = S[0]; Syy = S[1]; Szz = S[2]; Sxy = S[3]; Sxz = S[4]; Syz = S[5]; }
    const double dni = 1. / dn;
    const double vx = M_x * dni, vy = M_y * dni, vz = M_z * dni;
    const double smsq = M_x * M_x + M_y * M_y + M_z * M_z;
    double P = ph.gm1 * (Ener - 0.5 * dni * smsq);
    if constexpr (TAIT) P = tait_pressure(ph, dn, P);
    if (P < ph.P_floor) P = ph.P_floor;
    if constexpr (DIR == 0) {
        F[RH] = M_x;
        F[MX] = M_x * vx + P + P_xx - Sxx;
        F[MY] = M_y * vx + P_xy - Sxy;
        F[MZ] = M_z * vx + P_xz - Sxz;
        F[EN] = (Ener + P) * vx + (P_xx - Sxx) * vx + (P_xy - Sxy) * vy + (P_xz - Sxz) * vz;
        if constexpr (NQ == 11) {
            F[PXX] = ph.c4d3cv * vx; F[PYY] = -ph.c2d3cv * vx; F[PZZ] = -ph.c2d3cv * vx;
            F[PXY] = ph.colvis * vy; F[PXZ] = ph.colvis * vz; F[PYZ] = 0.;
        }
    } else if constexpr (DIR == 1) {
        F[RH] = M_y;
        F[MX] = M_x * vy + P_xy - Sxy;
        F[MY] = M_y * vy + P + P_yy - Syy;
        F[MZ] = M_z * vy + P_yz - Syz;
        F[EN] = (Ener + P) * vy + (P_yy - Syy) * vy + (P_xy - Sxy) * vx + (P_yz - Syz) * vz;
        if constexpr (NQ == 11) {
            F[PXX] = -ph.c2d3cv * vy; F[PYY] = ph.c4d3cv * vy; F[PZZ] = -ph.c2d3cv * vy;
            F[PXY] = ph.colvis * vx; F[PXZ] = 0.; F[PYZ] = ph.colvis * vz;
        }
    } else {
        F[RH] = M_z;
        F[MX] = M_x * vz + P_xz - Sxz;
        F[MY] = M_y * vz + P_yz - Syz;
        F[MZ] = M_z * vz + P + P_zz - Szz;
        F[EN] = (Ener + P) * vz + (P_zz - Szz) * vz + (P_xz - Sxz) * vx + (P_yz - Syz) * vy;
        if constexpr (NQ == 11) {
            F[PXX] = -ph.c2d3cv * vz; F[PYY] = -ph.c2d3cv * vz; F[PZZ] = ph.c4d3cv * vz;
            F[PXY] = 0.; F[PXZ] = ph.colvis * vx; F[PYZ] = ph.colvis * vy;
        }
    }
}

// Full 10-moment flux (ivis = 2): the six stress fields hold the second
// moments E_ij = rho u_i u_j + P_ij.  The momentum flux is E_{i,DIR} with the
// isotropic 10-moment pressure P_10 = tr(P)/3 swapped for the equation-of-state
// pressure P_5; the second moments carry the flux of the 10-moment equations,
// written out per direction as in the reference.
template <int DIR>
HD_DEV void flux_point10(const Phys& ph, const double* Q, double* F) {
    const double dn = Q[RH], M_x = Q[MX], M_y = Q[MY], M_z = Q[MZ], Ener = Q[EN];
    const double P_xx = Q[PXX], P_yy = Q[PYY], P_zz = Q[PZZ], P_xy = Q[PXY], P_xz = Q[PXZ], P_yz = Q[PYZ];
    const double dni = 1. / dn;
    const double vx = M_x * dni, vy = M_y * dni, vz = M_z * dni;
    const double smsq = M_x * M_x + M_y * M_y + M_z * M_z;
    double P = ph.gm1 * (Ener - 0.5 * dni * smsq);
    if (ph.ieos == 2) P = tait_pressure(ph, dn, P);
    if (P < ph.P_floor) P = ph.P_floor;
    double P_10 = (1. / 3.) * (P_xx + P_yy + P_zz - dni * smsq);
    if (P_10 < ph.P_floor) P_10 = ph.P_floor;
    const double P_5 = P;
    if constexpr (DIR == 0) {
        F[RH] = M_x;
        F[MX] = P_xx - P_10 + P_5; F[MY] = P_xy; F[MZ] = P_xz;
        F[EN] = (Ener + P) * vx + P_xx * vx + P_xy * vy + P_xz * vz;
        F[PXX] = vx * (3 * P_xx - 2 * vx * M_x);
        F[PYY] = 2 * vy * (P_xy - vy * M_x) + vx * P_yy;
        F[PZZ] = 2 * vz * (P_xz - vz * M_x) + vx * P_zz;
        F[PXY] = 2 * vx * (P_xy - vy * M_x) + vy * P_xx;
        F[PXZ] = 2 * vx * (P_xz - vz * M_x) + vz * P_xx;
        F[PYZ] = vx * P_yz + vy * P_xz + vz * P_xy - 2 * vy * vz * M_x;
    } else if constexpr (DIR == 1) {
        F[RH] = M_y;
        F[MX] = P_xy; F[MY] = P_yy - P_10 + P_5; F[MZ] = P_yz;
        F[EN] = (Ener + P) * vy + P_yy * vy + P_xy * vx + P_yz * vz;
        F[PXX] = 2 * vx * (P_xy - vx * M_y) + vy * P_xx;
        F[PYY] = vy * (3 * P_yy - 2 * vy * M_y);
        F[PZZ] = 2 * vz * (P_yz - vz * M_y) + vy * P_zz;
        F[PXY] = 2 * vy * (P_xy - vx * M_y) + vx * P_yy;
        F[PXZ] = vx * P_yz + vy * P_xz + vz * P_xy - 2 * vx * vz * M_y;
        F[PYZ] = 2 * vy * (P_yz - vz * M_y) + vz * P_yy;
    } else {
        F[RH] = M_z;
        F[MX] = P_xz; F[MY] = P_yz; F[MZ] = P_zz - P_10 + P_5;
        F[EN] = (Ener + P) * vz + P_zz * vz + P_xz * vx + P_yz * vy;
        F[PXX] = 2 * vx * (P_xz - vx * M_z) + vz * P_xx;
        F[PYY] = 2 * vy * (P_yz - vy * M_z) + vz * P_yy;
        F[PZZ] = vz * (3 * P_zz - 2 * vz * M_z);
        F[PXY] = vx * P_yz + vy * P_xz + vz * P_xy - 2 * vx * vy * M_z;
        F[PXZ] = 2 * vz * (P_xz - vx * M_z) + vx * P_zz;
        F[PYZ] = 2 * vz * (P_yz - vy * M_z) + vy * P_zz;
    }
}

// freezing speed of the 10-moment equations: |u_dir| + sqrt(gamma*P/rho) with
// P = tr(P_ij)/3 from the second moments (cfcal, ivis = 2)
template <int DIR>
HD_DEV double cfcal10(const Phys& ph, const double* Q) {
    const double dn = Q[RH], dni = 1. / dn;
    const double vx = Q[MX] * dni, vy = Q[MY] * dni, vz = Q[MZ] * dni;
    const double P = (1. / 3.) * (Q[PXX] + Q[PYY] + Q[PZZ] - dn * (vx * vx + vy * vy + vz * vz));
    const double cs = ph.ieos == 2 ? sqrt(7.2 * ph.P1 * pow(dn, 6.2)) : sqrt(ph.gamma * P * dni);
    const double v = DIR == 0 ? vx : DIR == 1 ? vy : vz;
    return fabs(v) + cs;
}

template <int DIR>
HD_DEV void dir_fields(int& ipar, int& ip1, int& ip2) {
    if constexpr (DIR == 0) { ipar = MX; ip1 = MY; ip2 = MZ; }
    else if constexpr (DIR == 1) { ipar = MY; ip1 = MZ; ip2 = MX; }
    else { ipar = MZ; ip1 = MX; ip2 = MY; }
}

// HLLC at one face point.  QL/QR: conserved traces (>= 5 fields), FL/FR their
// physical fluxes (same direction).  Writes out[0..4] in field order
// rh,mx,my,mz,en.  On entry out[] holds the fallback (central) flux, which is
// kept when no branch of Eq. (26) applies (only possible with NaN speeds).
template <int DIR>
HD_DEV void hllc_point(const Phys& ph, const double* QL, const double* QR, const double* FL,
                       const double* FR, double* out, double rho_i_last = 0.) {
    int ipar, ip1, ip2;
    dir_fields<DIR>(ipar, ip1, ip2);
    const double gm1 = ph.gm1, g = ph.gamma;
    // primitive states (src/flux.f90:510-520)
    const double rL = QL[RH], rR = QR[RH];
    const double riL = 1.0 / rL, riR = 1.0 / rR;
    const double uL = QL[ipar] * riL, vL = QL[ip1] * riL, wL = QL[ip2] * riL;
    const double uR = QR[ipar] * riR, vR = QR[ip1] * riR, wR = QR[ip2] * riR;
    double pL = gm1 * (QL[EN] - 0.5 * rL * (uL * uL + vL * vL + wL * wL));
    double pR = gm1 * (QR[EN] - 0.5 * rR * (uR * uR + vR * vR + wR * wR));
    // Davis wave speeds (src/flux.f90:522-585)
    double cL, cR;
    if (ph.ieos == 2) {
        // water: the reference's sound speed takes rho_i as its loop leaves it,
        // 1/rho of the face's last right state (rho_i_last) at every point
        pL = tait_pressure(ph, rL, pL);
        pR = tait_pressure(ph, rR, pR);
        cL = sqrt(7.2 * ph.P1 * pow(rL, 6.2) + pL * rho_i_last);
        cR = sqrt(7.2 * ph.P1 * pow(rR, 6.2) + pR * rho_i_last);
    } else {
        cL = sqrt(g * pL / rL);
        cR = sqrt(g * pR / rR);
    }
    const double SL = fmin(uL - cL, uR - cR);
    const double SR = fmax(uR + cR, uL + cL);
    // contact speed and star pressure, Eqs. (34),(36) of Batten 1997
    double sm_num;
    {
        // no contraction: the x-high outflow ghost makes this exactly zero
#pragma clang fp contract(off)
        sm_num = rR * uR * (SR - uR) - rL * uL * (SL - uL);
        sm_num = sm_num + pL - pR;
    }
    double sm_den = rR * (SR - uR) - rL * (SL - uL);
    if (sm_den == 0.0) sm_den = ph.rh_floor;
    const double SM = sm_num / sm_den;
    const double pstar = rL * (uL - SL) * (uL - SM) + pL;
    // Eq. (26) in the reference's sequential-if order: the last true branch wins
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
        // star state of the selected side, Eqs. (35)-(40), and F* = F + S (U* - U), Eq. (29)
        const bool left = (sel == 1);
        const double* Qk = left ? QL : QR;
        const double* Fk = left ? FL : FR;
        const double Sk = left ? SL : SR;
        const double uk = left ? uL : uR, rk = left ? rL : rR, pk = left ? pL : pR;
        double den = Sk - SM;
        if (den == 0.0) den = ph.rh_floor;
        const double slrm_i = 1.0 / den;
        const double sq = Sk - uk;
        const double s0 = rk * sq * slrm_i;
        const double s1 = (sq * Qk[ipar] + pstar - pk) * slrm_i;
        const double s2 = sq * Qk[ip1] * slrm_i;
        const double s3 = sq * Qk[ip2] * slrm_i;
        const double s4 = (sq * Qk[EN] - pk * uk + pstar * SM) * slrm_i;
        out[RH] = Fk[RH] + Sk * (s0 - Qk[RH]);
        out[ipar] = Fk[ipar] + Sk * (s1 - Qk[ipar]);
        out[ip1] = Fk[ip1] + Sk * (s2 - Qk[ip1]);
        out[ip2] = Fk[ip2] + Sk * (s3 - Qk[ip2]);
        out[EN] = Fk[EN] + Sk * (s4 - Qk[EN]);
    }
}

// freezing speed |u_dir| + c (cfcal, ivis in {0,1}, ieos = 1)
template <int DIR>
HD_DEV double cfcal(const Phys& ph, const double* Q) {
    const double dn = Q[RH], dni = 1. / dn;
    const double vx = Q[MX] * dni, vy = Q[MY] * dni, vz = Q[MZ] * dni;
    const double P = ph.gm1 * (Q[EN] - 0.5 * dn * (vx * vx + vy * vy + vz * vz));
    const double cs = sqrt(ph.gamma * P * dni);
    const double v = DIR == 0 ? vx : DIR == 1 ? vy : vz;
    return fabs(v) + cs;
}

// Roe flux (Eulderink & Mellema) for rh..en at one point; when the Roe sound
// speed squared is not positive, local Lax-Friedrichs with cfcal speeds, as the
// legacy flux_cal does (dg_3D_hydro_lqc_SEAN.f90:1255-1289).
template <int DIR, int NQ = 5>
HD_DEV void roe_point(const Phys& ph, const double* QL, const double* QR, const double* FL,
                      const double* FR, double* out) {
    int ipar, ip1, ip2;
    dir_fields<DIR>(ipar, ip1, ip2);
    const double gm1 = ph.gm1, cp = ph.gamma / gm1;
    double sk[2], sw[2][4];
    for (int s = 0; s < 2; s++) {
        const double* Q = s ? QR : QL;
        const double dnii = 1. / Q[RH];
        const double a = Q[ipar] * dnii, b = Q[ip1] * dnii, c = Q[ip2] * dnii;
        const double vsq = a * a + b * b + c * c;
        const double Pri = gm1 * (Q[EN] - 0.5 * Q[RH] * vsq);
        sk[s] = sqrt(Q[RH]);
        sw[s][0] = sk[s] * a; sw[s][1] = sk[s] * b; sw[s][2] = sk[s] * c;
        sw[s][3] = 0.5 * sk[s] * (vsq + cp * Pri / Q[RH]);
    }
    const double Qri = 1. / QL[RH];
    const double delrho = QR[RH] * Qri - 1.;
    const double dele = (QR[EN] - QL[EN]) * Qri;
    const double dm1 = (QR[ipar] - QL[ipar]) * Qri;
    const double dm2 = (QR[ip1] - QL[ip1]) * Qri;
    const double dm3 = (QR[ip2] - QL[ip2]) * Qri;
    const double dsk = 0.5 * (sk[1] + sk[0]);
    const double vc0 = 0.5 * (sw[1][0] + sw[0][0]) / dsk;
    const double vc1 = 0.5 * (sw[1][1] + sw[0][1]) / dsk;
    const double vc2 = 0.5 * (sw[1][2] + sw[0][2]) / dsk;
    const double hc = 0.5 * (sw[1][3] + sw[0][3]) / dsk;
    const double vcsq = vc0 * vc0 + vc1 * vc1 + vc2 * vc2;
    const double asq = gm1 * (hc - 0.5 * vcsq);
    if (!(asq > 0.0)) {
        double cw;
        if constexpr (NQ >= 11)
            cw = ph.ivis == 2 ? fmax(cfcal10<DIR>(ph, QL), cfcal10<DIR>(ph, QR))
                              : fmax(cfcal<DIR>(ph, QL), cfcal<DIR>(ph, QR));
        else
            cw = fmax(cfcal<DIR>(ph, QL), cfcal<DIR>(ph, QR));
        const int idx[5] = {RH, MX, MY, MZ, EN};
        for (int e = 0; e < 5; e++) {
            const int f = idx[e];
            out[f] = 0.5 * (FL[f] + FR[f]) - 0.5 * cw * (QR[f] - QL[f]);
        }
        return;
    }
    const double sc = sqrt(asq);
    const double l0 = fabs(vc0 - sc), l1 = fabs(vc0 + sc), l2 = fabs(vc0);
    const double ea2i = gm1 / asq;
    const double vcdel = vc0 * dm1 + vc1 * dm2 + vc2 * dm3 - dele;
    const double a9e = 0.5 * ea2i * (0.5 * vcsq * delrho - vcdel);
    const double sk9 = 0.5 * (dm1 - vc0 * delrho) / sc;
    const double a0 = a9e - sk9, a1 = a9e + sk9;
    const double a2 = ea2i * ((hc - vcsq) * delrho + vcdel);
    const double a3 = dm2 - vc1 * delrho, a4 = dm3 - vc2 * delrho;
    // Delta F = sum_j a_j |lambda_j| e_j with e_j columns over rows
    // (rho, energy, parallel, perp1, perp2)
    const double t0 = a0 * l0, t1 = a1 * l1, t2 = a2 * l2, t3 = a3 * l2, t4 = a4 * l2;
    double d_rho = t0 + t1 + t2;
    double d_en = t0 * (hc - sc * vc0) + t1 * (hc + sc * vc0) + t2 * (0.5 * vcsq) + t3 * vc1 + t4 * vc2;
    double d_par = t0 * (vc0 - sc) + t1 * (vc0 + sc) + t2 * vc0;
    double d_p1 = t0 * vc1 + t1 * vc1 + t2 * vc1 + t3;
    double d_p2 = t0 * vc2 + t1 * vc2 + t2 * vc2 + t4;
    const double rl = QL[RH];
    out[RH] = 0.5 * (FL[RH] + FR[RH] - d_rho * rl);
    out[EN] = 0.5 * (FL[EN] + FR[EN] - d_en * rl);
    out[ipar] = 0.5 * (FL[ipar] + FR[ipar] - d_par * rl);
    out[ip1] = 0.5 * (FL[ip1] + FR[ip1] - d_p1 * rl);
    out[ip2] = 0.5 * (FL[ip2] + FR[ip2] - d_p2 * rl);
}


// Reverse mode of roe_point for the Euler fields: mu (cotangent of out[rh..en])
// -> bQL, bQR (+=, Euler fields) and bFL, bFR (+=, the cotangents of the two
// physical fluxes, Euler fields).  The forward intermediates are recomputed
// and swept backwards line by line; the branches (Roe or the Lax-Friedrichs
// fallback, the larger freezing speed, the signs inside |lambda|) are those of
// the forward evaluation at the same states.
template <int DIR>
HD_DEV void roe_point_vjp(const Phys& ph, const double* QL, const double* QR, const double* mu, double* bQL,
                          double* bQR, double* bFL, double* bFR) {
    int ipar, ip1, ip2;
    dir_fields<DIR>(ipar, ip1, ip2);
    const int idx[5] = {RH, MX, MY, MZ, EN};
    for (int e = 0; e < 5; e++) {
        bFL[idx[e]] += 0.5 * mu[idx[e]];
        bFR[idx[e]] += 0.5 * mu[idx[e]];
    }
    const double gm1 = ph.gm1, cp = ph.gamma / gm1;
    // ---- forward, as roe_point
    double sk[2], sw[2][4], av[2], bv[2], cv[2], vsq[2], Pri[2], dnii[2];
    for (int s = 0; s < 2; s++) {
        const double* Q = s ? QR : QL;
        dnii[s] = 1. / Q[RH];
        av[s] = Q[ipar] * dnii[s]; bv[s] = Q[ip1] * dnii[s]; cv[s] = Q[ip2] * dnii[s];
        vsq[s] = av[s] * av[s] + bv[s] * bv[s] + cv[s] * cv[s];
        Pri[s] = gm1 * (Q[EN] - 0.5 * Q[RH] * vsq[s]);
        sk[s] = sqrt(Q[RH]);
        sw[s][0] = sk[s] * av[s]; sw[s][1] = sk[s] * bv[s]; sw[s][2] = sk[s] * cv[s];
        sw[s][3] = 0.5 * sk[s] * (vsq[s] + cp * Pri[s] / Q[RH]);
    }
    const double Qri = 1. / QL[RH];
    const double delrho = QR[RH] * Qri - 1.;
    const double dele = (QR[EN] - QL[EN]) * Qri;
    const double dm1 = (QR[ipar] - QL[ipar]) * Qri;
    const double dm2 = (QR[ip1] - QL[ip1]) * Qri;
    const double dm3 = (QR[ip2] - QL[ip2]) * Qri;
    const double dsk = 0.5 * (sk[1] + sk[0]);
    const double vc0 = 0.5 * (sw[1][0] + sw[0][0]) / dsk;
    const double vc1 = 0.5 * (sw[1][1] + sw[0][1]) / dsk;
    const double vc2 = 0.5 * (sw[1][2] + sw[0][2]) / dsk;
    const double hc = 0.5 * (sw[1][3] + sw[0][3]) / dsk;
    const double vcsq = vc0 * vc0 + vc1 * vc1 + vc2 * vc2;
    const double asq = gm1 * (hc - 0.5 * vcsq);
    if (!(asq > 0.0)) {
        // Lax-Friedrichs fallback: out = 0.5 (FL + FR) - 0.5 cw (QR - QL),
        // cw the larger of the two freezing speeds |v_DIR| + sqrt(gamma P / rho)
        const double cwL = cfcal<DIR>(ph, QL), cwR = cfcal<DIR>(ph, QR);
        // fmax semantics: a NaN speed (negative pressure) loses to the other side
        const bool left = (cwL >= cwR) || (cwR != cwR);
        const double cw = left ? cwL : cwR;
        double bcw = 0.;
        for (int e = 0; e < 5; e++) {
            const int f = idx[e];
            bQR[f] += -0.5 * cw * mu[f];
            bQL[f] += 0.5 * cw * mu[f];
            bcw += -0.5 * mu[f] * (QR[f] - QL[f]);
        }
        // cfcal: P = gm1 (E - 0.5 rho |v|^2), cs = sqrt(gamma P / rho), |v_DIR| + cs
        const double* Q = left ? QL : QR;
        double* bQ = left ? bQL : bQR;
        const double rho = Q[RH], ri = 1. / rho;
        const double vx = Q[MX] * ri, vy = Q[MY] * ri, vz = Q[MZ] * ri;
        const double P = gm1 * (Q[EN] - 0.5 * rho * (vx * vx + vy * vy + vz * vz));
        const double cs = sqrt(ph.gamma * P * ri);
        const double vd = DIR == 0 ? vx : DIR == 1 ? vy : vz;
        double bv3[3] = {0., 0., 0.};
        bv3[DIR] += bcw * (vd > 0. ? 1. : (vd < 0. ? -1. : 0.));
        const double bcs = bcw;
        double bP = 0., bri = 0.;
        if (cs > 0.) {
            bP += bcs * ph.gamma * ri / (2. * cs);
            bri += bcs * ph.gamma * P / (2. * cs);
        }
        bQ[EN] += gm1 * bP;
        double brho = -0.5 * gm1 * (vx * vx + vy * vy + vz * vz) * bP;
        const double vvv[3] = {vx, vy, vz};
        for (int a = 0; a < 3; a++) bv3[a] += -gm1 * rho * vvv[a] * bP;
        for (int a = 0; a < 3; a++) {
            bQ[MX + a] += bv3[a] * ri;
            bri += bv3[a] * Q[MX + a];
        }
        bQ[RH] += brho - ri * ri * bri;
        return;
    }
    const double sc = sqrt(asq);
    const double x0 = vc0 - sc, x1 = vc0 + sc;
    const double l0 = fabs(x0), l1 = fabs(x1), l2 = fabs(vc0);
    const double ea2i = gm1 / asq;
    const double vcdel = vc0 * dm1 + vc1 * dm2 + vc2 * dm3 - dele;
    const double Z = 0.5 * vcsq * delrho - vcdel;
    const double a9e = 0.5 * ea2i * Z;
    const double Y = dm1 - vc0 * delrho;
    const double sk9 = 0.5 * Y / sc;
    const double a0 = a9e - sk9, a1 = a9e + sk9;
    const double X = (hc - vcsq) * delrho + vcdel;
    const double a2 = ea2i * X;
    const double a3 = dm2 - vc1 * delrho, a4 = dm3 - vc2 * delrho;
    const double t0 = a0 * l0, t1 = a1 * l1, t2 = a2 * l2, t3 = a3 * l2, t4 = a4 * l2;
    const double S = t0 + t1 + t2;
    const double d_rho = S;
    const double d_en = t0 * (hc - sc * vc0) + t1 * (hc + sc * vc0) + t2 * (0.5 * vcsq) + t3 * vc1 + t4 * vc2;
    const double d_par = t0 * (vc0 - sc) + t1 * (vc0 + sc) + t2 * vc0;
    const double d_p1 = t0 * vc1 + t1 * vc1 + t2 * vc1 + t3;
    const double d_p2 = t0 * vc2 + t1 * vc2 + t2 * vc2 + t4;
    const double rl = QL[RH];
    // ---- reverse
    // out = 0.5 (FL + FR - d * rl)
    const double bd_rho = -0.5 * mu[RH] * rl, bd_en = -0.5 * mu[EN] * rl, bd_par = -0.5 * mu[ipar] * rl;
    const double bd_p1 = -0.5 * mu[ip1] * rl, bd_p2 = -0.5 * mu[ip2] * rl;
    bQL[RH] += -0.5 * (mu[RH] * d_rho + mu[EN] * d_en + mu[ipar] * d_par + mu[ip1] * d_p1 + mu[ip2] * d_p2);
    double bt0 = 0., bt1 = 0., bt2 = 0., bt3 = 0., bt4 = 0.;
    double bhc = 0., bsc = 0., bvc0 = 0., bvc1 = 0., bvc2 = 0., bvcsq = 0.;
    // d_rho = t0 + t1 + t2
    bt0 += bd_rho; bt1 += bd_rho; bt2 += bd_rho;
    // d_en
    bt0 += bd_en * (hc - sc * vc0);
    bt1 += bd_en * (hc + sc * vc0);
    bt2 += bd_en * (0.5 * vcsq);
    bt3 += bd_en * vc1;
    bt4 += bd_en * vc2;
    bhc += bd_en * (t0 + t1);
    bsc += bd_en * (-t0 * vc0 + t1 * vc0);
    bvc0 += bd_en * (-t0 * sc + t1 * sc);
    bvcsq += bd_en * 0.5 * t2;
    bvc1 += bd_en * t3;
    bvc2 += bd_en * t4;
    // d_par
    bt0 += bd_par * (vc0 - sc);
    bt1 += bd_par * (vc0 + sc);
    bt2 += bd_par * vc0;
    bvc0 += bd_par * S;
    bsc += bd_par * (t1 - t0);
    // d_p1, d_p2
    bt0 += bd_p1 * vc1 + bd_p2 * vc2;
    bt1 += bd_p1 * vc1 + bd_p2 * vc2;
    bt2 += bd_p1 * vc1 + bd_p2 * vc2;
    bvc1 += bd_p1 * S;
    bvc2 += bd_p2 * S;
    bt3 += bd_p1;
    bt4 += bd_p2;
    // t_j = a_j * l_j
    const double ba0 = bt0 * l0, ba1 = bt1 * l1, ba2 = bt2 * l2, ba3 = bt3 * l2, ba4 = bt4 * l2;
    const double bl0 = bt0 * a0, bl1 = bt1 * a1, bl2 = bt2 * a2 + bt3 * a3 + bt4 * a4;
    double bdm1 = 0., bdm2 = 0., bdm3 = 0., bdele = 0., bdelrho = 0., bvcdel = 0., bea2i = 0.;
    // a4 = dm3 - vc2 delrho; a3 = dm2 - vc1 delrho
    bdm3 += ba4; bvc2 += -ba4 * delrho; bdelrho += -ba4 * vc2;
    bdm2 += ba3; bvc1 += -ba3 * delrho; bdelrho += -ba3 * vc1;
    // a2 = ea2i X, X = (hc - vcsq) delrho + vcdel
    bea2i += ba2 * X;
    {
        const double bX = ba2 * ea2i;
        bhc += bX * delrho;
        bvcsq += -bX * delrho;
        bdelrho += bX * (hc - vcsq);
        bvcdel += bX;
    }
    // a0 = a9e - sk9, a1 = a9e + sk9
    const double ba9e = ba0 + ba1, bsk9 = ba1 - ba0;
    // sk9 = 0.5 Y / sc, Y = dm1 - vc0 delrho
    {
        const double bY = bsk9 * 0.5 / sc;
        bsc += -bsk9 * sk9 / sc;
        bdm1 += bY;
        bvc0 += -bY * delrho;
        bdelrho += -bY * vc0;
    }
    // a9e = 0.5 ea2i Z, Z = 0.5 vcsq delrho - vcdel
    bea2i += ba9e * 0.5 * Z;
    {
        const double bZ = ba9e * 0.5 * ea2i;
        bvcsq += bZ * 0.5 * delrho;
        bdelrho += bZ * 0.5 * vcsq;
        bvcdel += -bZ;
    }
    // vcdel = vc0 dm1 + vc1 dm2 + vc2 dm3 - dele
    bvc0 += bvcdel * dm1; bdm1 += bvcdel * vc0;
    bvc1 += bvcdel * dm2; bdm2 += bvcdel * vc1;
    bvc2 += bvcdel * dm3; bdm3 += bvcdel * vc2;
    bdele += -bvcdel;
    // ea2i = gm1 / asq
    double basq = -bea2i * ea2i / asq;
    // |x| with the forward signs
    {
        const double s0 = x0 > 0. ? 1. : (x0 < 0. ? -1. : 0.);
        const double s1 = x1 > 0. ? 1. : (x1 < 0. ? -1. : 0.);
        const double s2 = vc0 > 0. ? 1. : (vc0 < 0. ? -1. : 0.);
        bvc0 += bl0 * s0 + bl1 * s1 + bl2 * s2;
        bsc += -bl0 * s0 + bl1 * s1;
    }
    // sc = sqrt(asq); asq = gm1 (hc - 0.5 vcsq)
    basq += bsc / (2. * sc);
    bhc += basq * gm1;
    bvcsq += -0.5 * gm1 * basq;
    // vcsq
    bvc0 += 2. * vc0 * bvcsq; bvc1 += 2. * vc1 * bvcsq; bvc2 += 2. * vc2 * bvcsq;
    // vc_k = 0.5 (swR_k + swL_k) / dsk, hc likewise with sw_3
    double bsw[4];
    const double vcs[4] = {vc0, vc1, vc2, hc};
    const double bvcs[4] = {bvc0, bvc1, bvc2, bhc};
    double bdsk = 0.;
    for (int k = 0; k < 4; k++) {
        bsw[k] = 0.5 * bvcs[k] / dsk;
        bdsk += -bvcs[k] * vcs[k] / dsk;
    }
    // dm's, dele, delrho and Qri = 1/QL[RH]
    double bQri = 0.;
    bQR[ipar] += bdm1 * Qri; bQL[ipar] += -bdm1 * Qri; bQri += bdm1 * (QR[ipar] - QL[ipar]);
    bQR[ip1] += bdm2 * Qri; bQL[ip1] += -bdm2 * Qri; bQri += bdm2 * (QR[ip1] - QL[ip1]);
    bQR[ip2] += bdm3 * Qri; bQL[ip2] += -bdm3 * Qri; bQri += bdm3 * (QR[ip2] - QL[ip2]);
    bQR[EN] += bdele * Qri; bQL[EN] += -bdele * Qri; bQri += bdele * (QR[EN] - QL[EN]);
    bQR[RH] += bdelrho * Qri; bQri += bdelrho * QR[RH];
    bQL[RH] += -bQri * Qri * Qri;
    // per side: dsk = 0.5 (skR + skL); sw_0..2 = sk (a, b, c);
    // sw_3 = 0.5 sk (vsq + cp Pri / rho); Pri = gm1 (E - 0.5 rho vsq); sk = sqrt(rho)
    for (int s = 0; s < 2; s++) {
        const double* Q = s ? QR : QL;
        double* bQ = s ? bQR : bQL;
        const double rho = Q[RH];
        const double W = vsq[s] + cp * Pri[s] / rho;
        double bsk = 0.5 * bdsk + bsw[0] * av[s] + bsw[1] * bv[s] + bsw[2] * cv[s] + bsw[3] * 0.5 * W;
        double ba = bsw[0] * sk[s], bb = bsw[1] * sk[s], bc = bsw[2] * sk[s];
        const double bW = bsw[3] * 0.5 * sk[s];
        double bvsq = bW;
        double bPri = bW * cp / rho;
        double brho = -bW * cp * Pri[s] / (rho * rho);
        brho += bsk / (2. * sk[s]);
        bQ[EN] += gm1 * bPri;
        brho += -0.5 * gm1 * vsq[s] * bPri;
        bvsq += -0.5 * gm1 * rho * bPri;
        ba += 2. * av[s] * bvsq; bb += 2. * bv[s] * bvsq; bc += 2. * cv[s] * bvsq;
        bQ[ipar] += ba * dnii[s]; bQ[ip1] += bb * dnii[s]; bQ[ip2] += bc * dnii[s];
        const double bdnii = ba * Q[ipar] + bb * Q[ip1] + bc * Q[ip2];
        brho += -bdnii * dnii[s] * dnii[s];
        bQ[RH] += brho;
    }
}

}  // namespace hd
