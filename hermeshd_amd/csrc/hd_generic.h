// hd_generic.h -- table-driven HIP kernels of the DG operator, valid for every
// supported (nbasis, iquad, nQ, Riemann solver).  One thread per cell or per
// face; loads and stores are coalesced because consecutive threads take
// consecutive i (the reference's fastest index).  These kernels keep the
// reference's arithmetic order (sums over modes and points in order) and serve
// every configuration; the fused tile kernels specialise the hot
// configurations.
#pragma once
#include "hd_physics.h"
#include "hd_tables.h"

namespace hd {

struct Grid {
    int nx, ny, nz, nQ;
    long N;                 // cells
    double dxi, dyi, dzi;
    int bc[6];
};

struct LimParams {
    double rh_floor, epsP, gm1;
};

// Workgroup b runs on XCD b mod 8.  Giving each XCD a contiguous eighth of the
// cells keeps a cell's y and z neighbours (nx and nx*ny cells away) in the
// same XCD's L2 instead of another's.
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned b = blockIdx.x, G = gridDim.x;
    return (G & 7u) == 0 ? (b & 7u) * (G >> 3) + (b >> 3) : b;
}

__device__ __forceinline__ long qoff(const Grid& g, long cell, int f, int r) {
    return ((long)r * g.nQ + f) * g.N + cell;
}

// ---------------------------------------------------------------------------
// limiter (src/flux.f90:1077-1164): one thread per cell, in place; cells the
// limiter leaves unchanged are not written.
// ---------------------------------------------------------------------------
template <int NB, int QP>
__global__ __launch_bounds__(256) void k_limiter(double* __restrict__ q, const Tables* __restrict__ T,
                                                 Grid g, LimParams lp) {
    constexpr int NPGE = 6 * QP * QP;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    double m[5][NB];
#pragma unroll
    for (int f = 0; f < 5; f++)
#pragma unroll
        for (int r = 0; r < NB; r++) m[f][r] = q[qoff(g, c, f, r)];
    bool changed = false;
    if (m[RH][0] < lp.rh_floor) {
#pragma unroll
        for (int r = 1; r < NB; r++)
#pragma unroll
            for (int f = 0; f < 5; f++) m[f][r] = 0.0;
        m[RH][0] = lp.rh_floor;
        changed = true;
    } else {
        const double epsi = lp.rh_floor;
        // face points in bf_faces order: xm, xp, ym, yp, zm, zp
        double Qrhmin = 0.;
        for (int p = 0; p < NPGE; p++) {
            const double* b = T->bfc[p / (QP * QP)][p % (QP * QP)];
            double s = 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) s += b[r] * m[RH][r];
            if (p == 0 || s < Qrhmin) Qrhmin = s;
        }
        if (Qrhmin < epsi) {
            double theta = (epsi - m[RH][0]) / (Qrhmin - m[RH][0]);
            if (theta > 1.) theta = 1.;
            if (theta < 0.) theta = 0.;
#pragma unroll
            for (int r = 1; r < NB; r++) m[RH][r] = fabs(theta) * m[RH][r];
            changed = true;
        }
        const double Pave =
            lp.gm1 * (m[EN][0] - 0.5 * (m[MX][0] * m[MX][0] + m[MY][0] * m[MY][0] + m[MZ][0] * m[MZ][0]) / m[RH][0]);
        if (Pave < lp.epsP) {
#pragma unroll
            for (int r = 1; r < NB; r++)
#pragma unroll
                for (int f = 0; f < 5; f++) m[f][r] = 0.0;
            changed = true;
        } else {
            double theta = 1.;
            for (int p = 0; p < NPGE; p++) {
                const double* b = T->bfc[p / (QP * QP)][p % (QP * QP)];
                double e[5];
#pragma unroll
                for (int f = 0; f < 5; f++) {
                    double s = 0.;
#pragma unroll
                    for (int r = 0; r < NB; r++) s += b[r] * m[f][r];
                    e[f] = s;
                }
                const double dni = 1. / e[RH];
                const double P = lp.gm1 * (e[EN] - 0.5 * dni * (e[MX] * e[MX] + e[MY] * e[MY] + e[MZ] * e[MZ]));
                if (P < lp.epsP && Pave != P) theta = fmin(theta, (Pave - lp.epsP) / (Pave - P));
            }
            if (theta > 1.) theta = 1.;
            if (theta < 0.) theta = 0.;
            if (theta != 1.) {
#pragma unroll
                for (int r = 1; r < NB; r++)
#pragma unroll
                    for (int f = 0; f < 5; f++) m[f][r] = theta * m[f][r];
                changed = true;
            }
        }
    }
    if (changed) {
#pragma unroll
        for (int f = 0; f < 5; f++)
#pragma unroll
            for (int r = 0; r < NB; r++) q[qoff(g, c, f, r)] = m[f][r];
    }
}

// ---------------------------------------------------------------------------
// boundary states (src/boundary.f90:40-338): ext from the interior trace of
// the boundary cell.  side: 0 lo, 1 hi;  dir 0/1/2.
// ---------------------------------------------------------------------------
template <int NF, int NQ>
__device__ __forceinline__ void apply_bc(int kind, int dir, int side, const double (&in)[NF][NQ],
                                         double (&ex)[NF][NQ], const double* inflow = nullptr, long jk = 0,
                                         long plane = 0) {
    const int normal = MX + dir;
#pragma unroll
    for (int p = 0; p < NF; p++)
#pragma unroll
        for (int f = 0; f < NQ; f++) ex[p][f] = in[p][f];
    if (kind == BC_INFLOW) {
        // x_bc_inflow (src/boundary.f90:571-592); inflow is (ny,nz,nface,nQ), jk = j + ny*k
#pragma unroll
        for (int p = 0; p < NF; p++) {
            ex[p][RH] = inflow[jk + plane * (p + (long)NF * RH)];
            ex[p][MX] = inflow[jk + plane * (p + (long)NF * MX)];
            ex[p][MY] = 0.0;
            ex[p][MZ] = 0.0;
            ex[p][EN] = inflow[jk + plane * (p + (long)NF * EN)];
        }
    } else if (kind == BC_NOSLIP) {
#pragma unroll
        for (int p = 0; p < NF; p++) ex[p][MX] = ex[p][MY] = ex[p][MZ] = 0.0;
    } else if (kind == BC_WALL) {
#pragma unroll
        for (int p = 0; p < NF; p++) ex[p][normal] = 0.0;
    } else if (kind == BC_OUTFLOW) {
        if (dir == 0 && side == 1) {
            // xhibc_outflow (src/boundary.f90:246-282) after copying the interior
#pragma unroll
            for (int p = 0; p < NF; p++)
                if (ex[p][MX] > 0.0) ex[p][MX] = -in[p][MX];
        } else if (side == 0) {
            double v = ex[0][normal];
#pragma unroll
            for (int p = 1; p < NF; p++) v = fmax(v, ex[p][normal]);
            if (v > 0.0)
#pragma unroll
                for (int p = 0; p < NF; p++) ex[p][normal] = 0.0;
        } else {
            double v = ex[0][normal];
#pragma unroll
            for (int p = 1; p < NF; p++) v = fmin(v, ex[p][normal]);
            if (v < 0.0)
#pragma unroll
                for (int p = 0; p < NF; p++) ex[p][normal] = 0.0;
        }
    }
}

// physical flux of the configured equations at one point
template <int NQ, int DIR>
__device__ __forceinline__ void gflux(const Phys& ph, const double* Q, double* F) {
    if constexpr (NQ == 11) {
        if (ph.ivis == 2) {
            flux_point10<DIR>(ph, Q, F);
            return;
        }
    }
    if (ph.ieos == 2)
        flux_point<NQ, DIR, true>(ph, Q, F, nullptr);
    else
        flux_point<NQ, DIR>(ph, Q, F, nullptr);
}

template <int NB, int NF, int NQ>
__device__ __forceinline__ void load_trace(const double* __restrict__ q, const Grid& g, long cell,
                                           const double (*bf)[10], double (&out)[NF][NQ]) {
    double m[NB];
#pragma unroll
    for (int f = 0; f < NQ; f++) {
#pragma unroll
        for (int r = 0; r < NB; r++) m[r] = q[qoff(g, cell, f, r)];
#pragma unroll
        for (int p = 0; p < NF; p++) {
            double s = 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) s += bf[p][r] * m[r];
            out[p][f] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// face fluxes (calc_flux_x/y/z, src/flux.f90:197-450): one thread per face of
// direction DIR, including both domain faces; output layout
// fl[((f*NF + p) * nfaces) + face], face = i + ni*(j + nj*k).
// ---------------------------------------------------------------------------
template <int NB, int QP, int NQ, int RM, int DIR>
__global__ __launch_bounds__(128) void k_face_flux(const double* __restrict__ q, double* __restrict__ fl,
                                                   const Tables* __restrict__ T, Grid g, Phys ph,
                                                   const double* __restrict__ inflow) {
    constexpr int NF = QP * QP;
    const int ni = g.nx + (DIR == 0), nj = g.ny + (DIR == 1), nk = g.nz + (DIR == 2);
    const long nfaces = (long)ni * nj * nk;
    const long fid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (fid >= nfaces) return;
    const int i = (int)(fid % ni);
    const int j = (int)((fid / ni) % nj);
    const int k = (int)(fid / ((long)ni * nj));
    const int n_ax = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
    const int a = DIR == 0 ? i : DIR == 1 ? j : k;  // face index along the axis
    const double(*bp)[10] = T->bfc[2 * DIR + 1];
    const double(*bm)[10] = T->bfc[2 * DIR];
    auto cell_at = [&](int ii, int jj, int kk) -> long { return (long)ii + (long)g.nx * ((long)jj + (long)g.ny * kk); };
    const int il = i - (DIR == 0), jl = j - (DIR == 1), kl = k - (DIR == 2);
    double QL[NF][NQ], QR[NF][NQ];
    const bool periodic = g.bc[2 * DIR] == BC_PERIODIC;
    if (a > 0) load_trace<NB, NF, NQ>(q, g, cell_at(il, jl, kl), bp, QL);
    if (a < n_ax) load_trace<NB, NF, NQ>(q, g, cell_at(i, j, k), bm, QR);
    if (a == 0) {
        if (periodic) {
            const int ii = DIR == 0 ? g.nx - 1 : i, jj = DIR == 1 ? g.ny - 1 : j, kk = DIR == 2 ? g.nz - 1 : k;
            load_trace<NB, NF, NQ>(q, g, cell_at(ii, jj, kk), bp, QL);
        } else {
            apply_bc<NF, NQ>(g.bc[2 * DIR], DIR, 0, QR, QL, inflow, (long)j + (long)g.ny * k, (long)g.ny * g.nz);
        }
    }
    if (a == n_ax) {
        if (periodic) {
            const int ii = DIR == 0 ? 0 : i, jj = DIR == 1 ? 0 : j, kk = DIR == 2 ? 0 : k;
            load_trace<NB, NF, NQ>(q, g, cell_at(ii, jj, kk), bm, QR);
        } else {
            apply_bc<NF, NQ>(g.bc[2 * DIR + 1], DIR, 1, QL, QR);
        }
    }
#pragma unroll
    for (int p = 0; p < NF; p++) {
        double FL[NQ], FR[NQ], out[NQ];
        gflux<NQ, DIR>(ph, QL[p], FL);
        gflux<NQ, DIR>(ph, QR[p], FR);
#pragma unroll
        for (int f = 0; f < NQ; f++) out[f] = 0.5 * (FL[f] + FR[f]);
        if constexpr (NQ == 11) {
            // ivis 2: local Lax-Friedrichs dissipation on every field with the
            // 10-moment freezing speeds (src/flux.f90:243-263)
            if (ph.ivis == 2) {
                const double cw = fmax(cfcal10<DIR>(ph, QL[p]), cfcal10<DIR>(ph, QR[p]));
#pragma unroll
                for (int f = 0; f < NQ; f++) out[f] = 0.5 * (FL[f] + FR[f]) - 0.5 * cw * (QR[p][f] - QL[p][f]);
            }
        }
        if constexpr (RM == RIEMANN_HLLC) hllc_point<DIR>(ph, QL[p], QR[p], FL, FR, out, 1.0 / QR[NF - 1][RH]);
        if constexpr (RM == RIEMANN_ROE) roe_point<DIR, NQ>(ph, QL[p], QR[p], FL, FR, out);
#pragma unroll
        for (int f = 0; f < NQ; f++) fl[((long)f * NF + p) * nfaces + fid] = out[f];
    }
}

struct StageParams {
    const double* dt;     // device scalar
    double coll;          // relaxation (ivis 1: implicit factor; ivis 0: explicit source)
    double coll_q;        // 13 moments: relaxation rate of the heat flux
    int ivis;
    int combine;          // 0: none, 1: 0.5*(qn+qs), 2: 0.75*qn+0.25*qs, 3: qn/3+2qs/3
    int mode;             // 0: stage update, 1: write rhs = -glflux+source
    double bf[3];         // body force
    int* err;             // NaN flag
    unsigned long long* bad;
    int limit_out;        // fused kernels: limit the output (it is the next stage's input)
    LimParams lp;
    const float* noise[4];   // stochastic stress of the step (single precision): x, y, z face families, volume points
    long nfb;                // face lattice size (nx+1)(ny+1)(nz+1)
    const double* inflow;    // x-low inflow exterior state (ny,nz,nface,nQ), Fortran order
    int rows_zfast;          // row-shared stage kernel: workgroup order with z fastest (else y fastest)
};

// ---------------------------------------------------------------------------
// cell kernel: innerintegral2 (src/flux.f90:751-928), glflux assembly
// (src/flux.f90:958-978), source_calc (src/sources.f90:12-84) and
// advance_time_level (src/integrator.f90:125-252), then the SSP combination of
// RK2/RK3 (src/integrator.f90:62-86).
// ---------------------------------------------------------------------------
template <int NB, int QP, int NQ>
__global__ __launch_bounds__(128) void k_cell(const double* __restrict__ qin, const double* __restrict__ qn,
                                              double* __restrict__ qout, const double* __restrict__ flx,
                                              const double* __restrict__ fly, const double* __restrict__ flz,
                                              const Tables* __restrict__ T, Grid g, Phys ph, StageParams sp) {
    constexpr int NF = QP * QP, NPG = QP * QP * QP;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / ((long)g.nx * g.ny));
    double m[NQ][NB];
#pragma unroll
    for (int f = 0; f < NQ; f++)
#pragma unroll
        for (int r = 0; r < NB; r++) m[f][r] = qin[qoff(g, c, f, r)];
    // ---- volume integral ----
    constexpr int NS = (NB > 4) ? 12 : 3;
    double acc[NQ][NS];
#pragma unroll
    for (int f = 0; f < NQ; f++)
#pragma unroll
        for (int s = 0; s < NS; s++) acc[f][s] = 0.;
    for (int p = 0; p < NPG; p++) {
        const double* b = T->bint[p];
        double Qp[NQ], Fx[NQ], Fy[NQ], Fz[NQ];
#pragma unroll
        for (int f = 0; f < NQ; f++) {
            double s = 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) s += b[r] * m[f][r];
            Qp[f] = s;
        }
        gflux<NQ, 0>(ph, Qp, Fx);
        gflux<NQ, 1>(ph, Qp, Fy);
        gflux<NQ, 2>(ph, Qp, Fz);
        const double w = T->w3[p];
#pragma unroll
        for (int f = 0; f < NQ; f++) {
            acc[f][0] += w * Fx[f];
            acc[f][1] += w * Fy[f];
            acc[f][2] += w * Fz[f];
            if constexpr (NB > 4) {
                acc[f][3] += w * b[KZ] * Fy[f];
                acc[f][4] += w * b[KY] * Fz[f];
                acc[f][5] += w * b[KZ] * Fx[f];
                acc[f][6] += w * b[KX] * Fz[f];
                acc[f][7] += w * b[KY] * Fx[f];
                acc[f][8] += w * b[KX] * Fy[f];
                if constexpr (NB == 10) {
                    acc[f][9] += w * 3. * b[KX] * Fx[f];
                    acc[f][10] += w * 3. * b[KY] * Fy[f];
                    acc[f][11] += w * 3. * b[KZ] * Fz[f];
                } else {
                    acc[f][9] += w * b[KYZ] * Fx[f];
                    acc[f][10] += w * b[KZX] * Fy[f];
                    acc[f][11] += w * b[KXY] * Fz[f];
                }
            }
        }
    }
    const double dxi = g.dxi, dyi = g.dyi, dzi = g.dzi;
    // ---- face flux locations ----
    const long nfx = (long)(g.nx + 1) * g.ny * g.nz, nfy = (long)g.nx * (g.ny + 1) * g.nz,
               nfz = (long)g.nx * g.ny * (g.nz + 1);
    const long fxm = (long)i + (long)(g.nx + 1) * ((long)j + (long)g.ny * k);
    const long fym = (long)i + (long)g.nx * ((long)j + (long)(g.ny + 1) * k);
    const long fzm = (long)i + (long)g.nx * ((long)j + (long)g.ny * k);
    const long fxp = fxm + 1, fyp = fym + g.nx, fzp = fzm + (long)g.nx * g.ny;
    const double dt = *sp.dt;
    const double faci = 1. / (1. + dt * sp.coll);
    const double faciq = 1. / (1. + dt * sp.coll_q);
    bool bad = false;
    // NaN check of mode 1, SSP combination, store
    auto emit = [&](int f, int r, double v) {
        const long o = qoff(g, c, f, r);
        if (r == 0 && v != v) bad = true;
        if (sp.combine == 1) v = 0.5 * (qn[o] + v);
        else if (sp.combine == 2) v = 0.75 * qn[o] + 0.25 * v;
        else if (sp.combine == 3) v = (1. / 3.) * qn[o] + (2. / 3.) * v;
        qout[o] = v;
    };
    // ivis 2: glflux and source of the diagonal second moments, kept for the
    // trace-coupled implicit update
    constexpr int ND = NQ == 11 ? 3 : 1;
    double dglf[ND][NB], dsrc[ND][NB];
#pragma unroll
    for (int f = 0; f < NQ; f++) {
        double xm[NF], xp[NF], ym[NF], yp[NF], zm[NF], zp[NF];
#pragma unroll
        for (int p = 0; p < NF; p++) {
            xm[p] = flx[((long)f * NF + p) * nfx + fxm]; xp[p] = flx[((long)f * NF + p) * nfx + fxp];
            ym[p] = fly[((long)f * NF + p) * nfy + fym]; yp[p] = fly[((long)f * NF + p) * nfy + fyp];
            zm[p] = flz[((long)f * NF + p) * nfz + fzm]; zp[p] = flz[((long)f * NF + p) * nfz + fzp];
        }
        double intr[NB];
        intr[0] = 0.;
        intr[KX] = 0.25 * T->cb[KX] * dxi * acc[f][0];
        intr[KY] = 0.25 * T->cb[KY] * dyi * acc[f][1];
        intr[KZ] = 0.25 * T->cb[KZ] * dzi * acc[f][2];
        if constexpr (NB > 4) {
            intr[KYZ] = 0.25 * T->cb[KYZ] * (dyi * acc[f][3] + dzi * acc[f][4]);
            intr[KZX] = 0.25 * T->cb[KZX] * (dxi * acc[f][5] + dzi * acc[f][6]);
            intr[KXY] = 0.25 * T->cb[KXY] * (dxi * acc[f][7] + dyi * acc[f][8]);
            if constexpr (NB == 10) {
                intr[KXX] = 0.25 * T->cb[KXX] * (dxi * acc[f][9]);
                intr[KYY] = 0.25 * T->cb[KYY] * (dyi * acc[f][10]);
                intr[KZZ] = 0.25 * T->cb[KZZ] * (dzi * acc[f][11]);
            } else {
                intr[KXYZ] = 0.25 * T->cb[KXYZ] * (dxi * acc[f][9] + dyi * acc[f][10] + dzi * acc[f][11]);
            }
        }
        double glf[NB];
        {
            double sx = 0., sy = 0., sz = 0.;
#pragma unroll
            for (int p = 0; p < NF; p++) {
                sx += T->w2[p] * (xp[p] - xm[p]);
                sy += T->w2[p] * (yp[p] - ym[p]);
                sz += T->w2[p] * (zp[p] - zm[p]);
            }
            glf[0] = 0.25 * (dxi * sx + dyi * sy + dzi * sz);
        }
#pragma unroll
        for (int r = 1; r < NB; r++) {
            double sx = 0., sy = 0., sz = 0.;
#pragma unroll
            for (int p = 0; p < NF; p++) {
                sx += T->w2[p] * (T->bfc[F_XP][p][r] * xp[p] - T->bfc[F_XM][p][r] * xm[p]);
                sy += T->w2[p] * (T->bfc[F_YP][p][r] * yp[p] - T->bfc[F_YM][p][r] * ym[p]);
                sz += T->w2[p] * (T->bfc[F_ZP][p][r] * zp[p] - T->bfc[F_ZM][p][r] * zm[p]);
            }
            glf[r] = 0.25 * T->cb[r] * (dxi * sx + dyi * sy + dzi * sz) - intr[r];
        }
        // ---- source ----
        double src[NB];
#pragma unroll
        for (int r = 0; r < NB; r++) src[r] = 0.;
        if constexpr (NQ >= 11) {
            if (sp.ivis == 2 && f >= PXX) {
                // explicit part of the 10-moment relaxation (src/sources.f90:55-73)
                double s[NB];
#pragma unroll
                for (int r = 0; r < NB; r++) s[r] = 0.;
                for (int p = 0; p < NPG; p++) {
                    const double* b = T->bint[p];
                    double dn = 0., mx = 0., my = 0., mz = 0.;
#pragma unroll
                    for (int rr = 0; rr < NB; rr++) {
                        dn += b[rr] * m[RH][rr];
                        mx += b[rr] * m[MX][rr];
                        my += b[rr] * m[MY][rr];
                        mz += b[rr] * m[MZ][rr];
                    }
                    const double dni = 1. / dn;
                    const double vx = mx * dni, vy = my * dni, vz = mz * dni;
                    const double vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
                    const double colldn = sp.coll * dn;
                    double val;
                    if (f == PXX) val = colldn * (2 * vx2 - vy2 - vz2) * (1. / 3.);
                    else if (f == PYY) val = colldn * (2 * vy2 - vz2 - vx2) * (1. / 3.);
                    else if (f == PZZ) val = colldn * (2 * vz2 - vx2 - vy2) * (1. / 3.);
                    else if (f == PXY) val = colldn * vx * vy;
                    else if (f == PXZ) val = colldn * vx * vz;
                    else val = colldn * vy * vz;
#pragma unroll
                    for (int r = 0; r < NB; r++) s[r] += T->bwint[p][r] * val;
                }
#pragma unroll
                for (int r = 0; r < NB; r++) src[r] = 0.125 * T->cb[r] * s[r];
            }
            // 13 moments: the heat flux relaxes at its own rate
            const double rate = (NQ == 13 && f >= T_QX) ? sp.coll_q : sp.coll;
            if (sp.ivis == 0 && f >= PXX && rate != 0.) {
                for (int r = 0; r < NB; r++) {
                    double s = 0.;
                    for (int p = 0; p < NPG; p++) {
                        double qv = 0.;
#pragma unroll
                        for (int rr = 0; rr < NB; rr++) qv += T->bint[p][rr] * m[f][rr];
                        s += T->bwint[p][r] * (-rate * qv);
                    }
                    src[r] = 0.125 * T->cb[r] * s;
                }
            }
        }
        if (f >= MX && f <= MZ && sp.bf[f - MX] != 0.)
#pragma unroll
            for (int r = 0; r < NB; r++) src[r] += sp.bf[f - MX] * m[RH][r];
        if (f == EN && (sp.bf[0] != 0. || sp.bf[1] != 0. || sp.bf[2] != 0.))
#pragma unroll
            for (int r = 0; r < NB; r++) src[r] += sp.bf[0] * m[MX][r] + sp.bf[1] * m[MY][r] + sp.bf[2] * m[MZ][r];
        // ---- update ----
        if (sp.mode == 1) {
#pragma unroll
            for (int r = 0; r < NB; r++) qout[qoff(g, c, f, r)] = -(glf[r] - src[r]);
            continue;
        }
        if constexpr (NQ == 11) {
            if (sp.ivis == 2 && f >= PXX) {
                if (f <= PZZ) {
#pragma unroll
                    for (int r = 0; r < NB; r++) {
                        dglf[f - PXX][r] = glf[r];
                        dsrc[f - PXX][r] = src[r];
                    }
                    if (f == PZZ) {
                        // backward-Euler relaxation of the second moments towards
                        // their trace (src/integrator.f90:181-228)
#pragma unroll
                        for (int r = 0; r < NB; r++) {
                            const double Q_sum = sp.coll * dt * (m[PXX][r] + m[PYY][r] + m[PZZ][r]) * (1. / 3.);
                            const double glf_sum =
                                sp.coll * dt * dt * (dglf[0][r] + dglf[1][r] + dglf[2][r]) * (1. / 3.);
                            const double src_sum =
                                sp.coll * dt * dt * (dsrc[0][r] + dsrc[1][r] + dsrc[2][r]) * (1. / 3.);
#pragma unroll
                            for (int d = 0; d < 3; d++)
                                emit(PXX + d, r,
                                     faci * (m[PXX + d][r] - dt * dglf[d][r] + dt * dsrc[d][r] + Q_sum - glf_sum +
                                             src_sum));
                        }
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < NB; r++) emit(f, r, (m[f][r] - dt * glf[r] + dt * src[r]) * faci);
                }
                continue;
            }
        }
#pragma unroll
        for (int r = 0; r < NB; r++) {
            double v = m[f][r] - dt * (glf[r] - src[r]);
            if (sp.ivis == 1 && f >= PXX) v = (NQ == 13 && f >= T_QX ? faciq : faci) * v;
            emit(f, r, v);
        }
    }
    if (bad) {
        atomicOr(sp.err, 1);
        atomicMin(sp.bad, (unsigned long long)c);
    }
}

// ---------------------------------------------------------------------------
// Cell kernel of the 13-moment model, one field per wave: a workgroup is 13
// waves over the same 64 consecutive cells, and wave f computes field f (its
// volume integral, face lift, source and update).  A cell's 130 modes and 156
// volume accumulators never sit in one lane: each lane holds its field's 10
// modes and 12 accumulators, and the 13 point values of each volume point are
// exchanged through LDS.  Loads and
// stores of a wave touch one field of 64 consecutive cells.  The arithmetic
// of every value is that of k_cell.  The volume points are exchanged one
// plane of the rule (QP*QP points) at a time, and their primitives (1/rho,
// velocity, pressure) are computed once per point and shared the same way.
// ---------------------------------------------------------------------------
// Primitives of the 13-moment flux at one point (flux_point13): 1/rho,
// velocity, the floored pressure and the heat-flux flux col_q*T.
constexpr int NPRIM13 = 6;
__device__ __forceinline__ void prim13(const Phys& ph, const double* Q, double* pr) {
    const double dn = Q[RH], M_x = Q[MX], M_y = Q[MY], M_z = Q[MZ], Ener = Q[EN];
    const double dni = 1. / dn;
    const double vx = M_x * dni, vy = M_y * dni, vz = M_z * dni;
    const double smsq = M_x * M_x + M_y * M_y + M_z * M_z;
    double P = ph.gm1 * (Ener - 0.5 * dni * smsq);
    if (P < ph.P_floor) P = ph.P_floor;
    pr[0] = dni; pr[1] = vx; pr[2] = vy; pr[3] = vz; pr[4] = P;
    pr[5] = ph.colq * (P * dni);
}

// x, y and z flux of field f at one point (flux_point13, one field at a
// time; f is uniform over the wave).  The point's state and primitives are
// read from the exchange arrays of k_cell13 inside each case, so a wave loads
// only what its field's flux uses: qc points at field 0 of the point (fields
// qs apart), pc at primitive 0 (dni, vx, vy, vz, P, col_q*T; 64 apart).
template <int QS>
__device__ __forceinline__ void flux13_field(const Phys& ph, const double* qc, const double* pc, int f, double& Fx,
                                             double& Fy, double& Fz) {
    auto Q = [&](int e) { return qc[e * QS]; };
    auto V = [&](int e) { return pc[e * 64]; };
    switch (f) {
        case RH: Fx = Q(MX); Fy = Q(MY); Fz = Q(MZ); break;
        case MX: {
            const double M_x = Q(MX), vx = V(1), vy = V(2), vz = V(3), P = V(4);
            Fx = M_x * vx + P + Q(T_PXX); Fy = M_x * vy + Q(T_PXY); Fz = M_x * vz + Q(T_PXZ);
            break;
        }
        case MY: {
            const double M_y = Q(MY), vx = V(1), vy = V(2), vz = V(3), P = V(4);
            Fx = M_y * vx + Q(T_PXY); Fy = M_y * vy + P + Q(T_PYY); Fz = M_y * vz + Q(T_PYZ);
            break;
        }
        case MZ: {
            const double M_z = Q(MZ), vx = V(1), vy = V(2), vz = V(3), P = V(4);
            const double T_zz = -(Q(T_PXX) + Q(T_PYY));
            Fx = M_z * vx + Q(T_PXZ); Fy = M_z * vy + Q(T_PYZ); Fz = M_z * vz + P + T_zz;
            break;
        }
        case EN: {
            const double Ener = Q(EN), vx = V(1), vy = V(2), vz = V(3), P = V(4);
            const double T_xx = Q(T_PXX), T_yy = Q(T_PYY), T_zz = -(T_xx + T_yy);
            const double T_xy = Q(T_PXY), T_xz = Q(T_PXZ), T_yz = Q(T_PYZ);
            Fx = (Ener + P) * vx + T_xx * vx + T_xy * vy + T_xz * vz + Q(T_QX);
            Fy = (Ener + P) * vy + T_yy * vy + T_xy * vx + T_yz * vz + Q(T_QY);
            Fz = (Ener + P) * vz + T_zz * vz + T_xz * vx + T_yz * vy + Q(T_QZ);
            break;
        }
        case T_PXX: { const double vx = V(1), vy = V(2), vz = V(3);
                      Fx = ph.c4d3cv * vx; Fy = -ph.c2d3cv * vy; Fz = -ph.c2d3cv * vz; break; }
        case T_PYY: { const double vx = V(1), vy = V(2), vz = V(3);
                      Fx = -ph.c2d3cv * vx; Fy = ph.c4d3cv * vy; Fz = -ph.c2d3cv * vz; break; }
        case T_PXY: Fx = ph.colvis * V(2); Fy = ph.colvis * V(1); Fz = 0.; break;
        case T_PXZ: Fx = ph.colvis * V(3); Fy = 0.; Fz = ph.colvis * V(1); break;
        case T_PYZ: Fx = 0.; Fy = ph.colvis * V(3); Fz = ph.colvis * V(2); break;
        default: {
            const double ct = V(5);
            Fx = f == T_QX ? ct : 0.;
            Fy = f == T_QY ? ct : 0.;
            Fz = f == T_QZ ? ct : 0.;
        }
    }
}

// Face lift of one field in direction D (k_cell13): t[r] accumulates
// di*sum_p w2_p (bfc[+][p][r] fl(p,+) - bfc[-][p][r] fl(p,-)) over the
// directions, with the sums of k_cell.  The point loop stays rolled, so each
// point's table values are loaded as they are used; the face values of the
// next point are loaded one iteration ahead (the last iteration reloads its
// own point, in bounds).
template <int NB, int NF, int D>
__device__ __forceinline__ void lift13_dir(const double* __restrict__ fl, long nfl, long fm, long fp, double di,
                                           int f, const Tables* __restrict__ T, double (&t)[NB]) {
    constexpr int FM = 2 * D, FP = 2 * D + 1;  // F_XM, F_XP, F_YM, F_YP, F_ZM, F_ZP
    double sr[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) sr[r] = 0.;
    double vm = fl[((long)f * NF) * nfl + fm];
    double vp = fl[((long)f * NF) * nfl + fp];
#pragma unroll 1
    for (int p = 0; p < NF; p++) {
        const int pn = p + 1 < NF ? p + 1 : p;
        const double vmn = fl[((long)f * NF + pn) * nfl + fm];
        const double vpn = fl[((long)f * NF + pn) * nfl + fp];
        const double w = T->w2[p];
        sr[0] += w * (vp - vm);
#pragma unroll
        for (int r = 1; r < NB; r++) sr[r] += w * (T->bfc[FP][p][r] * vp - T->bfc[FM][p][r] * vm);
        vm = vmn;
        vp = vpn;
    }
#pragma unroll
    for (int r = 0; r < NB; r++) t[r] = D == 0 ? di * sr[r] : t[r] + di * sr[r];
}

constexpr int CELL13_WAVES = 13;

// ---------------------------------------------------------------------------
// get_min_dt (src/hermeshd.f90:145-198): per-block max of the wave speed over
// cells with rho > rh_mult*rh_floor, then a single-block finish that adds the
// sound speed of the last cell, as the reference does.
// ---------------------------------------------------------------------------
// sound speed of get_min_dt: ideal gas, or the water EOS (src/hermeshd.f90:165-166)
struct CflEos {
    double gamma;
    int ieos;
    double P1, T_floor;
};

__device__ __forceinline__ double cell_sound_speed(const double* q, const Grid& g, long c, const CflEos& e,
                                                   double* vmag0) {
    const double dn = q[qoff(g, c, RH, 0)];
    const double dni = 1. / dn;
    const double vx = q[qoff(g, c, MX, 0)] * dni, vy = q[qoff(g, c, MY, 0)] * dni, vz = q[qoff(g, c, MZ, 0)] * dni;
    const double cs = e.ieos == 2 ? sqrt(7.2 * e.P1 * pow(dn, 6.2) + e.T_floor)
                                  : sqrt(e.gamma * (q[qoff(g, c, EN, 0)] * dni - 0.5 * (vx * vx + vy * vy + vz * vz)));
    *vmag0 = fmax(fmax(fabs(vx) + cs, fabs(vy) + cs), fabs(vz) + cs);
    return cs;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

static __global__ __launch_bounds__(256) void k_wavespeed(const double* __restrict__ q, Grid g, CflEos eos,
                                                   double dn_min, double* __restrict__ partial) {
    __shared__ double red[4];
    double vmag = 0.;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N; c += (long)gridDim.x * blockDim.x) {
        double v0;
        cell_sound_speed(q, g, c, eos, &v0);
        const double dn = q[qoff(g, c, RH, 0)];
        if (v0 > vmag && dn > dn_min) vmag = v0;
    }
    vmag = wave_max(vmag);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = vmag;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) v = fmax(v, red[w]);
        partial[blockIdx.x] = v;
    }
}

static __global__ __launch_bounds__(256) void k_dt_finish(const double* __restrict__ q, Grid g, CflEos eos,
                                                   const double* __restrict__ partial, int npart, double cflm,
                                                   double* __restrict__ dt_out, double* __restrict__ t_io) {
    __shared__ double red[4];
    double v = 0.;
    for (int b = threadIdx.x; b < npart; b += blockDim.x) v = fmax(v, partial[b]);
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double vmag = red[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) vmag = fmax(vmag, red[w]);
        double v0;
        const double cs_last = cell_sound_speed(q, g, g.N - 1, eos, &v0);
        const double vmax = (vmag + cs_last) * g.dxi;
        const double dt = cflm / vmax;
        *dt_out = dt;
        if (t_io) *t_io = *t_io + dt;
    }
}

}  // namespace hd
