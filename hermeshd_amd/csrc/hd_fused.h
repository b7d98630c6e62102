// hd_fused.h -- one fused HIP kernel per SSP-RK stage: volume integral, face
// traces, Riemann fluxes on all six faces, surface lift, sources, update and
// the SSP combination, for one cell per thread with the state in registers.
//
// Operator (the reference's glflux + innerintegral2 + source_calc +
// advance_time_level, src/flux.f90:751-1022, src/sources.f90:12-84,
// src/integrator.f90:125-252), with dQ/dt = -glflux + source:
//   RHS[f][r] = sum_d dxi_d * ( sum_g Cvol(d,r,g) F_d(Q(g))[f]
//                             + sum_p Cface(-d,r,p) Fhat_{-d}(p)[f]
//                             + sum_p Cface(+d,r,p) Fhat_{+d}(p)[f] )  + source
//   Cvol  = 0.25 cb_r w3_g dphi_r/dxi_d(g),  Cface(-d) = +0.25 cb_r w2_p phi_r(-d,p),
//   Cface(+d) = -0.25 cb_r w2_p phi_r(+d,p).
// The basis is the Legendre product of src/basis_funcs.f90 with the mode order
// of src/initialize.f90:126-194; quadrature nodes and weights are the
// reference's values (1/sqrt(3); sqrt(15)/5, 5/9, 8/9) as exact hex literals.
// Each interior face's flux is evaluated by both adjacent cells (same inputs,
// same arithmetic, so both sides use the same value).
#pragma once
#include "hd_generic.h"
#include "hd_prim.h"

namespace hd {
namespace fz {

// ---------------------------------------------------------------- rules ----
template <int QP>
struct Rule;
template <>
struct Rule<2> {
    static HD_DEV constexpr double x(int i) { return i == 0 ? -0x1.279a74590331dp-1 : 0x1.279a74590331dp-1; }
    static HD_DEV constexpr double w(int) { return 1.0; }
    static HD_DEV constexpr double p2(int) { return 0x1.0p-53; }  // 1.5*s*s - 0.5 at s = 1/sqrt(3)
    static HD_DEV constexpr double w2(int, int) { return 1.0; }
    static HD_DEV constexpr double w3(int, int, int) { return 1.0; }
};
template <>
struct Rule<3> {
    static HD_DEV constexpr double x(int i) {
        return i == 0 ? -0x1.8c97ef43f7248p-1 : (i == 1 ? 0.0 : 0x1.8c97ef43f7248p-1);
    }
    static HD_DEV constexpr double w(int i) { return i == 1 ? 0x1.c71c71c71c71cp-1 : 0x1.1c71c71c71c72p-1; }
    static HD_DEV constexpr double p2(int i) { return i == 1 ? -0.5 : 0x1.999999999999cp-2; }
    // set_weights_3D: w2(i,j) = w1(j)*w1(i); w3(i; j,k) = w2(j,k)*w1(i)
    static HD_DEV constexpr double w2(int i, int j) { return w(j) * w(i); }
    static HD_DEV constexpr double w3(int i, int j, int k) { return w2(j, k) * w(i); }
};

// ---------------------------------------------------------------- basis ----
// Legendre degrees (a,b,c) in (x,y,z) of mode r
template <int NB>
HD_DEV constexpr int deg(int r, int ax) {
    // 1, x, y, z, yz, zx, xy, then xyz (NB 8) or P2x, P2y, P2z (NB 10); pure
    // arithmetic so that no lookup array is ever materialised on the device
    return r == 0 ? 0
         : r <= 3 ? (ax == r - 1 ? 1 : 0)
         : r <= 6 ? (ax == r - 4 ? 0 : 1)
         : (NB == 8 ? 1 : (ax == r - 7 ? 2 : 0));
}
template <int NB>
HD_DEV constexpr double cbasis(int r) {
    return double((2 * deg<NB>(r, 0) + 1) * (2 * deg<NB>(r, 1) + 1) * (2 * deg<NB>(r, 2) + 1));
}
template <int QP>
HD_DEV constexpr double Lval(int a, int i) { return a == 0 ? 1.0 : (a == 1 ? Rule<QP>::x(i) : Rule<QP>::p2(i)); }
// value at the face where the coordinate is -1 (side 0) or +1 (side 1)
HD_DEV constexpr double Lface(int a, int side) { return a == 1 ? (side ? 1.0 : -1.0) : 1.0; }
template <int QP>
HD_DEV constexpr double Dval(int a, int i) { return a == 0 ? 0.0 : (a == 1 ? 1.0 : 3.0 * Rule<QP>::x(i)); }

// interior point g = (i,j,k), x index slowest
template <int NB, int QP>
HD_DEV constexpr double phi_int(int r, int i, int j, int k) {
    return Lval<QP>(deg<NB>(r, 0), i) * Lval<QP>(deg<NB>(r, 1), j) * Lval<QP>(deg<NB>(r, 2), k);
}
// 0.25*cb_r*w3_g*d(phi_r)/d(xi_d) at g
template <int NB, int QP>
HD_DEV constexpr double cvol(int d, int r, int i, int j, int k) {
    const double a = d == 0 ? Dval<QP>(deg<NB>(r, 0), i) : Lval<QP>(deg<NB>(r, 0), i);
    const double b = d == 1 ? Dval<QP>(deg<NB>(r, 1), j) : Lval<QP>(deg<NB>(r, 1), j);
    const double c = d == 2 ? Dval<QP>(deg<NB>(r, 2), k) : Lval<QP>(deg<NB>(r, 2), k);
    return 0.25 * cbasis<NB>(r) * Rule<QP>::w3(i, j, k) * (a * b * c);
}
// face point p = (ii,jj), jj fastest: x faces (y=x[jj], z=x[ii]); y faces
// (x=x[ii], z=x[jj]); z faces (x=x[jj], y=x[ii])  (src/basis_funcs.f90:346-567)
template <int NB, int QP>
HD_DEV constexpr double phi_face(int dir, int side, int r, int ii, int jj) {
    const int a = deg<NB>(r, 0), b = deg<NB>(r, 1), c = deg<NB>(r, 2);
    if (dir == 0) return Lface(a, side) * Lval<QP>(b, jj) * Lval<QP>(c, ii);
    if (dir == 1) return Lval<QP>(a, ii) * Lface(b, side) * Lval<QP>(c, jj);
    return Lval<QP>(a, jj) * Lval<QP>(b, ii) * Lface(c, side);
}
template <int NB, int QP>
HD_DEV constexpr double cface(int dir, int side, int r, int ii, int jj) {
    return (side ? -0.25 : 0.25) * cbasis<NB>(r) * Rule<QP>::w2(ii, jj) * phi_face<NB, QP>(dir, side, r, ii, jj);
}

// Face tables in constant memory for the runtime point loops:
//   bf[dir][side][p][r] = phi_r on that face at point p
//   cf[dir][side][p][r] = cface
template <int NB, int QP>
struct FaceTab {
    double bf[3][2][QP * QP][NB];
    double cf[3][2][QP * QP][NB];
};
template <int NB, int QP>
constexpr FaceTab<NB, QP> make_face_tab() {
    FaceTab<NB, QP> t{};
    for (int d = 0; d < 3; d++)
        for (int s = 0; s < 2; s++)
            for (int ii = 0; ii < QP; ii++)
                for (int jj = 0; jj < QP; jj++)
                    for (int r = 0; r < NB; r++) {
                        t.bf[d][s][ii * QP + jj][r] = phi_face<NB, QP>(d, s, r, ii, jj);
                        t.cf[d][s][ii * QP + jj][r] = cface<NB, QP>(d, s, r, ii, jj);
                    }
    return t;
}
static __constant__ FaceTab<4, 2> c_face_4 = make_face_tab<4, 2>();
static __constant__ FaceTab<8, 2> c_face_8 = make_face_tab<8, 2>();
static __constant__ FaceTab<10, 3> c_face_10 = make_face_tab<10, 3>();
template <int NB>
__device__ __forceinline__ const FaceTab<NB, (NB == 10 ? 3 : 2)>& face_tab();
template <>
__device__ __forceinline__ const FaceTab<4, 2>& face_tab<4>() { return c_face_4; }
template <>
__device__ __forceinline__ const FaceTab<8, 2>& face_tab<8>() { return c_face_8; }
template <>
__device__ __forceinline__ const FaceTab<10, 3>& face_tab<10>() { return c_face_10; }

// ------------------------------------------------------ face traces ------
// On a face the basis collapses to a polynomial in the two tangential
// coordinates (row = the ii coordinate, col = the jj coordinate of the face
// point order above): phi_r = Lface(normal degree) * L(dr, row) * L(dc, col),
// with (dr, dc) in {(0,0), (1,0), (0,1), (1,1), (2,0), (0,2)}.  A side's
// trace is therefore sum_k K_k b_k(p) with NK = 3 (nb 4), 4 (nb 8) or 6
// (nb 10) coefficients K_k = sum of +-m_r over the modes of pair k: NK
// products per point instead of NB.  Both cells of a face evaluate a side's
// trace with this same arithmetic, so they still agree bitwise.
template <int NB>
HD_DEV constexpr int face_nk() { return NB == 10 ? 6 : (NB == 8 ? 4 : 3); }
HD_DEV constexpr int pair_index(int dr, int dc) {
    return dr == 0 ? (dc == 0 ? 0 : (dc == 1 ? 2 : 5)) : (dr == 1 ? (dc == 0 ? 1 : 3) : 4);
}
HD_DEV constexpr int pair_dr(int k) { return k == 1 || k == 3 ? 1 : (k == 4 ? 2 : 0); }
HD_DEV constexpr int pair_dc(int k) { return k == 2 || k == 3 ? 1 : (k == 5 ? 2 : 0); }
// tangential axes of a face: row = ii coordinate, col = jj coordinate
HD_DEV constexpr int row_axis(int dir) { return dir == 0 ? 2 : (dir == 1 ? 0 : 1); }
HD_DEV constexpr int col_axis(int dir) { return dir == 0 ? 1 : (dir == 1 ? 2 : 0); }

// b_k at face point p = (ii, jj), jj fastest, and w2_p*b_k (lifts)
template <int NB, int QP>
struct FacePoly {
    double b[QP * QP][face_nk<NB>()];
    double wb[QP * QP][face_nk<NB>()];
};
template <int NB, int QP>
constexpr FacePoly<NB, QP> make_face_poly() {
    FacePoly<NB, QP> t{};
    for (int ii = 0; ii < QP; ii++)
        for (int jj = 0; jj < QP; jj++)
            for (int k = 0; k < face_nk<NB>(); k++) {
                t.b[ii * QP + jj][k] = Lval<QP>(pair_dr(k), ii) * Lval<QP>(pair_dc(k), jj);
                t.wb[ii * QP + jj][k] = Rule<QP>::w2(ii, jj) * t.b[ii * QP + jj][k];
            }
    return t;
}
static __constant__ FacePoly<4, 2> c_fpoly_4 = make_face_poly<4, 2>();
static __constant__ FacePoly<8, 2> c_fpoly_8 = make_face_poly<8, 2>();
static __constant__ FacePoly<10, 3> c_fpoly_10 = make_face_poly<10, 3>();
template <int NB>
__device__ __forceinline__ const FacePoly<NB, (NB == 10 ? 3 : 2)>& face_poly();
template <>
__device__ __forceinline__ const FacePoly<4, 2>& face_poly<4>() { return c_fpoly_4; }
template <>
__device__ __forceinline__ const FacePoly<8, 2>& face_poly<8>() { return c_fpoly_8; }
template <>
__device__ __forceinline__ const FacePoly<10, 3>& face_poly<10>() { return c_fpoly_10; }

// K[f][k] of the side SIDE (0: coordinate -1, 1: +1) of a DIR face of a cell with modes m
template <int NB, int NQ, int DIR, int SIDE, int NK>
__device__ __forceinline__ void face_coeffs(const double (&m)[NQ][NB], double (&K)[NQ][NK]) {
#pragma unroll
    for (int f = 0; f < NQ; f++) {
#pragma unroll
        for (int k = 0; k < NK; k++) {
            double s = 0.;
            bool first = true;
#pragma unroll
            for (int r = 0; r < NB; r++) {
                if (pair_index(deg<NB>(r, row_axis(DIR)), deg<NB>(r, col_axis(DIR))) != k) continue;
                const bool neg = Lface(deg<NB>(r, DIR), SIDE) < 0.0;
                if (first) s = neg ? -m[f][r] : m[f][r];
                else s = neg ? s - m[f][r] : s + m[f][r];
                first = false;
            }
            K[f][k] = s;
        }
    }
}

template <int NQ, int NK>
__device__ __forceinline__ void face_eval(const double (&K)[NQ][NK], const double* __restrict__ b, double* out) {
#pragma unroll
    for (int f = 0; f < NQ; f++) {
        double s = K[f][0];
#pragma unroll
        for (int k = 1; k < NK; k++) s += K[f][k] * b[k];
        out[f] = s;
    }
}

// Keeps the scheduler from interleaving fully unrolled quadrature points,
// which would stretch live ranges past the register file.
__device__ __forceinline__ void point_barrier() { __builtin_amdgcn_sched_barrier(0); }

// ------------------------------------------------------------ volume -------
// innerintegral2 (src/flux.f90:751-928): Q at every Gauss point, the three
// directional fluxes from one set of primitives, and the projection onto
// d(phi_r)/d(xi_d) with compile-time weights (zero weights drop out).
template <int NB, int QP, int NQ, bool NZ>
__device__ __forceinline__ void volume_term(const double (&m)[NQ][NB], double (&rhs)[NQ][NB], const Phys& ph,
                                            double dxi, double dyi, double dzi, const StageParams& sp, long c,
                                            long N) {
#pragma unroll
    for (int i = 0; i < QP; i++) {
        // the stress of the plane's QP*QP points, loaded together so that
        // the points do not each wait for memory
        float Sv[QP * QP][6];
        if constexpr (NZ) {
            constexpr long npg = QP * QP * QP;
#pragma unroll
            for (int jk = 0; jk < QP * QP; jk++)
#pragma unroll
                for (int e = 0; e < 6; e++) Sv[jk][e] = sp.noise[3][((long)e * npg + i * QP * QP + jk) * N + c];
        }
#pragma unroll
        for (int j = 0; j < QP; j++)
#pragma unroll
            for (int k = 0; k < QP; k++) {
                double Qp[NQ];
#pragma unroll
                for (int f = 0; f < NQ; f++) {
                    double s = 0.;
#pragma unroll
                    for (int r = 0; r < NB; r++) {
                        const double b = phi_int<NB, QP>(r, i, j, k);
                        if (b != 0.0) s += b * m[f][r];
                    }
                    Qp[f] = s;
                }
                const pr::Prim st = pr::prim<NQ>(ph, Qp);
                double Fx[NQ], Fy[NQ], Fz[NQ];
                double S[6];
                const double* Sp = nullptr;
                if constexpr (NZ) {
                    // volume point (i,j,k) of the cell: stream id p*N + c
#pragma unroll
                    for (int e = 0; e < 6; e++) S[e] = Sv[j * QP + k][e];
                    Sp = S;
                }
                pr::flux<NQ, 0>(ph, Qp, st, Fx, Sp);
                pr::flux<NQ, 1>(ph, Qp, st, Fy, Sp);
                pr::flux<NQ, 2>(ph, Qp, st, Fz, Sp);
#pragma unroll
                for (int f = 0; f < NQ; f++) {
                    const double fx = dxi * Fx[f], fy = dyi * Fy[f], fz = dzi * Fz[f];
#pragma unroll
                    for (int r = 1; r < NB; r++) {
                        const double cx = cvol<NB, QP>(0, r, i, j, k);
                        const double cy = cvol<NB, QP>(1, r, i, j, k);
                        const double cz = cvol<NB, QP>(2, r, i, j, k);
                        if (cx != 0.0) rhs[f][r] += cx * fx;
                        if (cy != 0.0) rhs[f][r] += cy * fy;
                        if (cz != 0.0) rhs[f][r] += cz * fz;
                    }
                }
                point_barrier();
            }
    }
}

// ------------------------------------------------------------- faces -------
struct CellPos {
    int i, j, k;
    long c;
};

// The planes are reached by stepping one pointer a plane at a time; the
// empty asm hides the step count from the compiler, which would otherwise
// keep all NB*NQ plane offsets live in scalar registers (and spill them).
template <int NB, int NQ>
__device__ __forceinline__ void load_modes(const double* __restrict__ q, long N, long c, double (&m)[NQ][NB]) {
    // a global (address space 1) pointer, so the loads stay global loads
    const __attribute__((address_space(1))) double* p = (const __attribute__((address_space(1))) double*)(q + c);
#pragma unroll
    for (int r = 0; r < NB; r++)
#pragma unroll
        for (int f = 0; f < NQ; f++) {
            m[f][r] = *p;
            p += N;
            asm volatile("" : "+v"(p));
        }
}
template <int NB, int NQ>
__device__ __forceinline__ void load_modes(const double* __restrict__ q, const Grid& g, long c, double (&m)[NQ][NB]) {
    load_modes<NB, NQ>(q, g.N, c, m);
}

// Exterior state of an x-low inflow face point: inflow holds the reference's
// Qxlo_ext_def(ny,nz,nface,nQ) in Fortran order.
template <int NQ>
__device__ __forceinline__ void inflow_ghost(const double* __restrict__ inflow, const Grid& g, const CellPos& cp,
                                             int p, int nface, double* oth) {
    const long plane = (long)g.ny * g.nz;
    const long base = (long)cp.j + (long)g.ny * cp.k + plane * p;
    const long fs = plane * nface;  // stride of one field
    oth[RH] = inflow[base + fs * RH];
    oth[MX] = inflow[base + fs * MX];
    oth[MY] = 0.0;
    oth[MZ] = 0.0;
    oth[EN] = inflow[base + fs * EN];
}

// One face of the cell: DIR 0/1/2, SIDE 0 (minus) / 1 (plus).
// STORE: also write each point's scaled flux to xs[(p*NQ + f)*nx + i] for the
// x-neighbour's lift (the row-shared x faces of k_stage_xs).
// Modes of the neighbour across the DIR face on side SIDE (an interior or
// periodic face).
template <int NB, int NQ, int DIR, int SIDE>
__device__ __forceinline__ void load_neighbor(const double* __restrict__ q, const Grid& g, const CellPos& cp,
                                              double (&nm)[NQ][NB]) {
    const int n_ax = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
    const int a = DIR == 0 ? cp.i : DIR == 1 ? cp.j : cp.k;
    int an = SIDE ? a + 1 : a - 1;
    if (an < 0) an += n_ax;
    if (an >= n_ax) an -= n_ax;
    const long cn = DIR == 0 ? (cp.c - cp.i + an)
                  : DIR == 1 ? (cp.c + (long)(an - cp.j) * g.nx)
                             : (cp.c + (long)(an - cp.k) * g.nx * g.ny);
    load_modes<NB, NQ>(q, g, cn, nm);
}

// PRE: the neighbour's modes were loaded earlier (pre), so that their
// latency overlapped other work.
template <int NB, int QP, int NQ, int RM, bool NZ, int DIR, int SIDE, bool STORE = false, bool PRE = false>
__device__ __forceinline__ void face_term(const double* __restrict__ q, const Grid& g, const CellPos& cp,
                                          const double (&m)[NQ][NB], double (&rhs)[NQ][NB], const Phys& ph,
                                          double dxi_d, const StageParams& sp, double* xs = nullptr,
                                          const double (*pre)[NB] = nullptr) {
    constexpr int NF = QP * QP;
    const auto& T = face_tab<NB>();
    const int n_ax = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
    const int a = DIR == 0 ? cp.i : DIR == 1 ? cp.j : cp.k;
    const bool at_edge = SIDE ? (a == n_ax - 1) : (a == 0);
    const int bc = g.bc[2 * DIR + SIDE];
    const bool ext = at_edge && bc != BC_PERIODIC;
    constexpr int NK = face_nk<NB>();
    const auto& FP = face_poly<NB>();
    double ko[NQ][NK], kn[NQ][NK];
    face_coeffs<NB, NQ, DIR, SIDE, NK>(m, ko);
    if (!ext) {
        double nm[NQ][NB];
        if constexpr (PRE) {
#pragma unroll
            for (int f = 0; f < NQ; f++)
#pragma unroll
                for (int r = 0; r < NB; r++) nm[f][r] = pre[f][r];
        } else {
            load_neighbor<NB, NQ, DIR, SIDE>(q, g, cp, nm);
        }
        face_coeffs<NB, NQ, DIR, 1 - SIDE, NK>(nm, kn);
    }
    // outflow boundaries decide on the whole face (src/boundary.f90:213-338)
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
    long fid = 0;
    if constexpr (NZ) {
        // lattice id of this face; a periodic high face is the low face
        int af = a + SIDE;
        if (af == n_ax && bc == BC_PERIODIC) af = 0;
        const int fi = DIR == 0 ? af : cp.i, fj = DIR == 1 ? af : cp.j, fk = DIR == 2 ? af : cp.k;
        fid = (long)fi + (long)(g.nx + 1) * ((long)fj + (long)(g.ny + 1) * fk);
    }
    // the stress of all the face's points, loaded before the point loop
    float Sf[NZ ? NF : 1][6];
    if constexpr (NZ) {
#pragma unroll
        for (int p = 0; p < NF; p++)
#pragma unroll
            for (int e = 0; e < 6; e++) Sf[p][e] = sp.noise[DIR][((long)e * NF + p) * sp.nfb + fid];
    }
#pragma unroll 1
    for (int p = 0; p < NF; p++) {
        double own[NQ], oth[NQ];
        face_eval<NQ, NK>(ko, FP.b[p], own);
        if (!ext) {
            face_eval<NQ, NK>(kn, FP.b[p], oth);
        } else {
#pragma unroll
            for (int f = 0; f < NQ; f++) oth[f] = own[f];
            if (bc == BC_NOSLIP) {
                oth[MX] = oth[MY] = oth[MZ] = 0.0;
            } else if (bc == BC_WALL) {
                oth[MX + DIR] = 0.0;
            } else if (bc == BC_OUTFLOW) {
                if (DIR == 0 && SIDE == 1) {
                    if (oth[MX] > 0.0) oth[MX] = -own[MX];
                } else if (zero_normal) {
                    oth[MX + DIR] = 0.0;
                }
            }
            if constexpr (DIR == 0 && SIDE == 0) {
                // x_bc_inflow (src/boundary.f90:571-592): prescribed rh, mx, en;
                // no transverse momentum; the stress fields from the interior
                if (bc == BC_INFLOW) inflow_ghost<NQ>(sp.inflow, g, cp, p, QP * QP, oth);
            }
        }
        const double* QL = SIDE ? own : oth;
        const double* QR = SIDE ? oth : own;
        double S[6];
        const double* Sp = nullptr;
        if constexpr (NZ) {
            // p is a loop counter (not unrolled): select rather than index,
            // which keeps Sf in registers
#pragma unroll
            for (int e = 0; e < 6; e++) S[e] = Sf[0][e];
#pragma unroll
            for (int pp = 1; pp < NF; pp++)
                if (pp == p)
#pragma unroll
                    for (int e = 0; e < 6; e++) S[e] = Sf[pp][e];
            Sp = S;
        }
        double FL[NQ], FR[NQ], out[NQ];
        if constexpr (RM == RIEMANN_ROE) {
            flux_point<NQ, DIR>(ph, QL, FL, Sp);
            flux_point<NQ, DIR>(ph, QR, FR, Sp);
#pragma unroll
            for (int f = 0; f < NQ; f++) out[f] = 0.5 * (FL[f] + FR[f]);
            roe_point<DIR>(ph, QL, QR, FL, FR, out);
        } else {
            const pr::Prim sL = pr::prim<NQ>(ph, QL), sR = pr::prim<NQ>(ph, QR);
            pr::flux<NQ, DIR>(ph, QL, sL, FL, Sp);
            pr::flux<NQ, DIR>(ph, QR, sR, FR, Sp);
#pragma unroll
            for (int f = 0; f < NQ; f++) out[f] = 0.5 * (FL[f] + FR[f]);
            if constexpr (RM == RIEMANN_HLLC) pr::hllc<DIR>(ph, QL, sL, QR, sR, FL, FR, out);
        }
#pragma unroll
        for (int f = 0; f < NQ; f++) {
            const double v = dxi_d * out[f];
            if constexpr (STORE) xs[(long)(p * NQ + f) * g.nx + cp.i] = v;
#pragma unroll
            for (int r = 0; r < NB; r++) rhs[f][r] += T.cf[DIR][SIDE][p][r] * v;
        }
    }
}

// The -x face of a cell is the +x face of its left neighbour in the same row:
// its scaled point fluxes come from xs (written by that neighbour), bitwise
// what the cell would compute itself, and only the lift is done here.
template <int NB, int QP, int NQ>
__device__ __forceinline__ void x_lift_shared(const double* __restrict__ xs, const Grid& g, const CellPos& cp,
                                              double (&rhs)[NQ][NB]) {
    constexpr int NF = QP * QP;
    const auto& T = face_tab<NB>();
    const int il = cp.i == 0 ? g.nx - 1 : cp.i - 1;
#pragma unroll 1
    for (int p = 0; p < NF; p++) {
#pragma unroll
        for (int f = 0; f < NQ; f++) {
            const double v = xs[(long)(p * NQ + f) * g.nx + il];
#pragma unroll
            for (int r = 0; r < NB; r++) rhs[f][r] += T.cf[0][0][p][r] * v;
        }
    }
}

// ----------------------------------------------------------- limiter -------
// limiter (src/flux.f90:1077-1164), in place, compile-time face tables.  The
// pressure step takes the minimum face pressure first and divides once:
// (Pave-epsP)/(Pave-P) is monotone in P for P < epsP <= Pave, so this equals
// the reference's minimum over points.
// Exact early-out.  Every basis function is bounded by 1 in magnitude at the
// face points (products of L0 = 1, L1 and L2 on [-1,1]), so
// rho(p) >= rho_1 - sum_{r>1} |rho_r|, E(p) >= E_1 - sum |E_r|,
// |m(p)|^2 <= sum_a (sum_r |m_a,r|)^2, and the face pressure is bounded below
// by gm1*(E_lb - |m|^2_ub/(2 rho_lb)).  When both bounds clear the floors (with
// a relative margin far above rounding), the limiter's minima clear them too
// and it leaves the cell unchanged, so its face loops can be skipped.
template <int NB, int MQ>
__device__ __forceinline__ bool limiter_inactive(const double (&m)[MQ][NB], const LimParams& lp) {
    double sr = 0., se = 0., mm[3] = {0., 0., 0.};
#pragma unroll
    for (int r = 1; r < NB; r++) {
        sr += fabs(m[RH][r]);
        se += fabs(m[EN][r]);
#pragma unroll
        for (int a = 0; a < 3; a++) mm[a] += fabs(m[MX + a][r]);
    }
    const double rlb = m[RH][0] - sr;
    if (!(rlb > lp.rh_floor + 1e-10 * (fabs(m[RH][0]) + sr))) return false;
    double m2 = 0.;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double b = fabs(m[MX + a][0]) + mm[a];
        m2 += b * b;
    }
    const double Pave =
        lp.gm1 * (m[EN][0] - 0.5 * (m[MX][0] * m[MX][0] + m[MY][0] * m[MY][0] + m[MZ][0] * m[MZ][0]) / m[RH][0]);
    if (Pave < lp.epsP) return false;
    const double kin = 0.5 * m2 / rlb;
    const double plb = lp.gm1 * (m[EN][0] - se - kin);
    return plb > lp.epsP + 1e-10 * lp.gm1 * (fabs(m[EN][0]) + se + kin);
}

// Limits the modes m[0..4][*] of one cell in place; returns true if changed.
template <int NB, int QP, int MQ>
__device__ __forceinline__ bool limit_cell(double (&m)[MQ][NB], const LimParams& lp) {
    bool changed = false;
    if (m[RH][0] < lp.rh_floor) {
#pragma unroll
        for (int r = 1; r < NB; r++)
#pragma unroll
            for (int f = 0; f < 5; f++) m[f][r] = 0.0;
        m[RH][0] = lp.rh_floor;
        changed = true;
    } else if (limiter_inactive<NB, MQ>(m, lp)) {
        // no face point can reach a floor: the loops below would change nothing
    } else {
        const auto& T = face_tab<NB>();
        double rmin = 0.;
#pragma unroll 1
        for (int p = 0; p < 6 * QP * QP; p++) {
            const double* b = T.bf[p / (2 * QP * QP)][(p / (QP * QP)) % 2][p % (QP * QP)];
            double s = 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) s += b[r] * m[RH][r];
            rmin = (p == 0) ? s : (s < rmin ? s : rmin);
        }
        if (rmin < lp.rh_floor) {
            double theta = (lp.rh_floor - m[RH][0]) / (rmin - m[RH][0]);
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
            double pmin = 0.;
#pragma unroll 1
            for (int p = 0; p < 6 * QP * QP; p++) {
                const double* b = T.bf[p / (2 * QP * QP)][(p / (QP * QP)) % 2][p % (QP * QP)];
                double e[5];
#pragma unroll
                for (int f = 0; f < 5; f++) {
                    double s = 0.;
#pragma unroll
                    for (int r = 0; r < NB; r++) s += b[r] * m[f][r];
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
                if (theta != 1.) {
#pragma unroll
                    for (int r = 1; r < NB; r++)
#pragma unroll
                        for (int f = 0; f < 5; f++) m[f][r] = theta * m[f][r];
                    changed = true;
                }
            }
        }
    }
    return changed;
}

template <int NB, int QP>
__global__ __launch_bounds__(256, (NB == 4 ? 4 : (NB == 8 ? 2 : 1))) void k_limit(double* __restrict__ q, Grid g, LimParams lp) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    double m[5][NB];
#pragma unroll
    for (int r = 0; r < NB; r++)
#pragma unroll
        for (int f = 0; f < 5; f++) m[f][r] = q[((long)r * g.nQ + f) * g.N + c];
    if (limit_cell<NB, QP, 5>(m, lp)) {
#pragma unroll
        for (int r = 0; r < NB; r++)
#pragma unroll
            for (int f = 0; f < 5; f++) q[((long)r * g.nQ + f) * g.N + c] = m[f][r];
    }
}

// ------------------------------------------------------------ kernel -------
__device__ __forceinline__ void fence() { asm volatile("" ::: "memory"); }

// advance_time_level and the SSP combination of one cell (src/integrator.f90:
// 62-86, 210-232): rhs becomes m + dt*rhs (the stress fields scaled by faci
// under ivis 1), combined with the step-start state qn by COMB (0: none,
// 1: RK2 (qn+v)/2, 2: RK3 3qn/4 + v/4, 3: RK3 qn/3 + 2v/3).  Returns whether a
// mean value came out NaN.
template <int NB, int NQ, int COMB>
__device__ __forceinline__ bool advance(const double (&m)[NQ][NB], double (&rhs)[NQ][NB],
                                        const double* __restrict__ qn, long N, long c, double dt, double faci,
                                        bool relax) {
    double q0[COMB ? NQ : 1][COMB ? NB : 1];
    if constexpr (COMB != 0) load_modes<NB, NQ>(qn, N, c, q0);
    bool bad = false;
#pragma unroll
    for (int f = 0; f < NQ; f++)
#pragma unroll
        for (int r = 0; r < NB; r++) {
            double v = m[f][r] + dt * rhs[f][r];
            if (f >= PXX && relax) v = faci * v;
            if (r == 0 && v != v) bad = true;
            if constexpr (COMB == 1) v = 0.5 * (q0[f][r] + v);
            if constexpr (COMB == 2) v = 0.75 * q0[f][r] + 0.25 * v;
            if constexpr (COMB == 3) v = (1. / 3.) * q0[f][r] + (2. / 3.) * v;
            rhs[f][r] = v;  // reuse the accumulator as the output cell
        }
    return bad;
}

// One cell of one stage; `bad_id` is the index reported when mode 1 of the
// output is NaN.
// XS: the workgroup is one periodic row of cells (k_stage_xs); each x face is
// evaluated once, by its left cell, and shared through xs (LDS).
template <int NB, int QP, int NQ, int RM, bool NZ, bool XS = false>
__device__ __forceinline__ void stage_cell(const double* __restrict__ qin, const double* __restrict__ qn,
                                           double* __restrict__ qout, const Grid& g, const Phys& ph,
                                           const StageParams& sp, const CellPos& cp, long bad_id,
                                           double* xs = nullptr) {
    const long c = cp.c;
    double m[NQ][NB];
    load_modes<NB, NQ>(qin, g, c, m);
    double rhs[NQ][NB];
#pragma unroll
    for (int f = 0; f < NQ; f++)
#pragma unroll
        for (int r = 0; r < NB; r++) rhs[f][r] = 0.;
    // The fences keep each section's loads inside it, so the compiler does not
    // hoist all six neighbours' modes at once (which would spill).  The first
    // face's neighbour is loaded ahead of the volume term, whose arithmetic
    // covers that load's latency.
    double nm0[NQ][NB];
    if constexpr (XS) {
        load_neighbor<NB, NQ, 0, 1>(qin, g, cp, nm0);
    } else {
        if (!(cp.i == 0 && g.bc[0] != BC_PERIODIC)) load_neighbor<NB, NQ, 0, 0>(qin, g, cp, nm0);
    }
    volume_term<NB, QP, NQ, NZ>(m, rhs, ph, g.dxi, g.dyi, g.dzi, sp, c, g.N);
    fence();
    if constexpr (XS) {
        face_term<NB, QP, NQ, RM, NZ, 0, 1, true, true>(qin, g, cp, m, rhs, ph, g.dxi, sp, xs, nm0);
        __syncthreads();
        x_lift_shared<NB, QP, NQ>(xs, g, cp, rhs);
    } else {
        face_term<NB, QP, NQ, RM, NZ, 0, 0, false, true>(qin, g, cp, m, rhs, ph, g.dxi, sp, nullptr, nm0);
        fence();
        face_term<NB, QP, NQ, RM, NZ, 0, 1>(qin, g, cp, m, rhs, ph, g.dxi, sp);
    }
    fence();
    face_term<NB, QP, NQ, RM, NZ, 1, 0>(qin, g, cp, m, rhs, ph, g.dyi, sp);
    fence();
    face_term<NB, QP, NQ, RM, NZ, 1, 1>(qin, g, cp, m, rhs, ph, g.dyi, sp);
    fence();
    face_term<NB, QP, NQ, RM, NZ, 2, 0>(qin, g, cp, m, rhs, ph, g.dzi, sp);
    fence();
    face_term<NB, QP, NQ, RM, NZ, 2, 1>(qin, g, cp, m, rhs, ph, g.dzi, sp);
    fence();
    // sources (src/sources.f90:12-84): ivis 0 relaxes the stress explicitly
    // (-coll*P, projected exactly in modal space); body force rho*g, m.g
    if constexpr (NQ >= 11) {
        if (sp.ivis == 0 && sp.coll != 0.)
#pragma unroll
            for (int f = PXX; f <= PYZ; f++)
#pragma unroll
                for (int r = 0; r < NB; r++) rhs[f][r] += -sp.coll * m[f][r];
    }
    if (sp.bf[0] != 0. || sp.bf[1] != 0. || sp.bf[2] != 0.) {
#pragma unroll
        for (int r = 0; r < NB; r++) {
            rhs[MX][r] += sp.bf[0] * m[RH][r];
            rhs[MY][r] += sp.bf[1] * m[RH][r];
            rhs[MZ][r] += sp.bf[2] * m[RH][r];
            rhs[EN][r] += sp.bf[0] * m[MX][r] + sp.bf[1] * m[MY][r] + sp.bf[2] * m[MZ][r];
        }
    }
    // advance_time_level + SSP combination
    const double dt = *sp.dt;
    const double faci = 1. / (1. + dt * sp.coll);
    bool bad = false;
    if (sp.mode == 1) {
#pragma unroll
        for (int f = 0; f < NQ; f++)
#pragma unroll
            for (int r = 0; r < NB; r++) qout[((long)r * NQ + f) * g.N + c] = rhs[f][r];
        return;
    }
    // one branch per combination, outside the element loop, so that the
    // step-start values are loaded together rather than one per branch
    const bool relax = NQ >= 11 && sp.ivis == 1;
    switch (sp.combine) {
        case 1: bad = advance<NB, NQ, 1>(m, rhs, qn, g.N, c, dt, faci, relax); break;
        case 2: bad = advance<NB, NQ, 2>(m, rhs, qn, g.N, c, dt, faci, relax); break;
        case 3: bad = advance<NB, NQ, 3>(m, rhs, qn, g.N, c, dt, faci, relax); break;
        default: bad = advance<NB, NQ, 0>(m, rhs, qn, g.N, c, dt, faci, relax); break;
    }
    // When this output is the next stage's input, apply that stage's limiter
    // here: the reference limits the stage input in place before anything
    // reads it (src/integrator.f90:106-113), so the stored values are the same.
    if (sp.limit_out) limit_cell<NB, QP, NQ>(rhs, sp.lp);
#pragma unroll
    for (int f = 0; f < NQ; f++)
#pragma unroll
        for (int r = 0; r < NB; r++) qout[((long)r * NQ + f) * g.N + c] = rhs[f][r];
    if (bad) {
        atomicOr(sp.err, 1);
        atomicMin(sp.bad, (unsigned long long)bad_id);
    }
}


template <int NB, int QP, int NQ, int RM, bool NZ>
__global__ __launch_bounds__(256) void k_stage(const double* __restrict__ qin, const double* __restrict__ qn,
                                               double* __restrict__ qout, Grid g, Phys ph, StageParams sp) {
    const long c = (long)xcd_block() * blockDim.x + threadIdx.x;
    if (c >= g.N) return;
    CellPos cp;
    cp.c = c;
    cp.i = (int)(c % g.nx);
    cp.j = (int)((c / g.nx) % g.ny);
    cp.k = (int)(c / ((long)g.nx * g.ny));
    stage_cell<NB, QP, NQ, RM, NZ>(qin, qn, qout, g, ph, sp, cp, c);
}

// One periodic row of nx cells per workgroup (nx a multiple of 64, at most
// 256): the x faces are evaluated once and shared through LDS
// (NF*NQ*nx doubles of dynamic shared memory).  The grid covers the rows
// exactly, so every thread reaches the barrier.  Each XCD takes a contiguous
// range of workgroups; with rows_zfast the rows run with z fastest, so that
// a row's z neighbours are the adjacent workgroups and its y neighbours nz
// workgroups away, both inside the window of rows in flight on that XCD
// (whose L2 then serves the neighbour modes); otherwise y runs fastest.
template <int NB, int QP, int NQ, int RM, bool NZ>
__global__ __launch_bounds__(256) void k_stage_xs(const double* __restrict__ qin, const double* __restrict__ qn,
                                                  double* __restrict__ qout, Grid g, Phys ph, StageParams sp) {
    extern __shared__ double xs[];
    const long r = (long)xcd_block();
    const int j = sp.rows_zfast ? (int)(r / g.nz) : (int)(r % g.ny);
    const int k = sp.rows_zfast ? (int)(r % g.nz) : (int)(r / g.ny);
    const long row = (long)j + (long)g.ny * k;
    const long c = row * g.nx + threadIdx.x;
    CellPos cp;
    cp.c = c;
    cp.i = (int)threadIdx.x;
    cp.j = j;
    cp.k = k;
    stage_cell<NB, QP, NQ, RM, NZ, true>(qin, qn, qout, g, ph, sp, cp, c, xs);
}

// Face fluxes of the 13-moment model on a grid periodic in every direction,
// as face moments fl[((f*NK + k)*nfaces) + face] (lift13_mom).  A workgroup
// owns 64 consecutive faces of direction DIR, one wave per face point, so no
// lane holds more than one point's states.  Phase 1 turns the modes of the
// cells on both sides into face-polynomial coefficients (face_coeffs, NK per
// field and side) in LDS, one (side, field, face) item per thread.  Phase 2
// evaluates point p of every face on wave p (face_eval), then the physical
// fluxes and the Riemann solver of k_face_flux.  Phase 3 reduces the points
// of each (field, face) to the NK face moments.  Lanes past the last face
// reuse the last face and store nothing.
template <int NB, int QP, int RM, int DIR>
__global__ __launch_bounds__(QP * QP * 64) void k_face13(const double* __restrict__ q, double* __restrict__ fl,
                                                        Grid g, Phys ph) {
    constexpr int NQ = 13, NF = QP * QP, NK = face_nk<NB>();
    __shared__ double ks[2][NQ][NK][64];
    const int ni = g.nx + (DIR == 0), nj = g.ny + (DIR == 1);
    const long nfaces = (long)ni * nj * (g.nz + (DIR == 2));
    const long f0 = (long)xcd_block() * 64;  // XCD-contiguous ranges of faces
    // phase 1: NIT (side, field) items per thread, all their loads issued
    // before the first is used
    constexpr int NIT = (2 * NQ * 64 + NF * 64 - 1) / (NF * 64);
    // the item stride NF*64 is a multiple of 64, so every item of a thread
    // belongs to its lane's face: the two cells are found once per lane, in
    // 32-bit arithmetic (hd_create bounds the faces per direction by 2^31)
    long cell2[2];
    {
        const long fl64 = f0 + (long)(threadIdx.x & 63);
        const unsigned fid = (unsigned)(fl64 < nfaces ? fl64 : nfaces - 1);
        const unsigned fq = fid / (unsigned)ni;
        int ci = (int)(fid - fq * (unsigned)ni), cj = (int)(fq % (unsigned)nj), ck = (int)(fq / (unsigned)nj);
        int& a = DIR == 0 ? ci : DIR == 1 ? cj : ck;
        const int n_ax = DIR == 0 ? g.nx : DIR == 1 ? g.ny : g.nz;
        // side 0: the cell below the face (index a - 1, wrapped), side 1: above (index a, wrapped)
        const int a_lo = a == 0 ? n_ax - 1 : a - 1, a_hi = a == n_ax ? 0 : a;
        a = a_lo;
        cell2[0] = (long)(unsigned)(ci + g.nx * (cj + g.ny * ck));
        a = a_hi;
        cell2[1] = (long)(unsigned)(ci + g.nx * (cj + g.ny * ck));
    }
    double m1[NIT][1][NB];
#pragma unroll
    for (int u = 0; u < NIT; u++) {
        const int it = (int)threadIdx.x + u * NF * 64;
        if (it < 2 * NQ * 64) {
            const int side = (it >> 6) / NQ, fld = (it >> 6) % NQ;
            const long cell = side == 0 ? cell2[0] : cell2[1];
#pragma unroll
            for (int r = 0; r < NB; r++) m1[u][0][r] = q[((long)r * NQ + fld) * g.N + cell];
        }
    }
#pragma unroll
    for (int u = 0; u < NIT; u++) {
        const int it = (int)threadIdx.x + u * NF * 64;
        if (it < 2 * NQ * 64) {
            const int l = it & 63, side = (it >> 6) / NQ, fld = (it >> 6) % NQ;
            double K1[1][NK];
            if (side == 0) face_coeffs<NB, 1, DIR, 1, NK>(m1[u], K1);
            else face_coeffs<NB, 1, DIR, 0, NK>(m1[u], K1);
#pragma unroll
            for (int k = 0; k < NK; k++) ks[side][fld][k][l] = K1[0][k];
        }
    }
    __syncthreads();
    // lanes past the last face evaluate the last face again and store nothing
    const int p = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int l = (int)(threadIdx.x & 63);
    const auto& FP = face_poly<NB>();
    double QL[NQ], QR[NQ];
#pragma unroll
    for (int f = 0; f < NQ; f++) {
        double sl = ks[0][f][0][l], sr = ks[1][f][0][l];
#pragma unroll
        for (int k = 1; k < NK; k++) {
            sl += ks[0][f][k][l] * FP.b[p][k];
            sr += ks[1][f][k][l] * FP.b[p][k];
        }
        QL[f] = sl;
        QR[f] = sr;
    }
    double FL[NQ], FR[NQ], out[NQ];
    gflux<NQ, DIR>(ph, QL, FL);
    gflux<NQ, DIR>(ph, QR, FR);
#pragma unroll
    for (int f = 0; f < NQ; f++) out[f] = 0.5 * (FL[f] + FR[f]);
    if constexpr (RM == RIEMANN_HLLC) hllc_point<DIR>(ph, QL, QR, FL, FR, out);  // ideal gas: no last-state speed
    if constexpr (RM == RIEMANN_ROE) roe_point<DIR, NQ>(ph, QL, QR, FL, FR, out);
    // phase 3: the face moments M_k = sum_p wb_p[k] F_p of each field, the
    // point fluxes exchanged through ks once every wave has read it
    static_assert(NF <= 2 * NK, "the point fluxes fit in ks");
    double* vs = &ks[0][0][0][0];
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NQ; f++) vs[(f * NF + p) * 64 + l] = out[f];
    __syncthreads();
    for (int it = (int)threadIdx.x; it < NQ * NK * 64; it += NF * 64) {
        const int lf = it & 63;
        const int fk = __builtin_amdgcn_readfirstlane(it >> 6);
        const int f = fk / NK, k = fk - f * NK;
        double s = 0.;
#pragma unroll
        for (int pp = 0; pp < NF; pp++) s += FP.wb[pp][k] * vs[(f * NF + pp) * 64 + lf];
        if (f0 + lf < nfaces) fl[((long)f * NK + k) * nfaces + f0 + lf] = s;
    }
}

// Face lift of one field in direction D from the face moments of k_face13:
// sum_p w2_p phi_r(p) F_p on a face is Lface(deg_D(r), side) M_{k(r)}, with
// k(r) the mode's face-polynomial pair, so the sums of lift13_dir become
// t[r] += di (M+_{k(r)} - Lface(deg_D(r), 0) M-_{k(r)}), M+ of the cell's
// high face and M- of its low face.
template <int NB, int D>
HD_DEV constexpr int lift_k(int r) {
    return pair_index(deg<NB>(r, row_axis(D)), deg<NB>(r, col_axis(D)));
}
template <int NB, int D>
__device__ __forceinline__ void lift13_mom(const double* __restrict__ fl, long nfl, long fm, long fp, double di,
                                           int f, double (&t)[NB]) {
    constexpr int NK = face_nk<NB>();
    double mm[NK], mp[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) {
        mm[k] = fl[((long)f * NK + k) * nfl + fm];
        mp[k] = fl[((long)f * NK + k) * nfl + fp];
    }
#pragma unroll
    for (int r = 0; r < NB; r++) {
        const int k = lift_k<NB, D>(r);
        const double s = deg<NB>(r, D) == 1 ? mp[k] + mm[k] : mp[k] - mm[k];
        t[r] = D == 0 ? di * s : t[r] + di * s;
    }
}

}  // namespace fz


// MOM: the face arrays hold the face moments of fz::k_face13 (periodic
// grids); otherwise the point fluxes of k_face_flux.
template <int NB, int QP, bool MOM>
__global__ __launch_bounds__(CELL13_WAVES * 64) void k_cell13(
    const double* __restrict__ qin, const double* __restrict__ qn, double* __restrict__ qout,
    const double* __restrict__ flx, const double* __restrict__ fly, const double* __restrict__ flz,
    const Tables* __restrict__ T, Grid g, Phys ph, StageParams sp) {
    constexpr int NQ = 13, NF = QP * QP, NPG = QP * QP * QP;
    constexpr int NS = (NB > 4) ? 12 : 3;
    constexpr int NCH = NF;  // volume points per exchange: one plane of the 3D rule
    static_assert(NCH <= CELL13_WAVES, "one wave per point computes the primitives");
    __shared__ double qp_s[NQ][NCH][64];
    __shared__ double pr_s[NCH][NPRIM13][64];
    const int f = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    const long c = (long)xcd_block() * 64 + lane;  // XCD-contiguous ranges of cells
    const bool live = c < g.N;
    const long cc = live ? c : g.N - 1;  // lanes past the grid compute a copy of the last cell and store nothing
    // 32-bit position arithmetic: hd_create bounds the cells by 2^31
    const unsigned ucc = (unsigned)cc, cq = ucc / (unsigned)g.nx;
    const int i = (int)(ucc - cq * (unsigned)g.nx), j = (int)(cq % (unsigned)g.ny), k = (int)(cq / (unsigned)g.ny);
    double mf[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) mf[r] = qin[qoff(g, cc, f, r)];
    const double dxi = g.dxi, dyi = g.dyi, dzi = g.dzi;
    // face lift of field f first, its loads beside those of the modes (one
    // memory phase instead of two), one direction at a time:
    // t[r] = dxi*sx[r] + dyi*sy[r] + dzi*sz[r] with the sums of k_cell
    const long fxm = (long)i + (long)(g.nx + 1) * ((long)j + (long)g.ny * k);
    const long fym = (long)i + (long)g.nx * ((long)j + (long)(g.ny + 1) * k);
    const long fzm = (long)i + (long)g.nx * ((long)j + (long)g.ny * k);
    const long nfx = (long)(g.nx + 1) * g.ny * g.nz, nfy = (long)g.nx * (g.ny + 1) * g.nz,
               nfz = (long)g.nx * g.ny * (g.nz + 1);
    double t[NB];
    if constexpr (MOM) {
        fz::lift13_mom<NB, 0>(flx, nfx, fxm, fxm + 1, dxi, f, t);
        fz::lift13_mom<NB, 1>(fly, nfy, fym, fym + g.nx, dyi, f, t);
        fz::lift13_mom<NB, 2>(flz, nfz, fzm, fzm + (long)g.nx * g.ny, dzi, f, t);
    } else {
        lift13_dir<NB, NF, 0>(flx, nfx, fxm, fxm + 1, dxi, f, T, t);
        lift13_dir<NB, NF, 1>(fly, nfy, fym, fym + g.nx, dyi, f, T, t);
        lift13_dir<NB, NF, 2>(flz, nfz, fzm, fzm + (long)g.nx * g.ny, dzi, f, T, t);
    }
    // ---- volume integral: every wave reaches every barrier ----
    // glf[r] = 0.25 cb[r] (t[r] - sum_d di acc_d[r]) for r >= 1, so each
    // accumulator starts at -t[r] over its grid factor and t[1..] need not
    // stay live through the volume loop (the kernel runs at 128 VGPRs)
    double acc[NS];
    {
        const double dx = 1. / dxi, dy = 1. / dyi, dz = 1. / dzi;
        acc[0] = -dx * t[KX];
        acc[1] = -dy * t[KY];
        acc[2] = -dz * t[KZ];
        if constexpr (NB > 4) {
            acc[3] = -dy * t[KYZ];
            acc[4] = 0.;
            acc[5] = -dx * t[KZX];
            acc[6] = 0.;
            acc[7] = -dx * t[KXY];
            acc[8] = 0.;
            if constexpr (NB == 10) {
                acc[9] = -dx * t[KXX];
                acc[10] = -dy * t[KYY];
                acc[11] = -dz * t[KZZ];
            } else {
                acc[9] = -dx * t[KXYZ];
                acc[10] = 0.;
                acc[11] = 0.;
            }
        }
    }
    const double t0 = t[0];
    // volume integral, NCH points per exchange: each wave writes its field's
    // point values, then reads the 13 values of each point; every wave
    // reaches every barrier
#pragma unroll 1
    for (int p0 = 0; p0 < NPG; p0 += NCH) {
#pragma unroll
        for (int q = 0; q < NCH; q++) {
            const double* b = T->bint[p0 + q];
            double v = 0.;
#pragma unroll
            for (int r = 0; r < NB; r++) v += b[r] * mf[r];
            qp_s[f][q][lane] = v;
        }
        __syncthreads();
        // the primitives of point f of the plane, once per point
        if (f < NCH) {
            double Qp[NQ], pr[NPRIM13];
#pragma unroll
            for (int e = 0; e < NQ; e++) Qp[e] = qp_s[e][f][lane];
            prim13(ph, Qp, pr);
#pragma unroll
            for (int e = 0; e < NPRIM13; e++) pr_s[f][e][lane] = pr[e];
        }
        __syncthreads();
#pragma unroll 1
        for (int q = 0; q < NCH; q++) {
            const int p = p0 + q;
            const double* b = T->bint[p];
            double Fx, Fy, Fz;
            flux13_field<NCH * 64>(ph, &qp_s[0][q][lane], &pr_s[q][0][lane], f, Fx, Fy, Fz);
            const double w = T->w3[p];
            acc[0] += w * Fx;
            acc[1] += w * Fy;
            acc[2] += w * Fz;
            if constexpr (NB > 4) {
                acc[3] += w * b[KZ] * Fy;
                acc[4] += w * b[KY] * Fz;
                acc[5] += w * b[KZ] * Fx;
                acc[6] += w * b[KX] * Fz;
                acc[7] += w * b[KY] * Fx;
                acc[8] += w * b[KX] * Fy;
                if constexpr (NB == 10) {
                    acc[9] += w * 3. * b[KX] * Fx;
                    acc[10] += w * 3. * b[KY] * Fy;
                    acc[11] += w * 3. * b[KZ] * Fz;
                } else {
                    acc[9] += w * b[KYZ] * Fx;
                    acc[10] += w * b[KZX] * Fy;
                    acc[11] += w * b[KXY] * Fz;
                }
            }
        }
        __syncthreads();  // the next plane overwrites qp_s and pr_s
    }
    if (!live) return;  // no barrier below
    double intr[NB];
    intr[0] = 0.;
    intr[KX] = 0.25 * T->cb[KX] * dxi * acc[0];
    intr[KY] = 0.25 * T->cb[KY] * dyi * acc[1];
    intr[KZ] = 0.25 * T->cb[KZ] * dzi * acc[2];
    if constexpr (NB > 4) {
        intr[KYZ] = 0.25 * T->cb[KYZ] * (dyi * acc[3] + dzi * acc[4]);
        intr[KZX] = 0.25 * T->cb[KZX] * (dxi * acc[5] + dzi * acc[6]);
        intr[KXY] = 0.25 * T->cb[KXY] * (dxi * acc[7] + dyi * acc[8]);
        if constexpr (NB == 10) {
            intr[KXX] = 0.25 * T->cb[KXX] * (dxi * acc[9]);
            intr[KYY] = 0.25 * T->cb[KYY] * (dyi * acc[10]);
            intr[KZZ] = 0.25 * T->cb[KZZ] * (dzi * acc[11]);
        } else {
            intr[KXYZ] = 0.25 * T->cb[KXYZ] * (dxi * acc[9] + dyi * acc[10] + dzi * acc[11]);
        }
    }
    double glf[NB];
    glf[0] = 0.25 * t0;
#pragma unroll
    for (int r = 1; r < NB; r++) glf[r] = -intr[r];
    // ---- source: explicit relaxation (ivis 0), body force ----
    double src[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) src[r] = 0.;
    const double rate = f >= T_QX ? sp.coll_q : sp.coll;
    if (sp.ivis == 0 && f >= T_PXX && rate != 0.) {
#pragma unroll 1
        for (int r = 0; r < NB; r++) {
            double s = 0.;
#pragma unroll 1
            for (int p = 0; p < NPG; p++) {
                double qv = 0.;
#pragma unroll
                for (int rr = 0; rr < NB; rr++) qv += T->bint[p][rr] * mf[rr];
                s += T->bwint[p][r] * (-rate * qv);
            }
            src[r] = 0.125 * T->cb[r] * s;
        }
    }
    if (f >= MX && f <= MZ && sp.bf[f - MX] != 0.)
#pragma unroll
        for (int r = 0; r < NB; r++) src[r] += sp.bf[f - MX] * qin[qoff(g, c, RH, r)];
    if (f == EN && (sp.bf[0] != 0. || sp.bf[1] != 0. || sp.bf[2] != 0.))
#pragma unroll
        for (int r = 0; r < NB; r++)
            src[r] += sp.bf[0] * qin[qoff(g, c, MX, r)] + sp.bf[1] * qin[qoff(g, c, MY, r)] +
                      sp.bf[2] * qin[qoff(g, c, MZ, r)];
    // ---- update, SSP combination, NaN check of mode 1 ----
    if (sp.mode == 1) {
#pragma unroll
        for (int r = 0; r < NB; r++) qout[qoff(g, c, f, r)] = -(glf[r] - src[r]);
        return;
    }
    const double dt = *sp.dt;
    const double fac = f >= T_QX ? 1. / (1. + dt * sp.coll_q) : 1. / (1. + dt * sp.coll);
    bool bad = false;
#pragma unroll
    for (int r = 0; r < NB; r++) {
        double v = mf[r] - dt * (glf[r] - src[r]);
        if (sp.ivis == 1 && f >= T_PXX) v = fac * v;
        const long o = qoff(g, c, f, r);
        if (r == 0 && v != v) bad = true;
        if (sp.combine == 1) v = 0.5 * (qn[o] + v);
        else if (sp.combine == 2) v = 0.75 * qn[o] + 0.25 * v;
        else if (sp.combine == 3) v = (1. / 3.) * qn[o] + (2. / 3.) * v;
        qout[o] = v;
    }
    if (bad) {
        atomicOr(sp.err, 1);
        atomicMin(sp.bad, (unsigned long long)c);
    }
}

}  // namespace hd
