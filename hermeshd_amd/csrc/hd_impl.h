// hd_impl.h -- launch wrappers of one (nbasis, iquad, nQ, Riemann solver)
// configuration; included by the instantiation units only.
#pragma once
#include "hd_ctx.h"

namespace hd {

template <int NB, int QP, int NQ, int RM>
struct Impl {
    static void limiter(hd_ctx* c, double* q, hipStream_t s) {
        // compile-time face tables and an exact early-out for cells no floor
        // can reach; the table-driven k_limiter stays for HD_LIMITER=table
        if (c->generic && c->table_limiter)
            hipLaunchKernelGGL((k_limiter<NB, QP>), dim3(blocks_for(c->N, 256)), dim3(256), 0, s, q, c->dtab, c->grid, c->lp);
        else
            hipLaunchKernelGGL((fz::k_limit<NB, QP>), dim3(blocks_for(c->N, 256)), dim3(256), 0, s, q, c->grid, c->lp);
    }
    static void faces(hd_ctx* c, const double* q, hipStream_t s) {
        const Grid& g = c->grid;
        long nfx = (long)(g.nx + 1) * g.ny * g.nz, nfy = (long)g.nx * (g.ny + 1) * g.nz, nfz = (long)g.nx * g.ny * (g.nz + 1);
        if constexpr (NQ == 13) {
            // all-periodic grids: one wave per face point (fz::k_face13)
            bool periodic = !c->table_faces;
            for (int b = 0; b < 6; b++) periodic = periodic && g.bc[b] == BC_PERIODIC;
            if (periodic) {
                constexpr int NT = QP * QP * 64;
                hipLaunchKernelGGL((fz::k_face13<NB, QP, RM, 0>), dim3(blocks_for(nfx, 64)), dim3(NT), 0, s, q, c->fl[0], g, c->ph);
                hipLaunchKernelGGL((fz::k_face13<NB, QP, RM, 1>), dim3(blocks_for(nfy, 64)), dim3(NT), 0, s, q, c->fl[1], g, c->ph);
                hipLaunchKernelGGL((fz::k_face13<NB, QP, RM, 2>), dim3(blocks_for(nfz, 64)), dim3(NT), 0, s, q, c->fl[2], g, c->ph);
                return;
            }
        }
        hipLaunchKernelGGL((k_face_flux<NB, QP, NQ, RM, 0>), dim3(blocks_for(nfx, 128)), dim3(128), 0, s, q, c->fl[0], c->dtab, g, c->ph, c->inflow);
        hipLaunchKernelGGL((k_face_flux<NB, QP, NQ, RM, 1>), dim3(blocks_for(nfy, 128)), dim3(128), 0, s, q, c->fl[1], c->dtab, g, c->ph, c->inflow);
        hipLaunchKernelGGL((k_face_flux<NB, QP, NQ, RM, 2>), dim3(blocks_for(nfz, 128)), dim3(128), 0, s, q, c->fl[2], c->dtab, g, c->ph, c->inflow);
    }
    static void cell(hd_ctx* c, const double* qin, const double* qn, double* qout, const StageParams& sp, hipStream_t s) {
        if constexpr (NQ == 13) {
            // one field per wave, 64 cells per workgroup (every wave of a
            // workgroup reaches its barriers: the grid is padded, not clipped);
            // face moments from k_face13 on all-periodic grids (as faces())
            bool periodic = !c->table_faces;
            for (int b = 0; b < 6; b++) periodic = periodic && c->grid.bc[b] == BC_PERIODIC;
            if (periodic)
                hipLaunchKernelGGL((k_cell13<NB, QP, true>), dim3(blocks_for(c->N, 64)), dim3(CELL13_WAVES * 64), 0,
                                   s, qin, qn, qout, c->fl[0], c->fl[1], c->fl[2], c->dtab, c->grid, c->ph, sp);
            else
                hipLaunchKernelGGL((k_cell13<NB, QP, false>), dim3(blocks_for(c->N, 64)), dim3(CELL13_WAVES * 64), 0,
                                   s, qin, qn, qout, c->fl[0], c->fl[1], c->fl[2], c->dtab, c->grid, c->ph, sp);
            return;
        }
        hipLaunchKernelGGL((k_cell<NB, QP, NQ>), dim3(blocks_for(c->N, 128)), dim3(128), 0, s, qin, qn, qout, c->fl[0],
                           c->fl[1], c->fl[2], c->dtab, c->grid, c->ph, sp);
    }
    static void fused(hd_ctx* c, const double* qin, const double* qn, double* qout, const StageParams& sp, hipStream_t s) {
        const int bs = c->stage_bs;
        if (c->xs_bytes) {
            // one periodic row per workgroup, x faces shared through LDS
            const unsigned rows = (unsigned)(c->N / c->grid.nx);
            if constexpr (NQ >= 11) {
                if (sp.noise[0]) {
                    hipLaunchKernelGGL((fz::k_stage_xs<NB, QP, NQ, RM, true>), dim3(rows), dim3(c->grid.nx),
                                       c->xs_bytes, s, qin, qn, qout, c->grid, c->ph, sp);
                    return;
                }
            }
            hipLaunchKernelGGL((fz::k_stage_xs<NB, QP, NQ, RM, false>), dim3(rows), dim3(c->grid.nx), c->xs_bytes, s,
                               qin, qn, qout, c->grid, c->ph, sp);
            return;
        }
        if constexpr (NQ >= 11) {
            if (sp.noise[0]) {
                hipLaunchKernelGGL((fz::k_stage<NB, QP, NQ, RM, true>), dim3(blocks_for(c->N, bs)), dim3(bs), 0, s,
                                   qin, qn, qout, c->grid, c->ph, sp);
                return;
            }
        }
        hipLaunchKernelGGL((fz::k_stage<NB, QP, NQ, RM, false>), dim3(blocks_for(c->N, bs)), dim3(bs), 0, s, qin, qn,
                           qout, c->grid, c->ph, sp);
    }
    static void ens_limit(hd_ens* e, double* q, hipStream_t s) {
        hipLaunchKernelGGL((ens::k_ens_limit<NB, QP>), dim3(e->grid), dim3(e->bs), 0, s, q, e->dreal, e->dwork,
                           e->launch, e->lp);
    }
    static void ens_stage(hd_ens* e, const double* qin, const double* qn, double* qout, const StageParams& sp,
                          hipStream_t s) {
        if constexpr (NQ >= 11) {
            if (e->base.noise) {
                hipLaunchKernelGGL((ens::k_ens_stage<NB, QP, NQ, RM, true>), dim3(e->grid), dim3(e->bs), 0, s, qin,
                                   qn, qout, e->dreal, e->dwork, e->launch, e->ph, sp, e->noise, e->step_ctr,
                                   e->noise_floats);
                return;
            }
        }
        hipLaunchKernelGGL((ens::k_ens_stage<NB, QP, NQ, RM, false>), dim3(e->grid), dim3(e->bs), 0, s, qin, qn,
                           qout, e->dreal, e->dwork, e->launch, e->ph, sp, e->noise, e->step_ctr, e->noise_floats);
    }
    // defined in hd_impl_vjp.h; the units that include it instantiate it
    // there, and hd_k4_vjp.hip holds the nbasis 4 instances, so that those
    // compile with the default scheduler while hd_k4.hip takes another
    static void stage_vjp(hd_ctx* c, const double* q, const double* lam, double* lam_in, const adj::AdjParams& ap,
                          hipStream_t s);
    static Ops ops() {
        // the 13-moment model runs on its own face and cell kernels (no fused
        // stage, no ensembles); its transpose is split on periodic grids
        if constexpr (NQ == 13) return Ops{&limiter, &faces, &cell, nullptr, nullptr, nullptr, &stage_vjp};
        else return Ops{&limiter, &faces, &cell, &fused, &ens_limit, &ens_stage, &stage_vjp};
    }
};

template <int NB, int QP, int NQ>
bool pick_rm(int rm, Ops* o) {
    switch (rm) {
        case RIEMANN_CENTRAL: *o = Impl<NB, QP, NQ, RIEMANN_CENTRAL>::ops(); return true;
        case RIEMANN_HLLC: *o = Impl<NB, QP, NQ, RIEMANN_HLLC>::ops(); return true;
        case RIEMANN_ROE: *o = Impl<NB, QP, NQ, RIEMANN_ROE>::ops(); return true;
    }
    return false;
}

}  // namespace hd
