// hd_impl_vjp.h -- the stage transpose of Impl, out of the class so that a
// unit can leave its instantiation to another (hd_k4_vjp.hip).
#pragma once
#include "hd_impl.h"

namespace hd {

template <int NB, int QP, int NQ, int RM>
void Impl<NB, QP, NQ, RM>::stage_vjp(hd_ctx* c, const double* q, const double* lam, double* lam_in,
                                     const adj::AdjParams& ap, hipStream_t s) {
    if constexpr (NQ == 13) {
        // periodic grids: the split transpose (one wave per face point,
        // then one field per wave); bq2 exists only for those
        if (c->bq2[0]) {
            const Grid& g = c->grid;
            const long nfd[3] = {(long)(g.nx + 1) * g.ny * g.nz, (long)g.nx * (g.ny + 1) * g.nz,
                                 (long)g.nx * g.ny * (g.nz + 1)};
            constexpr int NT = QP * QP * 64;
            hipLaunchKernelGGL((adj::k_face13_vjp<NB, QP, RM, 0>), dim3(blocks_for(nfd[0], 64)), dim3(NT), 0, s, q,
                               lam, c->fl[0], c->bq2[0], g, c->ph, ap);
            hipLaunchKernelGGL((adj::k_face13_vjp<NB, QP, RM, 1>), dim3(blocks_for(nfd[1], 64)), dim3(NT), 0, s, q,
                               lam, c->fl[1], c->bq2[1], g, c->ph, ap);
            hipLaunchKernelGGL((adj::k_face13_vjp<NB, QP, RM, 2>), dim3(blocks_for(nfd[2], 64)), dim3(NT), 0, s, q,
                               lam, c->fl[2], c->bq2[2], g, c->ph, ap);
            c->vjp_blocks = (int)blocks_for(c->N, 64);
            hipLaunchKernelGGL((adj::k_cell13_vjp<NB, QP, RM>), dim3(c->vjp_blocks), dim3(CELL13_WAVES * 64), 0, s,
                               q, lam, lam_in, c->fl[0], c->fl[1], c->fl[2], c->bq2[0], c->bq2[1], c->bq2[2],
                               c->dtab, g, c->ph, ap);
            return;
        }
    }
    {
        // small grids are latency-bound: spread each cell over four waves
        if constexpr (adj::split_fits<NB, NQ>()) {
            if (c->N <= adj::kSplitCells) {
                c->vjp_blocks = (int)blocks_for(c->N, 64);
                hipLaunchKernelGGL((adj::k_stage_vjp_split<NB, QP, NQ, RM>), dim3(c->vjp_blocks), dim3(256), 0,
                                   s, q, lam, lam_in, c->grid, c->ph, ap);
                return;
            }
        }
        c->vjp_blocks = (int)blocks_for(c->N, 256);
        hipLaunchKernelGGL((adj::k_stage_vjp<NB, QP, NQ, RM>), dim3(c->vjp_blocks), dim3(256), 0, s, q, lam,
                           lam_in, c->grid, c->ph, ap);
    }
}

}  // namespace hd
