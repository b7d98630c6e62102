// hd_ensemble.h -- many independent realizations of one configuration, on
// boxes of different sizes, advanced by one launch per RK stage.
//
// The reference runs one realization per process (src/main.f90 time loop,
// src/random.f90 noise); here the realizations' states are packed back to
// back, each in the reference layout Q(nx,ny,nz,nQ,nbasis), and a persistent
// kernel walks a (realization, tile) work list.  A tile is blockDim.x cells:
// either a run of consecutive cells (linear) or a bx*by*bz brick.  With
// `xcd_chunk` the list is cut into 8 contiguous chunks, one per XCD, so that
// a realization's tiles share one L2 (workgroup b runs on XCD b mod 8).
// Noise is drawn per realization with the same counter layout as one
// context, keyed by the realization id, so results do not depend on the
// schedule.
#pragma once
#include "hd_fused.h"
#include "hd_noise.h"

namespace hd {
namespace ens {

struct EnsReal {
    Grid g;
    long qoff;     // offset of the realization's state in a packed array (doubles)
    long noff;     // offset of its noise block: 3 face families of 6*nface*nfb, then 6*npg*N volume
    long nfb;      // (nx+1)(ny+1)(nz+1)
    long cell0;    // first cell of the realization in the packed cell numbering
    int ntx, nty, ntz;  // tiles per axis (brick tiles)
    unsigned int realization;
};

struct Launch {
    int bx, by, bz;   // brick shape; bx*by*bz = blockDim.x
    int linear;       // 1: tiles are runs of blockDim.x consecutive cells
    int xcd_chunk;    // 1: XCD-contiguous chunks of the work list
    int nwork;
};

// cell of thread threadIdx.x in tile t of realization R; false when outside
__device__ __forceinline__ bool tile_cell(const EnsReal& R, const Launch& L, int t, fz::CellPos& cp) {
    const Grid& g = R.g;
    if (L.linear) {
        const long c = (long)t * blockDim.x + threadIdx.x;
        if (c >= g.N) return false;
        cp.c = c;
        cp.i = (int)(c % g.nx);
        cp.j = (int)((c / g.nx) % g.ny);
        cp.k = (int)(c / ((long)g.nx * g.ny));
        return true;
    }
    const int tx = t % R.ntx, ty = (t / R.ntx) % R.nty, tz = t / (R.ntx * R.nty);
    const int lx = threadIdx.x % L.bx, ly = (threadIdx.x / L.bx) % L.by, lz = threadIdx.x / (L.bx * L.by);
    cp.i = tx * L.bx + lx;
    cp.j = ty * L.by + ly;
    cp.k = tz * L.bz + lz;
    if (cp.i >= g.nx || cp.j >= g.ny || cp.k >= g.nz) return false;
    cp.c = cp.i + (long)g.nx * (cp.j + (long)g.ny * cp.k);
    return true;
}

// Walks the work list: plain striding, or 8 XCD-contiguous chunks.
template <typename F>
__device__ __forceinline__ void for_each_item(const Launch& L, F&& f) {
    if (L.xcd_chunk) {
        const int x = blockIdx.x & 7, per_xcd = (int)(gridDim.x >> 3), slot = (int)(blockIdx.x >> 3);
        const int lo = (int)(((long)L.nwork * x) >> 3), hi = (int)(((long)L.nwork * (x + 1)) >> 3);
        for (int w = lo + slot; w < hi; w += per_xcd) f(w);
    } else {
        for (int w = blockIdx.x; w < L.nwork; w += gridDim.x) f(w);
    }
}

template <int NB, int QP, int NQ, int RM, bool NZ>
__global__ __launch_bounds__(256) void k_ens_stage(const double* __restrict__ qin, const double* __restrict__ qn,
                                                   double* __restrict__ qout, const EnsReal* __restrict__ reals,
                                                   const int2* __restrict__ work, Launch L, Phys ph,
                                                   StageParams sp, const float* __restrict__ noise,
                                                   const unsigned long long* __restrict__ step, long nfloats) {
    constexpr int NF = QP * QP;
    for_each_item(L, [&](int w) {
        const int2 wk = work[w];
        const EnsReal R = reals[wk.x];
        fz::CellPos cp;
        if (!tile_cell(R, L, wk.y, cp)) return;
        StageParams s = sp;
        if constexpr (NZ) {
            const long fsz = 6L * NF * R.nfb;
            // the field of the current step: buffer (step & 1)
            const float* base = noise + ((*step & 1ull) ? nfloats : 0L) + R.noff;
            s.noise[0] = base;
            s.noise[1] = base + fsz;
            s.noise[2] = base + 2 * fsz;
            s.noise[3] = base + 3 * fsz;
            s.nfb = R.nfb;
        }
        fz::stage_cell<NB, QP, NQ, RM, NZ>(qin + R.qoff, qn + R.qoff, qout + R.qoff, R.g, ph, s, cp, R.cell0 + cp.c);
    });
}

template <int NB, int QP>
__global__ __launch_bounds__(256) void k_ens_limit(double* __restrict__ q, const EnsReal* __restrict__ reals,
                                                   const int2* __restrict__ work, Launch L, LimParams lp) {
    for_each_item(L, [&](int w) {
        const int2 wk = work[w];
        const EnsReal R = reals[wk.x];
        fz::CellPos cp;
        if (!tile_cell(R, L, wk.y, cp)) return;
        double* qr = q + R.qoff;
        const long N = R.g.N;
        const int nQ = R.g.nQ;
        double m[5][NB];
#pragma unroll
        for (int r = 0; r < NB; r++)
#pragma unroll
            for (int f = 0; f < 5; f++) m[f][r] = qr[((long)r * nQ + f) * N + cp.c];
        if (fz::limit_cell<NB, QP, 5>(m, lp)) {
#pragma unroll
            for (int r = 0; r < NB; r++)
#pragma unroll
                for (int f = 0; f < 5; f++) qr[((long)r * nQ + f) * N + cp.c] = m[f][r];
        }
    });
}

// Noise of every realization for the current step: flat over all points;
// `pts` holds the prefix sums of the per-realization point counts
// (3*nface*nfb + npg*N), each rounded up to a whole number of 256-thread
// blocks so that a block lies in one realization (nreal+1 entries).
constexpr long NOISE_BLOCK = 256;
struct EnsNoise {
    unsigned long long seed;
    const unsigned long long* step;
    int ahead;       // draws step *step + ahead
    long nfloats;    // floats of one step's field; step s lives in buffer s & 1
    const double* dt;
    double eta_sd, zeta_sd, dVi, kap_f, kap_v;
    int nface, npg, nreal;
    const long* pts;
    const EnsReal* reals;
    float* out;
};

// One chunk of NOISE_BLOCK points (chunk b covers points b*NOISE_BLOCK ...).
__device__ __forceinline__ void ens_noise_chunk(const EnsNoise& np, long b) {
    // the chunk's realization: a search on its first point, uniform over the
    // block
    const long t0 = b * NOISE_BLOCK;
    if (t0 >= np.pts[np.nreal]) return;
    int lo = 0, hi = np.nreal;  // pts[lo] <= t0 < pts[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (np.pts[mid] <= t0) lo = mid; else hi = mid;
    }
    const EnsReal& R = np.reals[lo];
    const long u = t0 - np.pts[lo] + threadIdx.x;
    const long nf = (long)np.nface * R.nfb, nv = (long)np.npg * R.g.N;
    if (u >= 3 * nf + nv) return;  // the padding of the last block
    const double sq = sqrt(np.dVi / *np.dt);
    const unsigned long long st = *np.step + (unsigned long long)np.ahead;
    long id, count;
    unsigned int stream;
    float* out = np.out + ((st & 1ull) ? np.nfloats : 0L) + R.noff;
    double kap;
    if (u < 3 * nf) {
        const int d = (u >= nf) + (u >= 2 * nf);
        id = u - (long)d * nf;
        count = nf;
        stream = (unsigned int)(d + 1);
        out += (long)d * 6 * nf;
        kap = np.kap_f;
    } else {
        id = u - 3 * nf;
        count = nv;
        stream = 4u;
        out += 3L * 6 * nf;
        kap = np.kap_v;
    }
    double S[6];
    noise::stress(np.seed, R.realization, st, stream, (unsigned long long)id, np.eta_sd * sq * kap,
                  np.zeta_sd * sq * kap, S);
#pragma unroll
    for (int e = 0; e < 6; e++) out[(long)e * count + id] = (float)S[e];  // single precision, as k_noise
}

// Blocks stride over the chunks: one chunk per block when the grid covers
// them all, or a small grid (one block per CU) that stays beside a running
// stage kernel and draws the next step's field in the stage's shadow.
static __global__ __launch_bounds__(NOISE_BLOCK) void k_ens_noise(EnsNoise np) {
    const long nchunk = np.pts[np.nreal] / NOISE_BLOCK;
    for (long b = blockIdx.x; b < nchunk; b += gridDim.x) ens_noise_chunk(np, b);
}

}  // namespace ens
}  // namespace hd
