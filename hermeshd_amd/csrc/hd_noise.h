// hd_noise.h -- the stochastic stress field of one time step.
//
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11) with counter (index lo, index hi | stream<<24,
// step lo, step hi) and key = seed mixed with the realization id; uniforms
// (u + 0.5) 2^-32 and Box-Muller pairs in single precision.  A value depends only on (seed,
// realization, step, stream, index), never on the launch shape.
//
// Stress at one point (random_stresses_pnts_r, src/random.f90:106-145, with
// the Landau-Lifshitz covariance): six normals g0..g5 from counters 2*id and
// 2*id+1; diagonal ae*(sqrt2*g_i - trace/3) + az*(g0+g1+g2), off-diagonals
// xy, xz, yz = ae*g3, ae*g4, ae*g5.  ae = eta_sd*sqrt(dVi/dt)*kappa with
// kappa = 1/sqrt(sum_p (w_p/W)^2) over the face (or volume) rule, so that a
// face- or cell-averaged stress has variance eta_sd^2/(dV dt).
//
// k_noise draws the whole step's field once: every face point of the three
// face families (stream 1..3, id = p*NFB + fid, fid = i + (nx+1)(j + (ny+1)k)
// over the (nx+1)(ny+1)(nz+1) face lattice) and every volume point (stream 4,
// id = p*N + cell).  Each point is drawn once and read by every stage of the
// step and by both cells of a face.  Layout: component-major,
// out[e*count + id], so that neighbouring cells read neighbouring words.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hd {
namespace noise {

__device__ __forceinline__ void philox4x32_10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0,
                                              uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
        const uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += W0; k1 += W1;
    }
}

// Single-precision Box-Muller (the reference draws its Gaussians in single
// precision, vsRngGaussian, src/random.f90:61,87).  The logarithm and the
// sine/cosine are written out here operation by operation, with no
// contraction into fused multiply-adds, from correctly rounded FP32
// operations only (+, -, *, /, sqrt, conversions), so that any IEEE FP32
// implementation of the same steps gives the same bits.
//   log a = e ln2 + 2 atanh(s),  a = m 2^e, m in [sqrt(1/2), sqrt(2)),
//           s = (m-1)/(m+1), atanh by its odd series to s^9 (|s| < 0.172)
//   sin/cos(pi x): x = n/2 + r, |r| <= 1/4 (quadrant n), Taylor series of
//           sin and cos of pi r to degree 9 and 10, then the quadrant's signs
__device__ __forceinline__ float bm_logf(float a) {
#pragma clang fp contract(off)
    uint32_t bits = __float_as_uint(a);
    int e = (int)((bits >> 23) & 0xffu) - 127;
    float m = __uint_as_float((bits & 0x007fffffu) | 0x3f800000u);
    if (m > 0x1.6a09e6p+0f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 0x1.c71c72p-4f;       // 1/9
    p = p * s2 + 0x1.24924ap-3f;    // 1/7
    p = p * s2 + 0x1.99999ap-3f;    // 1/5
    p = p * s2 + 0x1.555556p-2f;    // 1/3
    p = p * s2 + 1.0f;
    const float lm = 2.0f * s * p;
    return (float)e * 0x1.62e430p-1f + lm;  // ln 2
}

__device__ __forceinline__ void bm_sincospi(float x, float* sv, float* cv) {
#pragma clang fp contract(off)
    const float nf = floorf(2.0f * x + 0.5f);
    const int n = (int)nf;
    const float r = x - nf * 0.5f;
    const float z = r * 0x1.921fb6p+1f;  // pi
    const float z2 = z * z;
    float sp = 0x1.71de3ap-19f;          // 1/9!
    sp = sp * z2 - 0x1.a01a02p-13f;      // 1/7!
    sp = sp * z2 + 0x1.111112p-7f;       // 1/5!
    sp = sp * z2 - 0x1.555556p-3f;       // 1/3!
    sp = sp * z2 + 1.0f;
    const float sn = z * sp;
    float cp = -0x1.27e4fcp-22f;         // 1/10!
    cp = cp * z2 + 0x1.a01a02p-16f;      // 1/8!
    cp = cp * z2 - 0x1.6c16c2p-10f;      // 1/6!
    cp = cp * z2 + 0x1.555556p-5f;       // 1/4!
    cp = cp * z2 - 0.5f;
    const float cs = cp * z2 + 1.0f;
    const int qd = n & 3;
    *sv = qd == 0 ? sn : (qd == 1 ? cs : (qd == 2 ? -sn : -cs));
    *cv = qd == 0 ? cs : (qd == 1 ? -sn : (qd == 2 ? -cs : sn));
}

__device__ __forceinline__ void normals4(uint64_t seed, uint32_t realization, uint64_t step, uint32_t stream,
                                         uint64_t index, double* out) {
#pragma clang fp contract(off)
    uint32_t c0 = (uint32_t)index, c1 = (uint32_t)(index >> 32) | (stream << 24);
    uint32_t c2 = (uint32_t)step, c3 = (uint32_t)(step >> 32);
    philox4x32_10(c0, c1, c2, c3, (uint32_t)seed ^ (realization * 0x9E3779B9u), (uint32_t)(seed >> 32) ^ realization);
    const uint32_t u[4] = {c0, c1, c2, c3};
#pragma unroll
    for (int h = 0; h < 2; h++) {
        // uniforms in (0, 1]: (u + 1/2) 2^-32 in FP32
        const float a = ((float)u[2 * h] + 0.5f) * 0x1p-32f;
        const float b = ((float)u[2 * h + 1] + 0.5f) * 0x1p-32f;
        const float rad = sqrtf(-2.0f * bm_logf(a));
        float sv, cv;
        bm_sincospi(2.0f * b, &sv, &cv);
        out[2 * h] = (double)(rad * cv);
        out[2 * h + 1] = (double)(rad * sv);
    }
}

// stress (xx, yy, zz, xy, xz, yz) at point id of a stream
__device__ __forceinline__ void stress(uint64_t seed, uint32_t realization, uint64_t step, uint32_t stream, uint64_t id,
                                       double ae, double az, double S[6]) {
    double g[8];
    normals4(seed, realization, step, stream, 2 * id, g);
    normals4(seed, realization, step, stream, 2 * id + 1, g + 4);
    const double s2 = 1.4142135623730951;
    const double dxx = s2 * g[0], dyy = s2 * g[1], dzz = s2 * g[2];
    const double tr = (dxx + dyy + dzz) / 3.0;
    const double trW = g[0] + g[1] + g[2];
    S[0] = ae * (dxx - tr) + az * trW;
    S[1] = ae * (dyy - tr) + az * trW;
    S[2] = ae * (dzz - tr) + az * trW;
    S[3] = ae * g[3];
    S[4] = ae * g[4];
    S[5] = ae * g[5];
}

struct NoiseParams {
    unsigned long long seed;
    unsigned int realization;
    int nface, npg;
    const unsigned long long* step;  // device step counter
    const double* dt;                // device time step
    double eta_sd, zeta_sd, dVi, kap_f, kap_v;
    long nfb, N;                     // face lattice size, cells
    float* face[3];                  // [6][nface][nfb] per face family
    float* vol;                      // [6][npg][N]
};

static __global__ __launch_bounds__(256) void k_noise(NoiseParams np) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nf = (long)np.nface * np.nfb, nv = (long)np.npg * np.N;
    if (t >= 3 * nf + nv) return;
    const double sq = sqrt(np.dVi / *np.dt);
    const uint64_t step = *np.step;
    long id, count;
    uint32_t stream;
    float* out;
    double kap;
    if (t < 3 * nf) {
        const int d = (t >= nf) + (t >= 2 * nf);
        id = t - (long)d * nf;
        count = nf;
        stream = (uint32_t)(d + 1);
        out = np.face[d];
        kap = np.kap_f;
    } else {
        id = t - 3 * nf;
        count = nv;
        stream = 4u;
        out = np.vol;
        kap = np.kap_v;
    }
    double S[6];
    stress(np.seed, np.realization, step, stream, (uint64_t)id, np.eta_sd * sq * kap, np.zeta_sd * sq * kap, S);
    // stored in single precision, the reference's precision for its
    // Gaussian draws (vsRngGaussian, src/random.f90:61,87)
#pragma unroll
    for (int e = 0; e < 6; e++) out[(long)e * count + id] = (float)S[e];
}

static __global__ void k_set_step(unsigned long long* step, unsigned long long value, int advance) {
    if (threadIdx.x == 0) *step = advance ? *step + 1ull : value;
}

}  // namespace noise
}  // namespace hd
