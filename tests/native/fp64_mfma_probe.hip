// FP64 throughput on gfx950: chains of VALU FMAs against chains of
// v_mfma_f64_16x16x4f64, each alone, both in one wave, and in alternating
// waves.  4096 one-wave workgroups (4 waves per SIMD); each kernel is timed
// with HIP events after a short warm-up launch and reported in TFLOP/s
// (2 flops per FMA, 2*16*16*4 per MFMA).  It measures whether routing FP64
// contractions through the matrix cores can beat the vector FMA rate.
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -o fp64_mfma_probe tests/native/fp64_mfma_probe.hip && ./fp64_mfma_probe
#include <hip/hip_runtime.h>
#include <cstdio>
typedef double double4_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void k_valu(double* out, int iters, double x) {
    double a[8];
    for (int i = 0; i < 8; i++) a[i] = x + threadIdx.x + i;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) a[i] = fma(a[i], 0.999999, 1e-7);
    }
    double s = 0; for (int i = 0; i < 8; i++) s += a[i];
    out[blockIdx.x * 64 + threadIdx.x] = s;
}

__global__ __launch_bounds__(64) void k_mfma(double* out, int iters, double x) {
    double4_t c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
    double a = x + threadIdx.x * 1e-3, b = 1.0 - threadIdx.x * 1e-4;
    for (int it = 0; it < iters; it++) {
        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, b, c3, 0, 0, 0);
    }
    double s = 0;
    for (int i = 0; i < 4; i++) s += c0[i] + c1[i] + c2[i] + c3[i];
    out[blockIdx.x * 64 + threadIdx.x] = s;
}

// both in one wave: 4 MFMA + 8*R VALU FMA per iteration
template <int R>
__global__ __launch_bounds__(64) void k_both(double* out, int iters, double x) {
    double4_t c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
    double a = x + threadIdx.x * 1e-3, b = 1.0 - threadIdx.x * 1e-4;
    double v[8];
    for (int i = 0; i < 8; i++) v[i] = x + threadIdx.x + i;
    for (int it = 0; it < iters; it++) {
        c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, b, c3, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int i = 0; i < 8; i++) v[i] = fma(v[i], 0.999999, 1e-7);
    }
    double s = 0;
    for (int i = 0; i < 4; i++) s += c0[i] + c1[i] + c2[i] + c3[i];
    for (int i = 0; i < 8; i++) s += v[i];
    out[blockIdx.x * 64 + threadIdx.x] = s;
}

// mixed waves: even blocks MFMA, odd blocks VALU
__global__ __launch_bounds__(64) void k_mixed(double* out, int iters, double x) {
    if (blockIdx.x & 1) {
        double a[8];
        for (int i = 0; i < 8; i++) a[i] = x + threadIdx.x + i;
        for (int it = 0; it < 4 * iters; it++) {
#pragma unroll
            for (int i = 0; i < 8; i++) a[i] = fma(a[i], 0.999999, 1e-7);
        }
        double s = 0; for (int i = 0; i < 8; i++) s += a[i];
        out[blockIdx.x * 64 + threadIdx.x] = s;
    } else {
        double4_t c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
        double a = x + threadIdx.x * 1e-3, b = 1.0 - threadIdx.x * 1e-4;
        for (int it = 0; it < iters; it++) {
            c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, c2, 0, 0, 0);
            c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, b, c3, 0, 0, 0);
        }
        double s = 0;
        for (int i = 0; i < 4; i++) s += c0[i] + c1[i] + c2[i] + c3[i];
        out[blockIdx.x * 64 + threadIdx.x] = s;
    }
}

int main() {
    const int blocks = 256 * 4 * 4;  // 4 waves per SIMD
    double* out;
    if (hipMalloc(&out, (size_t)blocks * 64 * sizeof(double)) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    auto run = [&](const char* name, auto kern, int iters, double flop_per_iter_per_wave) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(64), 0, 0, out, 10, 1.0);
        (void)hipEventRecord(e0, 0);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(64), 0, 0, out, iters, 1.0);
        (void)hipEventRecord(e1, 0);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        const double flop = flop_per_iter_per_wave * iters * blocks;
        printf("%-22s %8.3f ms  %7.2f TFLOP/s\n", name, ms, flop / (ms * 1e-3) / 1e12);
    };
    const int it = 20000;
    run("valu fma", k_valu, it, 8.0 * 2 * 64);
    run("mfma f64 16x16x4", k_mfma, it, 4.0 * 2 * 16 * 16 * 4);
    run("both, 1 wave (8 fma)", k_both<1>, it, 4.0 * 2 * 16 * 16 * 4 + 8.0 * 2 * 64);
    run("both, 1 wave (32 fma)", k_both<4>, it, 4.0 * 2 * 16 * 16 * 4 + 32.0 * 2 * 64);
    run("mixed waves", k_mixed, it, 0.5 * (4.0 * 2 * 16 * 16 * 4) + 0.5 * (4 * 8.0 * 2 * 64));
    (void)hipFree(out);
    return 0;
}
