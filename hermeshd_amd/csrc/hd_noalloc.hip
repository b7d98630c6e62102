// hd_noalloc.hip -- hd_min_dt_static: the CFL time step of get_min_dt without a
// context and without allocating device memory.
//
// The wave-speed reduction writes one maximum per workgroup of 256 cells and a
// one-block kernel reduces those maxima.  Here they live in `scratch`, a
// file-scope device array of kScratchPartials doubles (512 KiB) that every GPU
// the library runs on holds until the process ends.  Its device address is
// looked up once per device (hipGetSymbolAddress, on the first call with that
// device current) and only that address is passed to kernels.  Every partial
// a call reads was written by the same call, so nothing has to be cleared;
// calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "hd_ctx.h"

namespace {

constexpr long kScratchPartials = 65536;  // one per 256 cells: grids up to 16,777,216 cells
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchPartials];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// device address of this device's `scratch` (looked up once per device)
int scratch_for_current_device(double** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return hd::fail(HD_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return HD_OK;
}

}  // namespace

extern "C" int hd_min_dt_static(const hd_config* cfg, const double* q, double* dt_dev, void* stream) {
    using namespace hd;
    if (!cfg || !q || !dt_dev) return fail(HD_ERR_INVALID, "NULL argument");
    int rc = validate(cfg);
    if (rc) return rc;
    const long N = (long)cfg->nx * cfg->ny * cfg->nz;
    const long nblk = (N + 255) / 256;
    if (nblk > kScratchPartials)
        return fail(HD_ERR_INVALID, "hd_min_dt_static handles at most " + std::to_string(kScratchPartials * 256) +
                                        " cells (its static scratch holds " + std::to_string(kScratchPartials) +
                                        " workgroup maxima); the grid has " + std::to_string(N));
    double* partial = nullptr;
    if ((rc = scratch_for_current_device(&partial))) return rc;
    Grid g;
    g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz; g.nQ = cfg->nQ; g.N = N;
    // init_spatial_params (src/initialize.f90:298-327), as hd_create computes them
    g.dxi = (double)cfg->nx / (cfg->lx / 2.0 - (-(cfg->lx / 2.0)));
    g.dyi = (double)cfg->ny / (cfg->ly / 2.0 - (-(cfg->ly / 2.0)));
    g.dzi = (double)cfg->nz / (cfg->lz / 2.0 - (-(cfg->lz / 2.0)));
    for (int s = 0; s < 6; s++) g.bc[s] = cfg->bc[s];
    CflEos e;
    e.gamma = cfg->gamma;
    e.ieos = cfg->ieos;
    e.P1 = cfg->eos_p1;
    e.T_floor = cfg->T_floor;
    hipStream_t s = (hipStream_t)stream;
    // one cell per thread: every workgroup writes its own partial, which the
    // finishing block reads back
    hipLaunchKernelGGL(k_wavespeed, dim3((unsigned)nblk), dim3(256), 0, s, q, g, e, 1.01 * cfg->rh_floor, partial);
    hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(256), 0, s, q, g, e, partial, (int)nblk, cfg->cflm, dt_dev,
                       (double*)nullptr);
    HIPCHK(hipGetLastError());
    return HD_OK;
}
