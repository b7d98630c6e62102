// hd_ctx.h -- the context behind the C ABI and the kernel dispatch table.
//
// A context owns the basis tables, geometry and scratch of one configuration
// (the reference keeps these as Fortran module globals; src/params.f90,
// src/basis_funcs.f90, src/boundary.f90, src/flux.f90:11-18).  The kernels of
// each (nbasis, iquad) family are instantiated in their own translation unit
// (hd_k4.hip, hd_k8.hip, hd_k10_5.hip, hd_k10_11.hip) so that they compile in
// parallel; hd_pick_* fills the dispatch table for one configuration.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/hermeshd_amd.h"
#include "hd_generic.h"

struct hd_ctx {
    hd_config cfg;
    int q, nb, nQ, nface;
    long N;
    hd::Grid grid;
    hd::Phys ph;
    hd::LimParams lp;
    hd::Tables htab;
    hd::Tables* dtab = nullptr;
    double* fl[3] = {nullptr, nullptr, nullptr};  // face fluxes per direction (table-driven kernels)
    double* partial = nullptr;                    // wave-speed partial maxima
    int npart = 0;
    int* err = nullptr;
    unsigned long long* bad = nullptr;
    double* scratch_dt = nullptr;  // holds 1.0: the time step hd_rhs gives the noise amplitude
    double* stage_buf = nullptr;   // third stage array of RK3 when the caller passes none
    // stochastic stress (cfg.noise): the step's field and the step counter
    unsigned long long* step_ctr = nullptr;
    double* noise_f[3] = {nullptr, nullptr, nullptr};
    double* noise_v = nullptr;
    long nfb = 0;                  // face lattice size (nx+1)(ny+1)(nz+1)
    double kap_f = 0., kap_v = 0.; // point amplitude factors of the face and volume rules
    bool generic = false;          // HD_KERNELS=generic selects the table-driven kernels
    int64_t bytes = 0;
    int device = 0;
    // stage timing (hd_timing_enable): event pairs recorded around each stage
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
    size_t ev_used = 0;
};

namespace hd {

struct Ops {
    void (*limiter)(hd_ctx*, double*, hipStream_t);
    void (*faces)(hd_ctx*, const double*, hipStream_t);
    void (*cell)(hd_ctx*, const double*, const double*, double*, const StageParams&, hipStream_t);
    void (*fused)(hd_ctx*, const double*, const double*, double*, const StageParams&, hipStream_t);
};

bool hd_pick_4(int nq, int rm, Ops* o);
bool hd_pick_8(int nq, int rm, Ops* o);
bool hd_pick_10_5(int rm, Ops* o);
bool hd_pick_10_11(int rm, Ops* o);

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace hd
