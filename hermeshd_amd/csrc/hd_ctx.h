// hd_ctx.h -- the context behind the C ABI and the kernel dispatch table.
//
// A context owns the basis tables, geometry and scratch of one configuration
// (the reference keeps these as Fortran module globals; src/params.f90,
// src/basis_funcs.f90, src/boundary.f90, src/flux.f90:11-18).  The kernels of
// each (nbasis, iquad) family are instantiated in their own translation unit
// (hd_k4.hip, hd_k8.hip, hd_k10_5.hip, hd_k10_11.hip, hd_k10_13.hip) so that they compile in
// parallel; hd_pick_* fills the dispatch table for one configuration.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/hermeshd_amd.h"
#include "hd_adjoint.h"
#include "hd_ensemble.h"

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
    // periodic 13-moment transpose: the right cells' face cotangents (the left
    // cells' reuse fl); allocated at the first such transpose
    double* bq2[3] = {nullptr, nullptr, nullptr};
    double* partial = nullptr;                    // wave-speed partial maxima
    int npart = 0;
    int* err = nullptr;
    unsigned long long* bad = nullptr;
    double* scratch_dt = nullptr;  // holds 1.0: the time step hd_rhs gives the noise amplitude
    double* stage_buf = nullptr;   // third stage array of RK3 when the caller passes none
    double* vjp_partial = nullptr; // per-block partial sums of the adjoint's parameter derivative
    int vjp_blocks = 0;            // blocks of the last transpose launch
    // stochastic stress (cfg.noise): the step's field and the step counter
    unsigned long long* step_ctr = nullptr;
    float* noise_f[3] = {nullptr, nullptr, nullptr};
    float* noise_v = nullptr;
    long nfb = 0;                  // face lattice size (nx+1)(ny+1)(nz+1)
    // x-low inflow exterior state (ny,nz,nface,nQ), set by hd_set_inflow
    double* inflow = nullptr;
    bool inflow_set = false;
    // immersed-solid mask (hd_set_mask), applied after every step
    unsigned char* mask = nullptr;
    bool mask_on = false;
    double rho_mask = 0.;
    double kap_f = 0., kap_v = 0.; // point amplitude factors of the face and volume rules
    bool generic = false;          // HD_KERNELS=generic selects the table-driven kernels
    // threads per workgroup of the fused stage kernel (HD_STAGE_BLOCK): one
    // wave per workgroup lets each SIMD start its next wave as soon as its
    // last one ends (the waves' run times differ with the limiter's early-out)
    int stage_bs = 64;
    // LDS bytes of the row-shared x-face stage kernel, 0 when it does not apply
    // (x not periodic, nx not a multiple of 64 up to 256, or too much LDS);
    // HD_STAGE_XS=0 switches it off
    unsigned xs_bytes = 0;
    // workgroup order of the row-shared kernel: z fastest (HD_STAGE_ROWS=zy, the
    // default) or y fastest (HD_STAGE_ROWS=yz)
    int rows_zfast = 1;
    // table-driven limiter kernel instead of fz::k_limit (HD_LIMITER=table)
    int table_limiter = 0;
    // one-thread-per-face k_face_flux for periodic 13-moment grids too (HD_FACES=table)
    int table_faces = 0;
    // fused per-cell transpose for periodic 13-moment grids too (HD_VJP=fused)
    int fused_vjp = 0;
    int64_t bytes = 0;
    int device = 0;
    // stage timing (hd_timing_enable): event pairs recorded around each stage
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending;
    size_t ev_used = 0;
};

// An ensemble: realizations of one configuration on boxes of different
// sizes, packed back to back (hd_ens_* entry points).
struct hd_ens {
    hd_config base;
    int nreal = 0, q = 0, nb = 0, nQ = 0, nface = 0, npg = 0;
    hd::Phys ph;
    hd::LimParams lp;
    double dxi = 0., dyi = 0., dzi = 0., kap_f = 0., kap_v = 0.;
    std::vector<hd::ens::EnsReal> hreal;
    std::vector<long> hpts;             // noise point prefix sums (nreal+1)
    std::vector<int2> hwork;
    hd::ens::EnsReal* dreal = nullptr;
    long* dpts = nullptr;
    int2* dwork = nullptr;
    size_t work_cap = 0;
    long state_doubles = 0, cells = 0, noise_floats = 0;
    float* noise = nullptr;
    double* one_dt = nullptr;          // 1.0: the noise time step of hd_ens_rhs
    double* stage_buf = nullptr;       // RK3 third array when the caller passes none
    unsigned long long* step_ctr = nullptr;
    int* err = nullptr;
    unsigned long long* bad = nullptr;
    // launch shape (hd_ens_set_launch / hd_ens_tune)
    int choice = -1, bs = 256, grid = 0, num_cus = 0;
    hd::ens::Launch launch{};
    // captured step (hd_ens_run)
    hipStream_t cap_stream = nullptr;
    // the next step's noise is drawn on side_stream while the captured
    // step's stages run (fork/join events inside the capture)
    hipStream_t side_stream = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    hipGraphExec_t gexec = nullptr;
    const void* gkey[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int gchoice = -1;
    int64_t bytes = 0;
    int device = 0;
};

namespace hd {

struct Ops {
    void (*limiter)(hd_ctx*, double*, hipStream_t);
    void (*faces)(hd_ctx*, const double*, hipStream_t);
    void (*cell)(hd_ctx*, const double*, const double*, double*, const StageParams&, hipStream_t);
    void (*fused)(hd_ctx*, const double*, const double*, double*, const StageParams&, hipStream_t);
    void (*ens_limit)(hd_ens*, double*, hipStream_t);
    void (*ens_stage)(hd_ens*, const double*, const double*, double*, const StageParams&, hipStream_t);
    // nullptr when the configuration has no adjoint (Roe)
    void (*stage_vjp)(hd_ctx*, const double*, const double*, double*, const adj::AdjParams&, hipStream_t);
};

// error text of the calling thread (hd_last_error) and the error return
extern thread_local std::string g_last_error;
int fail(int code, const std::string& msg);
#define HIPCHK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return hd::fail(HD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
// checks a configuration (the per-realization extents of an ensemble included)
int validate(const hd_config* c);
// 1/sqrt(sum_p (w_p/W)^2) of the 2-D (face) or 3-D (volume) product rule
double rule_kappa(const Tables& t, int dims);

// dispatch table of one (nbasis, iquad, nQ, Riemann solver); false if not built
bool pick_ops_for(int nb, int q, int nq, int rm, Ops* o);
bool hd_pick_4(int nq, int rm, Ops* o);
bool hd_pick_8(int nq, int rm, Ops* o);
bool hd_pick_10_5(int rm, Ops* o);
bool hd_pick_10_11(int rm, Ops* o);
bool hd_pick_10_13(int rm, Ops* o);

inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace hd
