// hd_capi.hip -- C ABI of the HERMESHD hot path on MI355X (gfx950).
//
// All entry points are stream-ordered and return an error code; none aborts.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include <cmath>
#include <cstdlib>

#include "hd_ctx.h"
#include "hd_noise.h"

namespace hd {
thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace hd

namespace hd {
// dispatch over (nbasis, iquad) x nQ x Riemann solver
bool pick_ops_for(int nb, int q, int nq, int rm, Ops* o) {
    if (nb == 4 && q == 2) return hd_pick_4(nq, rm, o);
    if (nb == 8 && q == 2) return hd_pick_8(nq, rm, o);
    if (nb == 10 && q == 3 && nq == 5) return hd_pick_10_5(rm, o);
    if (nb == 10 && q == 3 && nq == 11) return hd_pick_10_11(rm, o);
    if (nb == 10 && q == 3 && nq == 13) return hd_pick_10_13(rm, o);
    return false;
}

int validate(const hd_config* c) {
    if (!c) return fail(HD_ERR_INVALID, "config is NULL");
    if (c->nx < 1 || c->ny < 1 || c->nz < 1) return fail(HD_ERR_INVALID, "grid extents must be >= 1");
    const bool ok = (c->nbasis == 4 && c->iquad == 2) || (c->nbasis == 8 && c->iquad == 2) ||
                    (c->nbasis == 10 && c->iquad == 3);
    if (!ok) return fail(HD_ERR_INVALID, "(nbasis, iquad) must be (4,2), (8,2) or (10,3)");
    if (c->nQ != 5 && c->nQ != 11 && c->nQ != 13) return fail(HD_ERR_INVALID, "nQ must be 5, 11 or 13");
    if (c->ivis < 0 || c->ivis > 2) return fail(HD_ERR_INVALID, "ivis must be 0, 1 or 2");
    if (c->ivis >= 1 && c->nQ == 5) return fail(HD_ERR_INVALID, "ivis=1,2 need the stress fields (nQ=11 or 13)");
    if (c->ivis == 2 && c->nQ != 11) return fail(HD_ERR_INVALID, "ivis=2 needs the second moments (nQ=11)");
    if (c->nQ == 13 && (c->noise || c->ieos != 1))
        return fail(HD_ERR_INVALID, "the 13-moment model runs with an ideal gas and no noise");
    // the 13-moment kernels index cells and faces in 32 bits
    if (c->nQ == 13 && (long)(c->nx + 1) * (c->ny + 1) * (c->nz + 1) > 2147483647L)
        return fail(HD_ERR_INVALID, "the 13-moment model takes at most 2^31 - 1 faces per direction");
    if (c->ivis == 2 && c->noise) return fail(HD_ERR_INVALID, "ivis=2 has no stochastic stress");
    if (c->ieos != 1 && c->ieos != 2) return fail(HD_ERR_INVALID, "ieos must be 1 (ideal gas) or 2 (water)");
    if (c->ieos == 2 && c->riemann == HD_RIEMANN_ROE)
        return fail(HD_ERR_INVALID, "ieos=2 runs with the HLLC or central flux");
    if (c->riemann < 0 || c->riemann > 2) return fail(HD_ERR_INVALID, "riemann must be 0, 1 or 2");
    if (c->integrator != 2 && c->integrator != 3) return fail(HD_ERR_INVALID, "integrator must be 2 or 3");
    for (int d = 0; d < 3; d++) {
        for (int s = 0; s < 2; s++)
            if (c->bc[2 * d + s] < 0 || c->bc[2 * d + s] > 4) return fail(HD_ERR_INVALID, "unknown boundary kind");
        if ((c->bc[2 * d] == HD_BC_PERIODIC) != (c->bc[2 * d + 1] == HD_BC_PERIODIC))
            return fail(HD_ERR_INVALID, "a periodic axis must be periodic on both sides");
    }
    for (int s = 1; s < 6; s++)
        if (c->bc[s] == HD_BC_INFLOW) return fail(HD_ERR_INVALID, "inflow is an x-low boundary");
    if (c->noise && c->nQ != 11) return fail(HD_ERR_INVALID, "noise needs nQ=11 (the stress fields)");
    if (!(c->lx > 0 && c->ly > 0 && c->lz > 0)) return fail(HD_ERR_INVALID, "domain lengths must be positive");
    if (!(c->gamma > 1.0)) return fail(HD_ERR_INVALID, "gamma must exceed 1");
    return HD_OK;
}

// kappa = 1/sqrt(sum_p (w_p/W)^2) of a product rule in 2 (face) or 3 (volume) dimensions
double rule_kappa(const Tables& t, int dims) {
    const int n = dims == 2 ? t.nface : t.npg;
    const double* w = dims == 2 ? t.w2 : t.w3;
    double W = 0., s2 = 0.;
    for (int p = 0; p < n; p++) W += w[p];
    for (int p = 0; p < n; p++) s2 += (w[p] / W) * (w[p] / W);
    return 1.0 / std::sqrt(s2);
}

}  // namespace hd

namespace {

using namespace hd;

bool pick_ops(const hd_ctx* c, Ops* o) { return pick_ops_for(c->nb, c->q, c->nQ, c->cfg.riemann, o); }

int alloc(hd_ctx* c, void** p, size_t bytes) {
    HIPCHK(hipMalloc(p, bytes));
    c->bytes += (int64_t)bytes;
    return HD_OK;
}

StageParams stage_params(const hd_ctx* c, const double* dt_dev, int combine, int mode) {
    StageParams sp;
    sp.dt = dt_dev;
    sp.coll = c->cfg.coll;
    sp.coll_q = c->cfg.coll_q;
    sp.ivis = c->cfg.ivis;
    sp.combine = combine;
    sp.mode = mode;
    for (int d = 0; d < 3; d++) sp.bf[d] = c->cfg.body_force[d];
    sp.err = c->err;
    sp.bad = c->bad;
    sp.limit_out = 0;
    sp.lp = c->lp;
    for (int d = 0; d < 3; d++) sp.noise[d] = c->noise_f[d];
    sp.noise[3] = c->noise_v;
    sp.nfb = c->nfb;
    sp.inflow = c->inflow;
    sp.rows_zfast = c->rows_zfast;
    return sp;
}

// Draws the stochastic stress of the current step for time step *dt_dev.
int draw_noise(hd_ctx* c, const double* dt_dev, hipStream_t s) {
    if (!c->cfg.noise) return HD_OK;
    noise::NoiseParams np;
    np.seed = c->cfg.seed;
    np.realization = c->cfg.realization;
    np.nface = c->nface;
    np.npg = c->nface * c->q;
    np.step = c->step_ctr;
    np.dt = dt_dev;
    np.eta_sd = c->cfg.eta_sd;
    np.zeta_sd = c->cfg.zeta_sd;
    np.dVi = c->grid.dxi * c->grid.dyi * c->grid.dzi;
    np.kap_f = c->kap_f;
    np.kap_v = c->kap_v;
    np.nfb = c->nfb;
    np.N = c->N;
    for (int d = 0; d < 3; d++) np.face[d] = c->noise_f[d];
    np.vol = c->noise_v;
    const long total = 3L * c->nface * c->nfb + (long)np.npg * c->N;
    hipLaunchKernelGGL(noise::k_noise, dim3(blocks_for(total, 256)), dim3(256), 0, s, np);
    HIPCHK(hipGetLastError());
    return HD_OK;
}

// immersed solid (apply_cyl_in_2d_pipe_boundaries, src/boundary.f90:686-700,
// with the mode loop over nbasis as intended): masked cells get density
// rho_mask with no slope and zero momentum in every mode
__global__ __launch_bounds__(256) void k_apply_mask(double* __restrict__ q, const unsigned char* __restrict__ mask,
                                                    long N, int nQ, int nb, double rho_mask) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N || !mask[c]) return;
    for (int r = 0; r < nb; r++) {
        q[((long)r * nQ + RH) * N + c] = r == 0 ? rho_mask : 0.0;
        q[((long)r * nQ + MX) * N + c] = 0.0;
        q[((long)r * nQ + MY) * N + c] = 0.0;
        q[((long)r * nQ + MZ) * N + c] = 0.0;
    }
}

int apply_mask(hd_ctx* c, double* q, hipStream_t s) {
    if (!c->mask_on) return HD_OK;
    hipLaunchKernelGGL(k_apply_mask, dim3(blocks_for(c->N, 256)), dim3(256), 0, s, q, c->mask, c->N, c->nQ, c->nb,
                       c->rho_mask);
    HIPCHK(hipGetLastError());
    return HD_OK;
}

CflEos cfl_eos(const hd_ctx* c) {
    CflEos e;
    e.gamma = c->cfg.gamma;
    e.ieos = c->cfg.ieos;
    e.P1 = c->cfg.eos_p1;
    e.T_floor = c->cfg.T_floor;
    return e;
}

int advance_step(hd_ctx* c, hipStream_t s) {
    if (!c->cfg.noise) return HD_OK;
    hipLaunchKernelGGL(noise::k_set_step, dim3(1), dim3(64), 0, s, c->step_ctr, 0ull, 1);
    HIPCHK(hipGetLastError());
    return HD_OK;
}

hipEvent_t next_event(hd_ctx* c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
    }
    return c->ev_pool[c->ev_used++];
}

// One Euler stage with optional fused SSP combination.  in_limited: the input
// was already limited by the previous stage's epilogue.  limit_out: the output
// is the next stage's input, so the fused kernel limits it before storing.
int stage(hd_ctx* c, double* qin, const double* qn, double* qout, const double* dt_dev, int combine, hipStream_t s,
          bool in_limited = false, bool limit_out = false) {
    Ops o;
    if (!pick_ops(c, &o)) return fail(HD_ERR_UNSUPPORTED, "configuration not built");
    if (c->cfg.bc[0] == HD_BC_INFLOW && !c->inflow_set)
        return fail(HD_ERR_INVALID, "inflow boundary without its state (hd_set_inflow)");
    const bool fuse_lim = c->cfg.limiter && !c->generic;
    if (c->cfg.limiter && !(in_limited && fuse_lim)) o.limiter(c, qin, s);
    StageParams sp = stage_params(c, dt_dev, combine, 0);
    sp.limit_out = (limit_out && fuse_lim) ? 1 : 0;
    // timing brackets the stage kernel(s) only, not the input limiter
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
        e0 = next_event(c); e1 = next_event(c);
        if (!e0 || !e1) return fail(HD_ERR_HIP, "hipEventCreate failed");
        HIPCHK(hipEventRecord(e0, s));
    }
    if (c->generic) {
        o.faces(c, qin, s);
        o.cell(c, qin, qn, qout, sp, s);
    } else {
        o.fused(c, qin, qn, qout, sp, s);
    }
    HIPCHK(hipGetLastError());
    if (c->timing) {
        HIPCHK(hipEventRecord(e1, s));
        c->ev_pending.emplace_back(e0, e1);
    }
    return HD_OK;
}

}  // namespace

extern "C" {

const char* hd_last_error(void) { return g_last_error.c_str(); }

int hd_create(const hd_config* cfg, hd_ctx** out) {
    if (!out) return fail(HD_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    hd_ctx* c = new (std::nothrow) hd_ctx();
    if (!c) return fail(HD_ERR_INVALID, "out of host memory");
    c->cfg = *cfg;
    c->q = cfg->iquad; c->nb = cfg->nbasis; c->nQ = cfg->nQ; c->nface = c->q * c->q;
    c->N = (long)cfg->nx * cfg->ny * cfg->nz;
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return fail(HD_ERR_HIP, "no HIP device"); }
    Grid& g = c->grid;
    g.nx = cfg->nx; g.ny = cfg->ny; g.nz = cfg->nz; g.nQ = cfg->nQ; g.N = c->N;
    // init_spatial_params (src/initialize.f90:298-327) with one rank
    g.dxi = (double)cfg->nx / (cfg->lx / 2.0 - (-(cfg->lx / 2.0)));
    g.dyi = (double)cfg->ny / (cfg->ly / 2.0 - (-(cfg->ly / 2.0)));
    g.dzi = (double)cfg->nz / (cfg->lz / 2.0 - (-(cfg->lz / 2.0)));
    for (int s = 0; s < 6; s++) g.bc[s] = cfg->bc[s];
    Phys& ph = c->ph;
    ph.gamma = cfg->gamma; ph.gm1 = cfg->gamma - 1.;
    ph.P_floor = cfg->T_floor * cfg->rh_floor; ph.rh_floor = cfg->rh_floor;
    ph.colvis = cfg->colvis; ph.c2d3cv = (2. / 3.) * cfg->colvis; ph.c4d3cv = (4. / 3.) * cfg->colvis;
    ph.ivis = cfg->ivis;
    ph.ieos = cfg->ieos; ph.P1 = cfg->eos_p1; ph.Pbase = cfg->eos_pbase;
    ph.colq = cfg->col_q;
    c->lp.rh_floor = cfg->rh_floor; c->lp.epsP = cfg->rh_floor * cfg->T_floor; c->lp.gm1 = ph.gm1;
    build_tables(c->nb, c->q, &c->htab);
    int r = HD_OK;
    const size_t nf = (size_t)c->nface * c->nQ;
    const size_t nfx = (size_t)(g.nx + 1) * g.ny * g.nz, nfy = (size_t)g.nx * (g.ny + 1) * g.nz,
                 nfz = (size_t)g.nx * g.ny * (g.nz + 1);
    c->npart = 1024;
    // the full 10-moment equations (ivis 2) and the water EOS (ieos 2) run on the
    // table-driven kernels only, and the 13-moment model on the same face/cell
    // split with its own kernels; HD_KERNELS=generic selects the table-driven
    // kernels for any configuration
    const char* kern = std::getenv("HD_KERNELS");
    c->generic = (kern && std::strcmp(kern, "generic") == 0) || cfg->ivis == 2 || cfg->ieos == 2 ||
                 cfg->nQ == 13;
    if (c->generic && cfg->noise) {
        hd_destroy(c);
        return fail(HD_ERR_UNSUPPORTED, "the table-driven kernels (HD_KERNELS=generic) carry no stochastic stress");
    }
    // face-flux arrays: only the face/cell split writes them (the fused stage
    // kernel keeps its face fluxes in registers and LDS)
    if ((r = alloc(c, (void**)&c->dtab, sizeof(Tables))) ||
        (c->generic && ((r = alloc(c, (void**)&c->fl[0], nf * nfx * sizeof(double))) ||
                        (r = alloc(c, (void**)&c->fl[1], nf * nfy * sizeof(double))) ||
                        (r = alloc(c, (void**)&c->fl[2], nf * nfz * sizeof(double))))) ||
        (r = alloc(c, (void**)&c->partial, c->npart * sizeof(double))) ||
        (r = alloc(c, (void**)&c->err, sizeof(int))) ||
        (r = alloc(c, (void**)&c->bad, sizeof(unsigned long long))) ||
        (r = alloc(c, (void**)&c->scratch_dt, sizeof(double))) ||
        (r = alloc(c, (void**)&c->step_ctr, sizeof(unsigned long long))) ||
        (r = alloc(c, (void**)&c->vjp_partial, blocks_for(c->N, 64) * sizeof(double))) ||
        (cfg->bc[0] == HD_BC_INFLOW &&
         (r = alloc(c, (void**)&c->inflow, (size_t)g.ny * g.nz * c->nface * c->nQ * sizeof(double)))) ||
        (cfg->integrator == 3 &&
         (r = alloc(c, (void**)&c->stage_buf, (size_t)c->N * c->nQ * c->nb * sizeof(double))))) {
        std::string msg = g_last_error;
        hd_destroy(c);
        return fail(r, msg);
    }
    if (cfg->noise) {
        // the step's stress field: 6 components at every point of the three
        // face lattices and of the volume rule
        c->nfb = (long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
        c->kap_f = rule_kappa(c->htab, 2);
        c->kap_v = rule_kappa(c->htab, 3);
        const size_t nfw = (size_t)6 * c->nface * c->nfb, nvw = (size_t)6 * c->htab.npg * c->N;
        for (int d = 0; d < 3 && r == HD_OK; d++) r = alloc(c, (void**)&c->noise_f[d], nfw * sizeof(float));
        if (r == HD_OK) r = alloc(c, (void**)&c->noise_v, nvw * sizeof(float));
        if (r) {
            std::string msg = g_last_error;
            hd_destroy(c);
            return fail(r, msg);
        }
    }
    const double one = 1.0;
    if (hipMemcpy(c->dtab, &c->htab, sizeof(Tables), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->scratch_dt, &one, sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->step_ctr, 0, sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(c->err, 0, sizeof(int)) != hipSuccess ||
        hipMemset(c->bad, 0xff, sizeof(unsigned long long)) != hipSuccess) {
        hd_destroy(c);
        return fail(HD_ERR_HIP, "initialising context buffers failed");
    }
    if (const char* ro = std::getenv("HD_STAGE_ROWS")) c->rows_zfast = std::strcmp(ro, "yz") != 0;
    if (const char* lm = std::getenv("HD_LIMITER")) c->table_limiter = std::strcmp(lm, "table") == 0;
    if (const char* fc = std::getenv("HD_FACES")) c->table_faces = std::strcmp(fc, "table") == 0;
    if (const char* vj = std::getenv("HD_VJP")) c->fused_vjp = std::strcmp(vj, "fused") == 0;
    // workgroup size of the fused stage kernel: 64, 128 or 256 threads
    if (const char* bs = std::getenv("HD_STAGE_BLOCK")) {
        const int v = std::atoi(bs);
        if (v == 64 || v == 128 || v == 256) c->stage_bs = v;
    }
    {
        const size_t bytes = (size_t)c->nface * c->nQ * cfg->nx * sizeof(double);
        const char* xs = std::getenv("HD_STAGE_XS");
        const bool off = xs && std::strcmp(xs, "0") == 0;
        if (!off && cfg->bc[0] == HD_BC_PERIODIC && cfg->nx % 64 == 0 && cfg->nx <= 256 && bytes <= 65536)
            c->xs_bytes = (unsigned)bytes;
    }
    *out = c;
    return HD_OK;
}

int hd_destroy(hd_ctx* c) {
    if (!c) return HD_OK;
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    void* ptrs[] = {c->dtab,     c->fl[0],    c->fl[1],     c->fl[2],      c->partial,    c->err,
                    c->bad,      c->scratch_dt, c->stage_buf, c->step_ctr, c->noise_f[0], c->noise_f[1],
                    c->noise_f[2], c->noise_v,  c->vjp_partial, c->inflow, c->mask,
                    c->bq2[0],     c->bq2[1],   c->bq2[2]};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete c;
    return HD_OK;
}

int64_t hd_device_bytes(const hd_ctx* c) { return c ? c->bytes : 0; }

int hd_timing_enable(hd_ctx* c, int enable) {
    if (!c) return fail(HD_ERR_INVALID, "NULL context");
    c->timing = enable != 0;
    return HD_OK;
}

int hd_timing_read(hd_ctx* c, double* total_ms, int64_t* stages) {
    if (!c || !total_ms || !stages) return fail(HD_ERR_INVALID, "NULL argument");
    double sum = 0.;
    for (auto& pr : c->ev_pending) {
        HIPCHK(hipEventSynchronize(pr.second));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
        sum += ms;
    }
    *total_ms = sum;
    *stages = (int64_t)c->ev_pending.size();
    c->ev_pending.clear();
    c->ev_used = 0;
    return HD_OK;
}

int hd_limit(hd_ctx* c, double* q, void* stream) {
    if (!c || !q) return fail(HD_ERR_INVALID, "NULL argument");
    Ops o;
    if (!pick_ops(c, &o)) return fail(HD_ERR_UNSUPPORTED, "configuration not built");
    o.limiter(c, q, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return HD_OK;
}

int hd_rhs(hd_ctx* c, const double* q, double* rhs, void* stream) {
    if (!c || !q || !rhs) return fail(HD_ERR_INVALID, "NULL argument");
    Ops o;
    if (!pick_ops(c, &o)) return fail(HD_ERR_UNSUPPORTED, "configuration not built");
    hipStream_t s = (hipStream_t)stream;
    if (q == rhs) return fail(HD_ERR_INVALID, "q and rhs must differ");
    if (c->cfg.bc[0] == HD_BC_INFLOW && !c->inflow_set)
        return fail(HD_ERR_INVALID, "inflow boundary without its state (hd_set_inflow)");
    // rhs mode does not advance; the scratch scalar (1.0) is the time step of
    // the noise amplitude
    int rc = draw_noise(c, c->scratch_dt, s);
    if (rc) return rc;
    StageParams sp = stage_params(c, c->scratch_dt, 0, 1);
    if (c->generic) {
        o.faces(c, q, s);
        o.cell(c, q, q, rhs, sp, s);
    } else {
        o.fused(c, q, q, rhs, sp, s);
    }
    HIPCHK(hipGetLastError());
    return HD_OK;
}

int hd_euler_step(hd_ctx* c, double* q_in, double* q_out, const double* dt_dev, void* stream) {
    if (!c || !q_in || !q_out || !dt_dev) return fail(HD_ERR_INVALID, "NULL argument");
    if (q_in == q_out) return fail(HD_ERR_INVALID, "q_in and q_out must differ");
    int rc = draw_noise(c, dt_dev, (hipStream_t)stream);
    if (rc) return rc;
    return stage(c, q_in, q_in, q_out, dt_dev, 0, (hipStream_t)stream);
}

int hd_min_dt(hd_ctx* c, const double* q, double* dt_dev, void* stream) {
    if (!c || !q || !dt_dev) return fail(HD_ERR_INVALID, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const int nblk = (int)std::min<long>(c->npart, (c->N + 255) / 256);
    hipLaunchKernelGGL(k_wavespeed, dim3(nblk), dim3(256), 0, s, q, c->grid, cfl_eos(c), 1.01 * c->cfg.rh_floor,
                       c->partial);
    hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(256), 0, s, q, c->grid, cfl_eos(c), c->partial, nblk, c->cfg.cflm,
                       dt_dev, (double*)nullptr);
    HIPCHK(hipGetLastError());
    return HD_OK;
}

int hd_update(hd_ctx* c, double* q_io, double* q1, double* q2, const double* dt_dev, void* stream) {
    if (!c || !q_io || !q1 || !dt_dev) return fail(HD_ERR_INVALID, "NULL argument");
    if (q_io == q1 || (q2 && (q2 == q_io || q2 == q1))) return fail(HD_ERR_INVALID, "q_io, q1 and q2 must differ");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    // one stress field per step, shared by its stages
    if ((rc = draw_noise(c, dt_dev, s))) return rc;
    if (c->cfg.integrator == 3) {
        // RK3 (src/integrator.f90:74-86): Q1 = E(Q0); Q2 = 3/4 Q0 + 1/4 E(Q1); Q0 = 1/3 Q0 + 2/3 E(Q2).
        // A stage never writes the array it reads neighbours from.
        double* q2b = q2 ? q2 : c->stage_buf;
        if ((rc = stage(c, q_io, q_io, q1, dt_dev, 0, s, false, true))) return rc;
        if ((rc = stage(c, q1, q_io, q2b, dt_dev, 2, s, true, true))) return rc;
        if ((rc = stage(c, q2b, q_io, q_io, dt_dev, 3, s, true, false))) return rc;
    } else {
        // RK2 (src/integrator.f90:62-71): Q1 = E(Q0); Q0 = 1/2 (Q0 + E(Q1))
        if ((rc = stage(c, q_io, q_io, q1, dt_dev, 0, s, false, true))) return rc;
        if ((rc = stage(c, q1, q_io, q_io, dt_dev, 1, s, true, false))) return rc;
    }
    if ((rc = apply_mask(c, q_io, s))) return rc;
    return advance_step(c, s);
}

int hd_stage(hd_ctx* c, int32_t rk_order, int32_t stage_index, double* q_in, const double* q_n, double* q_out,
             const double* dt_dev, void* stream) {
    if (!c || !q_in || !q_out || !dt_dev) return fail(HD_ERR_INVALID, "NULL argument");
    if (q_in == q_out) return fail(HD_ERR_INVALID, "q_in and q_out must differ");
    int combine;
    if (rk_order == 2 && (stage_index == 1 || stage_index == 2)) combine = stage_index == 1 ? 0 : 1;
    else if (rk_order == 3 && stage_index >= 1 && stage_index <= 3) combine = stage_index == 1 ? 0 : stage_index;
    else return fail(HD_ERR_INVALID, "rk_order must be 2 (stages 1-2) or 3 (stages 1-3)");
    if (combine != 0 && !q_n) return fail(HD_ERR_INVALID, "q_n is required on combination stages");
    if (q_n == q_out && combine != 0) return fail(HD_ERR_INVALID, "q_n and q_out must differ");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if (stage_index == 1 && (rc = draw_noise(c, dt_dev, s))) return rc;
    if ((rc = stage(c, q_in, combine ? q_n : q_in, q_out, dt_dev, combine, s, false, false))) return rc;
    if (stage_index == rk_order) {
        if ((rc = apply_mask(c, q_out, s))) return rc;
        return advance_step(c, s);
    }
    return HD_OK;
}

int hd_set_inflow(hd_ctx* c, const double* def_dev, void* stream) {
    if (!c || !def_dev) return fail(HD_ERR_INVALID, "NULL argument");
    if (c->cfg.bc[0] != HD_BC_INFLOW) return fail(HD_ERR_INVALID, "the x-low boundary is not an inflow");
    const size_t n = (size_t)c->grid.ny * c->grid.nz * c->nface * c->nQ;
    HIPCHK(hipMemcpyAsync(c->inflow, def_dev, n * sizeof(double), hipMemcpyDeviceToDevice, (hipStream_t)stream));
    c->inflow_set = true;
    return HD_OK;
}

int hd_set_mask(hd_ctx* c, const uint8_t* mask_dev, double rho_mask, void* stream) {
    if (!c) return fail(HD_ERR_INVALID, "NULL context");
    if (!mask_dev) {
        c->mask_on = false;
        return HD_OK;
    }
    if (!c->mask) {
        int r = alloc(c, (void**)&c->mask, (size_t)c->N);
        if (r) return r;
    }
    HIPCHK(hipMemcpyAsync(c->mask, mask_dev, (size_t)c->N, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    c->rho_mask = rho_mask;
    c->mask_on = true;
    return HD_OK;
}

int hd_apply_mask(hd_ctx* c, double* q, void* stream) {
    if (!c || !q) return fail(HD_ERR_INVALID, "NULL argument");
    return apply_mask(c, q, (hipStream_t)stream);
}

// The split 13-moment transpose (periodic grids) writes the left cells' face
// cotangents into fl and the right cells' into bq2, allocated here once.
static int ensure_vjp13_buffers(hd_ctx* c) {
    if (c->cfg.nQ != 13 || c->fused_vjp || c->bq2[0]) return HD_OK;
    for (int b = 0; b < 6; b++)
        if (c->cfg.bc[b] != HD_BC_PERIODIC) return HD_OK;
    const Grid& g = c->grid;
    const size_t nf = (size_t)c->nface * c->nQ;
    const size_t nfd[3] = {(size_t)(g.nx + 1) * g.ny * g.nz, (size_t)g.nx * (g.ny + 1) * g.nz,
                           (size_t)g.nx * g.ny * (g.nz + 1)};
    for (int d = 0; d < 3; d++) {
        const int r = alloc(c, (void**)&c->bq2[d], nf * nfd[d] * sizeof(double));
        if (r) return r;
    }
    return HD_OK;
}

int hd_stage_vjp(hd_ctx* c, const double* q_in, const double* lam_out, double* lam_in, const double* dt_dev,
                 const hd_stage_vjp_args* args, void* stream) {
    if (!c || !q_in || !lam_out || !lam_in || !dt_dev || !args) return fail(HD_ERR_INVALID, "NULL argument");
    if (lam_in == lam_out || lam_in == q_in) return fail(HD_ERR_INVALID, "lam_in must differ from lam_out and q_in");
    if (c->cfg.noise) return fail(HD_ERR_UNSUPPORTED, "the adjoint does not differentiate the stochastic stress");
    Ops o;
    if (!pick_ops(c, &o)) return fail(HD_ERR_UNSUPPORTED, "configuration not built");
    if (!o.stage_vjp) return fail(HD_ERR_UNSUPPORTED, "configuration has no adjoint");
    if (c->cfg.ivis == 2 || c->cfg.ieos == 2)
        return fail(HD_ERR_UNSUPPORTED, "no adjoint for the full 10-moment equations or the water EOS");
    if (c->mask_on) return fail(HD_ERR_UNSUPPORTED, "the adjoint does not differentiate the immersed-solid mask");
    if (c->cfg.bc[0] == HD_BC_INFLOW && !c->inflow_set)
        return fail(HD_ERR_INVALID, "inflow boundary without its state (hd_set_inflow)");
    adj::AdjParams ap;
    ap.dt = dt_dev;
    ap.coll = c->cfg.coll;
    ap.coll_q = c->cfg.coll_q;
    ap.ivis = c->cfg.ivis;
    ap.a = args->a;
    ap.b = args->b;
    for (int d = 0; d < 3; d++) ap.bf[d] = c->cfg.body_force[d];
    ap.acc[0] = args->acc0; ap.acc[1] = args->acc1;
    ap.w[0] = args->w0; ap.w[1] = args->w1;
    ap.q_raw = c->cfg.limiter ? args->q_raw : nullptr;
    ap.lp = c->lp;
    ap.out = args->out;
    ap.qn = args->qn;
    if (ap.out && ap.a != 0. && !ap.qn) return fail(HD_ERR_INVALID, "qn is required with out when a != 0");
    ap.partial = args->dcoll_dev ? c->vjp_partial : nullptr;
    ap.rhs_only = 0;
    ap.inflow = c->inflow;
    hipStream_t s = (hipStream_t)stream;
    if (int r = ensure_vjp13_buffers(c)) return r;
    o.stage_vjp(c, q_in, lam_out, lam_in, ap, s);
    HIPCHK(hipGetLastError());
    if (args->dcoll_dev) {
        hipLaunchKernelGGL(adj::k_sum_partials, dim3(1), dim3(256), 0, s, c->vjp_partial, c->vjp_blocks,
                           args->dcoll_dev);
        HIPCHK(hipGetLastError());
    }
    return HD_OK;
}

int hd_rhs_vjp(hd_ctx* c, const double* q, const double* lam, double* out, void* stream) {
    if (!c || !q || !lam || !out) return fail(HD_ERR_INVALID, "NULL argument");
    if (out == lam || out == q) return fail(HD_ERR_INVALID, "out must differ from q and lam");
    if (c->cfg.noise) return fail(HD_ERR_UNSUPPORTED, "the adjoint does not differentiate the stochastic stress");
    Ops o;
    if (!pick_ops(c, &o)) return fail(HD_ERR_UNSUPPORTED, "configuration not built");
    if (!o.stage_vjp) return fail(HD_ERR_UNSUPPORTED, "configuration has no adjoint");
    if (c->cfg.ivis == 2 || c->cfg.ieos == 2)
        return fail(HD_ERR_UNSUPPORTED, "no adjoint for the full 10-moment equations or the water EOS");
    if (c->cfg.bc[0] == HD_BC_INFLOW && !c->inflow_set)
        return fail(HD_ERR_INVALID, "inflow boundary without its state (hd_set_inflow)");
    adj::AdjParams ap{};
    ap.dt = c->scratch_dt;
    ap.coll = c->cfg.coll;
    ap.coll_q = c->cfg.coll_q;
    ap.ivis = c->cfg.ivis;
    ap.a = 0.;
    ap.b = 1.;
    for (int d = 0; d < 3; d++) ap.bf[d] = c->cfg.body_force[d];
    ap.lp = c->lp;
    ap.rhs_only = 1;
    ap.inflow = c->inflow;
    if (int r = ensure_vjp13_buffers(c)) return r;
    o.stage_vjp(c, q, lam, out, ap, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return HD_OK;
}

int hd_set_step(hd_ctx* c, uint64_t step, void* stream) {
    if (!c) return fail(HD_ERR_INVALID, "NULL context");
    hipLaunchKernelGGL(noise::k_set_step, dim3(1), dim3(64), 0, (hipStream_t)stream, c->step_ctr,
                       (unsigned long long)step, 0);
    HIPCHK(hipGetLastError());
    return HD_OK;
}

int hd_step(hd_ctx* c, double* q_io, double* q1, double* q2, double* t_dev, double* dt_dev, void* stream) {
    if (!c || !q_io || !q1 || !dt_dev) return fail(HD_ERR_INVALID, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const int nblk = (int)std::min<long>(c->npart, (c->N + 255) / 256);
    hipLaunchKernelGGL(k_wavespeed, dim3(nblk), dim3(256), 0, s, q_io, c->grid, cfl_eos(c), 1.01 * c->cfg.rh_floor,
                       c->partial);
    hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(256), 0, s, q_io, c->grid, cfl_eos(c), c->partial, nblk,
                       c->cfg.cflm, dt_dev, t_dev);
    HIPCHK(hipGetLastError());
    return hd_update(c, q_io, q1, q2, dt_dev, stream);
}

int hd_check(hd_ctx* c, void* stream, int64_t* bad_cell) {
    if (!c) return fail(HD_ERR_INVALID, "NULL context");
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    int flag = 0;
    unsigned long long bad = ~0ull;
    HIPCHK(hipMemcpy(&flag, c->err, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&bad, c->bad, sizeof(bad), hipMemcpyDeviceToHost));
    if (bad_cell) *bad_cell = flag ? (int64_t)bad : -1;
    if (flag) {
        HIPCHK(hipMemset(c->err, 0, sizeof(int)));
        HIPCHK(hipMemset(c->bad, 0xff, sizeof(unsigned long long)));
        return fail(HD_ERR_NONFINITE, "NaN in mode 1 of a stage output at cell " + std::to_string(bad));
    }
    return HD_OK;
}

}  // extern "C"
