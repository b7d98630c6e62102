/*
 * hermeshd_amd.h -- C ABI of the MI355X (gfx950) HERMESHD hot path: the
 * discontinuous-Galerkin right-hand side and the SSP-RK stage update.
 *
 * All state pointers are device pointers to FP64 arrays laid out as the
 * reference's Fortran Q(nx,ny,nz,nQ,nbasis) (cell index i fastest; a C array
 * q[nbasis][nQ][nz][ny][nx]).  Streams are hipStream_t passed as void*
 * (NULL = the legacy default stream).  Every call is stream-ordered and
 * asynchronous unless stated; every call returns HD_OK or a negative code and
 * never aborts the process.  The text of the last error of the calling thread
 * is returned by hd_last_error().  One context serves one configuration; many
 * contexts may coexist.  Calls on one context must not overlap (they share
 * the context's scratch buffers).
 *
 * The reference exposes these operations through the f90wrap module
 * `hermeshd` (src/hermeshd.py:5-120) over module procedures of
 * src/hermeshd.f90, src/integrator.f90 and src/flux.f90; each entry point
 * below names the routine it replaces.
 */
#ifndef HERMESHD_AMD_H
#define HERMESHD_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define HD_OK 0
#define HD_ERR_INVALID (-1)     /* bad argument or configuration */
#define HD_ERR_HIP (-2)         /* HIP runtime failure */
#define HD_ERR_NONFINITE (-3)   /* NaN in mode 1 of a stage output (src/integrator.f90:234-250) */
#define HD_ERR_UNSUPPORTED (-4) /* valid but not built */

/* boundary kinds (src/boundary.f90:789-880) */
#define HD_BC_PERIODIC 0
#define HD_BC_OUTFLOW 1
#define HD_BC_WALL 2
#define HD_BC_NOSLIP 3
#define HD_BC_INFLOW 4   /* x-low only: prescribed rh, mx, en (x_bc_inflow, src/boundary.f90:571-592) */
/* face solvers for rh..en; other fields always use the central flux */
#define HD_RIEMANN_CENTRAL 0
#define HD_RIEMANN_HLLC 1
#define HD_RIEMANN_ROE 2

/* The compile-time switches of src/input.f90 and src/params.f90, at run time. */
typedef struct hd_config {
    int32_t nx, ny, nz;          /* cells (one device holds the whole grid) */
    int32_t nbasis, iquad, nQ;   /* (4,2), (8,2) or (10,3); nQ 5, 11 or 13.  nQ 13 is the
                                    13-moment model (no counterpart in the reference): fields
                                    rh,mx,my,mz,en,pxx,pyy,pxy,pxz,pyz,qx,qy,qz, traceless stress
                                    (pzz = -pxx-pyy) and Maxwell-Cattaneo heat flux; ivis 0 or 1,
                                    ideal gas, no noise */
    int32_t ivis, ieos;          /* ivis 0 explicit, 1 semi-implicit stress relaxation, 2 full
                                    10-moment equations (nQ 11, no noise, no adjoint or ensemble:
                                    table-driven kernels); ieos 1 ideal gas, 2 Murnaghan-Tait
                                    water (table-driven kernels, no Roe) */
    int32_t riemann;             /* HD_RIEMANN_* (input.f90 ihllc; Roe from the legacy solver) */
    int32_t limiter;             /* 1 = positivity limiter on every stage input */
    int32_t integrator;          /* 2 = heun (SSP-RK2), 3 = shu-osher (SSP-RK3) */
    int32_t bc[6];               /* xlo,xhi,ylo,yhi,zlo,zhi: HD_BC_* */
    int32_t noise;               /* 1 = stochastic stress in the fluxes (needs nQ = 11) */
    int32_t reserved0[2];        /* zero; keeps the doubles 8-byte aligned without padding */
    double lx, ly, lz;           /* domain lengths, grid centred on the origin */
    double gamma;                /* aindex */
    double rh_floor, T_floor;    /* floors; P_floor = T_floor*rh_floor */
    double coll, colvis;         /* relaxation rate, coll*viscosity */
    double cflm;                 /* CFL number of get_min_dt */
    double body_force[3];        /* constant acceleration (no counterpart in the reference) */
    double eta_sd, zeta_sd;      /* noise amplitudes sqrt(2 eta T), sqrt(zeta T / 3) */
    uint64_t seed;               /* Philox key */
    uint32_t realization, reserved1;
    double eos_p1, eos_pbase;    /* ieos 2: P = eos_p1*(rho^7.2 - 1) + eos_pbase + P_ideal
                                    (src/params.f90:97-105: P_1, P_base) */
    double coll_q, col_q;        /* nQ 13: heat-flux relaxation rate and coll_q*conductivity
                                    (no counterpart in the reference) */
} hd_config;

typedef struct hd_ctx hd_ctx;

/* Builds basis/quadrature tables, geometry and scratch on the current device.
   Replaces initializer (src/initialize.f90:24-61) minus MPI and output.
   HD_ERR_INVALID for an invalid config, checked before any device call; for
   nQ 13 that includes (nx+1)(ny+1)(nz+1) > 2^31 - 1 (32-bit face indices). */
int hd_create(const hd_config *cfg, hd_ctx **out);
int hd_destroy(hd_ctx *ctx);

/* Positivity limiter, in place.  Replaces limiter (src/flux.f90:1077-1164). */
int hd_limit(hd_ctx *ctx, double *q, void *stream);

/* rhs = -glflux + source for an (already limited) state.  Replaces calc_rhs
   (src/integrator.f90:116-122): glflux (src/flux.f90:932-1022) with
   exchange_flux/apply_boundaries, and source_calc (src/sources.f90:12-84).
   With noise on, the stochastic stress of the current step index enters the
   fluxes with the amplitude of a time step of 1. */
int hd_rhs(hd_ctx *ctx, const double *q, double *rhs, void *stream);

/* One forward-Euler stage: limiter on q_in (in place), then
   q_out = q_in - dt*(glflux - source) (+ implicit stress relaxation when ivis = 1, 2).
   dt_dev points to one double in device memory.  Replaces euler_step
   (src/integrator.f90:93-102) and advance_time_level (125-252). */
int hd_euler_step(hd_ctx *ctx, double *q_in, double *q_out, const double *dt_dev, void *stream);

/* dt from the CFL condition, written to dt_dev.  Replaces get_min_dt
   (src/hermeshd.f90:145-198), including its sound-speed formula. */
int hd_min_dt(hd_ctx *ctx, const double *q, double *dt_dev, void *stream);

/* hd_min_dt without a context and without allocating: the time step of the
   configuration cfg for state q, written to dt_dev.  No counterpart in the
   reference (get_min_dt, src/hermeshd.f90:145-198, is what it computes); it is
   for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops).  The workgroup maxima of the reduction live
   in a static device array of 65536 doubles (512 KiB, held by each GPU the
   library runs on until the process ends), so grids of more than 16,777,216
   cells are refused (HD_ERR_INVALID).  Calls that overlap on one GPU share
   that array and race; calls on one stream do not. */
int hd_min_dt_static(const hd_config *cfg, const double *q, double *dt_dev, void *stream);

/* One SSP-RK2 or SSP-RK3 step with a given dt.  Replaces RK2/RK3
   (src/integrator.f90:62-86).  q1 and q2 are stage storage as in the
   reference; q2 may be NULL (SSP-RK2 does not need it, and for SSP-RK3 the
   context's own stage array is used).  The SSP combinations are fused into
   the stage kernels.  On return q_io holds the new state; q1 and q2 hold
   unspecified values.  With noise on, one stochastic stress field is drawn
   for the step (key seed, realization, step index; amplitude from *dt_dev;
   Gaussian draws by an FP32 Box-Muller transform, the reference's precision
   for them, combined into the stress in FP64 and stored in FP32), used by
   every stage, and the step index then advances by one. */
int hd_update(hd_ctx *ctx, double *q_io, double *q1, double *q2, const double *dt_dev, void *stream);

/* dt = min_dt(q_io); update; t += dt.  t_dev and dt_dev are device doubles.
   Replaces step (src/hermeshd.f90:53-62). */
int hd_step(hd_ctx *ctx, double *q_io, double *q1, double *q2, double *t_dev, double *dt_dev,
            void *stream);

/* One stage of an SSP-RK step: limiter on q_in (in place), then
   q_out = a*q_n + b*E(q_in) with (a,b) of stage `stage_index` (1-based) of
   SSP-RK2 (rk_order 2: (0,1), (1/2,1/2)) or SSP-RK3 (rk_order 3: (0,1),
   (3/4,1/4), (1/3,2/3)).  q_out is not limited.  Replaces euler_step plus
   the combination lines of RK2/RK3 (src/integrator.f90:62-122).  With noise
   on, stage 1 draws the step's stress field and the last stage advances the
   step index. */
int hd_stage(hd_ctx *ctx, int32_t rk_order, int32_t stage_index, double *q_in, const double *q_n,
             double *q_out, const double *dt_dev, void *stream);

/* ---- adjoint (reverse mode) ------------------------------------------------
   No counterpart in the reference, which has no adjoint: these give the
   gradient of a run with respect to its initial state and the relaxation
   rate coll (used for the Poiseuille-channel sensitivity).  Central, HLLC and
   Roe face solvers; no noise.  On a 13-moment context periodic in every
   direction, the first hd_rhs_vjp or hd_stage_vjp allocates the face
   cotangents of the split transpose (3 arrays of nQ*iquad^2 doubles per face,
   counted by hd_device_bytes) unless HD_VJP=fused was set at hd_create. */

/* out = (dR/dq)^T lam, R the right-hand side of hd_rhs at state q. */
int hd_rhs_vjp(hd_ctx *ctx, const double *q, const double *lam, double *out, void *stream);

typedef struct hd_stage_vjp_args {
    double a, b;               /* the stage's weights: out = a*q_n + b*E(q_in) */
    const double *acc0, *acc1; /* NULL or arrays added to lam_in with weights w0, w1 */
    double w0, w1;
    const double *q_raw;       /* stage input before the limiter (used when the limiter is on) */
    const double *out, *qn;    /* forward stage output and q_n, for d/dcoll (out NULL: skip;
                                  qn is read only when a != 0, and then required) */
    double *dcoll_dev;         /* device double: += d(lam_out . out)/dcoll (NULL: skip) */
} hd_stage_vjp_args;

/* Transpose of one stage (hd_stage) at its (limited) input q_in, applied to
   lam_out, the cotangent of its output:
   lam_in = w0*acc0 + w1*acc1 + L'^T (B^T b lam_out + dt (dR/dq)^T B^T b lam_out),
   B the implicit stress scaling (ivis 1), L' the limiter with the factors of
   the forward pass at q_raw (active set frozen).  The cotangent a*lam_out of
   q_n is left to the caller (pass it through acc0/acc1 of the stage that
   produced q_n). */
int hd_stage_vjp(hd_ctx *ctx, const double *q_in, const double *lam_out, double *lam_in,
                 const double *dt_dev, const hd_stage_vjp_args *args, void *stream);

/* Sets the step index that keys the stochastic stress (0 after hd_create),
   so that a restarted run draws the same noise.  Replaces the stream set-up
   of random_init (src/random.f90:36-39). */
int hd_set_step(hd_ctx *ctx, uint64_t step, void *stream);

/* Exterior state of an HD_BC_INFLOW x-low boundary: def_dev points to the
   reference's Qxlo_ext_def(ny,nz,nface,nQ) (Fortran order, device memory),
   copied into the context on `stream`; rh, mx and en are used, the transverse
   momenta are 0 and the stress fields come from the interior.  Required
   before the first stage of an inflow configuration.  Replaces the array that
   set_xlobc_inflow fills (src/boundary.f90:595-622) for x_bc_inflow
   (571-592). */
int hd_set_inflow(hd_ctx *ctx, const double *def_dev, void *stream);

/* Immersed-solid mask: mask_dev is (nx,ny,nz) uint8 in Fortran order
   (1 = solid), copied into the context; NULL switches the mask off.  With a
   mask, hd_update, hd_step and the last stage of hd_stage reset the solid
   cells after the step: density rho_mask with no slope, zero momentum in
   every mode (apply_cyl_in_2d_pipe_boundaries, src/boundary.f90:686-700,
   and the mask of cyl_in_2d_pipe_mask, 625-648).  The adjoint entry points
   refuse a masked context. */
int hd_set_mask(hd_ctx *ctx, const uint8_t *mask_dev, double rho_mask, void *stream);

/* Applies the mask to q now (no-op without a mask). */
int hd_apply_mask(hd_ctx *ctx, double *q, void *stream);

/* Derived fields of output_vtk (src/output.f90:14-318) sampled at nvtk^3
   interior points of every cell (set_vtk_vals_3D, src/basis_funcs.f90:90-200;
   nvtk 1..4).  `fields` selects bits of HD_OUT_*; out (device, float32)
   receives one (nz*nvtk, ny*nvtk, nx*nvtk) plane per selected field, x fastest,
   in bit order.  scales (host, HD_OUT_COUNT entries, or NULL for 1) are the
   unit factors the reference multiplies by (n0, v0, te0/eV_per_K, 1/t0, ...);
   the log-density factor is applied inside the logarithm, as there.  The
   internal energy density is written as computed (the reference computes it
   and then writes its previous field's buffer). */
#define HD_OUT_DENSITY 0
#define HD_OUT_LOG_DENSITY 1
#define HD_OUT_VX 2
#define HD_OUT_VY 3
#define HD_OUT_VZ 4
#define HD_OUT_TEMPERATURE 5
#define HD_OUT_INTERNAL_ENERGY 6
#define HD_OUT_TOTAL_ENERGY 7
#define HD_OUT_ENTROPY 8
#define HD_OUT_PRESSURE 9
#define HD_OUT_PXX 10
#define HD_OUT_PYY 11
#define HD_OUT_PZZ 12
#define HD_OUT_VORTICITY 13
#define HD_OUT_COUNT 14
int hd_sample(hd_ctx *ctx, const double *q, int32_t nvtk, uint32_t fields, const double *scales, float *out,
              void *stream);

/* Waits for the context's work on `stream` and reports the NaN flag raised by
   any stage since the last check (HD_ERR_NONFINITE, with the first bad cell
   index in *bad_cell), then clears it.  The reference exits the process at
   this point (src/integrator.f90:234-250); here the caller decides. */
int hd_check(hd_ctx *ctx, void *stream, int64_t *bad_cell);

/* ------------------------------------------------------------------------
 * Ensembles: many independent realizations of one configuration on boxes of
 * different sizes, advanced together.  The reference runs one realization
 * per process (the time loop of src/main.f90:22-49 around step,
 * src/hermeshd.f90:53-62); each realization here computes exactly what one
 * context with the same configuration and realization id computes.
 * The states of all realizations are packed back to back in one device
 * array, each in the layout above; hd_ens_offsets gives where each starts.
 * `base` gives the physics, basis, boundaries and noise; its nx,ny,nz and
 * lx,ly,lz only fix the cell size (lx/nx, ...), shared by every realization.
 * dims holds nreal (nx,ny,nz) triples.  realization_ids (NULL: base
 * realization + index) key each realization's noise.
 * ------------------------------------------------------------------------ */
typedef struct hd_ens hd_ens;

int hd_ens_create(const hd_config *base, int32_t nreal, const int32_t *dims,
                  const uint32_t *realization_ids, hd_ens **out);
int hd_ens_destroy(hd_ens *ens);
/* doubles in one packed state array */
int64_t hd_ens_state_size(const hd_ens *ens);
/* nreal start offsets (in doubles) of the realizations in a packed array */
int hd_ens_offsets(const hd_ens *ens, int64_t *offsets);
int64_t hd_ens_device_bytes(const hd_ens *ens);

/* hd_limit, hd_rhs and hd_update over every realization; one launch per
   stage for all of them, one noise launch per step. */
int hd_ens_limit(hd_ens *ens, double *q, void *stream);
int hd_ens_rhs(hd_ens *ens, const double *q, double *rhs, void *stream);
int hd_ens_update(hd_ens *ens, double *q_io, double *q1, double *q2, const double *dt_dev,
                  void *stream);

/* nsteps steps of hd_ens_update, replayed from a HIP graph of one step that
   is captured on the first call (and again when an argument or the launch
   shape changes).  *dt_dev may change between calls.  With noise on, the call
   first draws the current step's field; each replayed step then draws the
   next step's field on the ensemble's own side stream while its stages run
   (the field is double-buffered), so results equal hd_ens_update's. */
int hd_ens_run(hd_ens *ens, double *q_io, double *q1, double *q2, const double *dt_dev,
               int64_t nsteps, void *stream);

int hd_ens_set_step(hd_ens *ens, uint64_t step, void *stream);
/* like hd_check; reports the realization and its cell */
int hd_ens_check(hd_ens *ens, void *stream, int64_t *bad_realization, int64_t *bad_cell);

/* Launch shapes of the stage kernel: tiles of consecutive cells or bricks,
   one workgroup per tile or a persistent grid walking the (realization,
   tile) list.  info: block size, brick bx,by,bz, linear flag, and
   2*(workgroups per CU, 0 = one per tile) + (XCD-contiguous chunks).  */
int32_t hd_ens_launch_count(void);
int hd_ens_launch_info(int32_t choice, int32_t info[6]);
int hd_ens_set_launch(hd_ens *ens, int32_t choice);
int hd_ens_get_launch(const hd_ens *ens);
/* Autotuner: times reps stages of every launch shape on q (read only; the
   output goes to scratch), keeps the fastest and returns it in *choice;
   us_per_stage (NULL or hd_ens_launch_count() doubles) receives the times.
   Synchronises the device. */
int hd_ens_tune(hd_ens *ens, double *q, double *scratch, const double *dt_dev, int32_t reps,
                void *stream, int32_t *choice, double *us_per_stage);

/* Stage timing for measurement: when enabled, every stage launched by
   hd_euler_step/hd_update/hd_step is bracketed by HIP events recorded on the
   caller's stream around the stage kernel (the input limiter of a step's
   first stage is outside the bracket).  hd_timing_read synchronises, returns the summed stage time
   in milliseconds and the number of stages since the last read, and resets. */
int hd_timing_enable(hd_ctx *ctx, int enable);
int hd_timing_read(hd_ctx *ctx, double *total_ms, int64_t *stages);

/* Number of bytes of device memory the context holds. */
int64_t hd_device_bytes(const hd_ctx *ctx);

const char *hd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
