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
/* face solvers for rh..en; other fields always use the central flux */
#define HD_RIEMANN_CENTRAL 0
#define HD_RIEMANN_HLLC 1
#define HD_RIEMANN_ROE 2

/* The compile-time switches of src/input.f90 and src/params.f90, at run time. */
typedef struct hd_config {
    int32_t nx, ny, nz;          /* cells (one device holds the whole grid) */
    int32_t nbasis, iquad, nQ;   /* (4,2), (8,2) or (10,3); nQ 5 or 11 */
    int32_t ivis, ieos;          /* ivis 0 explicit, 1 semi-implicit stress relaxation; ieos 1 */
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
} hd_config;

typedef struct hd_ctx hd_ctx;

/* Builds basis/quadrature tables, geometry and scratch on the current device.
   Replaces initializer (src/initialize.f90:24-61) minus MPI and output. */
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
   q_out = q_in - dt*(glflux - source) (+ stress relaxation when ivis = 1).
   dt_dev points to one double in device memory.  Replaces euler_step
   (src/integrator.f90:93-102) and advance_time_level (125-252). */
int hd_euler_step(hd_ctx *ctx, double *q_in, double *q_out, const double *dt_dev, void *stream);

/* dt from the CFL condition, written to dt_dev.  Replaces get_min_dt
   (src/hermeshd.f90:145-198), including its sound-speed formula. */
int hd_min_dt(hd_ctx *ctx, const double *q, double *dt_dev, void *stream);

/* One SSP-RK2 or SSP-RK3 step with a given dt.  Replaces RK2/RK3
   (src/integrator.f90:62-86).  q1 and q2 are stage storage as in the
   reference; q2 may be NULL (SSP-RK2 does not need it, and for SSP-RK3 the
   context's own stage array is used).  The SSP combinations are fused into
   the stage kernels.  On return q_io holds the new state; q1 and q2 hold
   unspecified values.  With noise on, one stochastic stress field is drawn
   for the step (key seed, realization, step index; amplitude from *dt_dev),
   used by every stage, and the step index then advances by one. */
int hd_update(hd_ctx *ctx, double *q_io, double *q1, double *q2, const double *dt_dev, void *stream);

/* dt = min_dt(q_io); update; t += dt.  t_dev and dt_dev are device doubles.
   Replaces step (src/hermeshd.f90:53-62). */
int hd_step(hd_ctx *ctx, double *q_io, double *q1, double *q2, double *t_dev, double *dt_dev,
            void *stream);

/* Sets the step index that keys the stochastic stress (0 after hd_create),
   so that a restarted run draws the same noise.  Replaces the stream set-up
   of random_init (src/random.f90:36-39). */
int hd_set_step(hd_ctx *ctx, uint64_t step, void *stream);

/* Waits for the context's work on `stream` and reports the NaN flag raised by
   any stage since the last check (HD_ERR_NONFINITE, with the first bad cell
   index in *bad_cell), then clears it.  The reference exits the process at
   this point (src/integrator.f90:234-250); here the caller decides. */
int hd_check(hd_ctx *ctx, void *stream, int64_t *bad_cell);

/* Stage timing for measurement: when enabled, every stage launched by
   hd_euler_step/hd_update/hd_step is bracketed by HIP events recorded on the
   caller's stream.  hd_timing_read synchronises, returns the summed stage time
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
