/*
 * hd_oracle.h -- CPU restatement of HERMESHD's DG right-hand side and SSP-RK
 * stage update.  TEST INFRASTRUCTURE ONLY: this library is the checker that
 * the HIP product path is compared against.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it; nothing in the product path
 * links or calls it.
 *
 * State layout: Q(nx,ny,nz,nQ,nbasis) in Fortran order, i.e. a C array
 * q[r][f][k][j][i] with i fastest (reference src/params.f90:163).
 * Field order rh,mx,my,mz,en[,pxx,pyy,pzz,pxy,pxz,pyz] (src/params.f90:42-47).
 * The 13-moment model (nQ = 13, no counterpart in the reference) orders its
 * fields rh,mx,my,mz,en,pxx,pyy,pxy,pxz,pyz,qx,qy,qz: the traceless stress
 * (pzz = -pxx - pyy) and the heat flux.
 */
#ifndef HD_ORACLE_H
#define HD_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_PERIODIC = 0, ORC_BC_OUTFLOW = 1, ORC_BC_WALL = 2, ORC_BC_NOSLIP = 3, ORC_BC_INFLOW = 4 };
enum { ORC_FLUX_CENTRAL = 0, ORC_FLUX_HLLC = 1, ORC_FLUX_ROE = 2 };

typedef struct orc_config {
    int nx, ny, nz;
    int nbasis, iquad, nQ;       /* (nbasis,iquad) in {(4,2),(8,2),(10,3)}; nQ in {5,11,13} */
    int ivis, ieos;              /* ivis in {0,1,2}; ieos 1 (ideal gas) or 2 (Murnaghan-Tait water) */
    int riemann;                 /* ORC_FLUX_* */
    int limiter;                 /* 1: positivity limiter on every stage input */
    int integrator;              /* 2 = heun (SSP-RK2), 3 = shu-osher (SSP-RK3) */
    int bc[6];                   /* xlo,xhi,ylo,yhi,zlo,zhi : ORC_BC_* */
    int noise;                   /* 1: stochastic stress enters the flux (needs nQ = 11) */
    int pad_;
    double lx, ly, lz;           /* domain lengths; grid is centred on the origin */
    double gamma;                /* aindex */
    double rh_floor, T_floor;    /* P_floor = T_floor*rh_floor */
    double coll, colvis;         /* relaxation rate and coll*viscosity */
    double cflm;                 /* CFL number used by get_min_dt */
    double body_force[3];        /* constant acceleration (not in the reference) */
    double eta_sd, zeta_sd;      /* noise amplitudes (stochastic stress) */
    uint64_t seed;               /* noise key */
    uint32_t realization, pad2_;
    double eos_p1, eos_pbase;    /* ieos = 2: P = P_1*(rho^7.2 - 1) + P_base + P_ideal */
    double coll_q, col_q;        /* nQ = 13: heat-flux relaxation rate and coll_q*conductivity */
} orc_config;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_config *cfg);
void orc_destroy(orc_ctx *c);
int orc_set_step_index(orc_ctx *c, uint64_t step);    /* noise counter; orc_update advances it by one */
int orc_set_noise_dt(orc_ctx *c, double dt);          /* dt of the noise amplitude in orc_rhs (default 1) */
/* linear-response hook: while `on`, every point of the stochastic stress is
   zero except component comp (xx,yy,zz,xy,xz,yz) of point id of a stream
   (1..3 faces, 4 volume), which is `value`; amplitudes are not applied */
int orc_set_noise_impulse(orc_ctx *c, int on, uint32_t stream, uint64_t id, int comp, double value);
/* prescribed exterior state of an x-low inflow boundary: def(ny,nz,nface,nQ)
   in Fortran order (Qxlo_ext_def); rh, mx and en are read */
int orc_set_inflow(orc_ctx *c, const double *def);
/* immersed-solid mask (nx,ny,nz) in Fortran order, 1 = solid; NULL clears.
   orc_update applies it after every step (orc_apply_mask) */
int orc_set_mask(orc_ctx *c, const uint8_t *mask, double rho_mask);
int orc_apply_mask(orc_ctx *c, double *q);
/* output_vtk's derived fields at nvtk^3 points per cell: out has one
   (nz*nvtk, ny*nvtk, nx*nvtk) plane per selected field (bits of `fields`,
   order density, log density, vx, vy, vz, temperature, internal energy,
   total energy, entropy, pressure, pxx, pyy, pzz, vorticity) */
int orc_sample(const orc_ctx *c, const double *q, int nvtk, uint32_t fields, const double *scales, float *out);

/* tables (for pinning tests); arrays sized by the caller */
int orc_tables(const orc_ctx *c, double *xq, double *w1, double *w2, double *w3,
               double *cbasis, double *bint, double *bxp, double *bxm, double *byp,
               double *bym, double *bzp, double *bzm, double *bfaces);

/* operators; all arrays are full state arrays unless stated */
int orc_limiter(orc_ctx *c, double *q);                              /* in place */
int orc_glflux(orc_ctx *c, const double *q, double *glflux);         /* no limiter */
int orc_source(orc_ctx *c, const double *q, double *source);
int orc_rhs(orc_ctx *c, const double *q, double *rhs);               /* -glflux+source, no limiter */
int orc_euler_step(orc_ctx *c, double *q_in, double *q_out, double dt);
double orc_min_dt(orc_ctx *c, const double *q);
int orc_update(orc_ctx *c, double *q_io, double *q1, double *q2, double dt);
int orc_step(orc_ctx *c, double *q_io, double *q1, double *q2, double *t, double *dt);

/* point kernels */
int orc_flux_points(const orc_ctx *c, const double *qp, double *fp, int dir, int npnts);
int orc_riemann_face(const orc_ctx *c, const double *qlr, const double *flr, double *fout,
                     int dir, int kind);
double orc_cfcal(const orc_ctx *c, const double *q, int dir);

/* Philox-4x32-10 (Salmon et al., SC'11) and the noise stream built on it */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_noise_normals(uint64_t seed, uint32_t realization, uint64_t step, uint32_t stream,
                       uint64_t index, double out[4]);
/* stochastic stress (xx,yy,zz,xy,xz,yz) of point id in a stream at the current step index */
int orc_noise_stress(const orc_ctx *c, uint32_t stream, uint64_t id, double amp_eta, double amp_zeta, double S[6]);

/* threads of the cell and face loops when built with OpenMP (returns the
   count in use; 1 without OpenMP) */
int orc_set_threads(int n);
const char *orc_last_error(void);
int64_t orc_last_bad_cell(void);

#ifdef __cplusplus
}
#endif
#endif
