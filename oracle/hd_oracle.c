/*
 * hd_oracle.c -- plain-C restatement of HERMESHD's hot path, used ONLY as the
 * checker for the HIP product path (tests/, the smoke check, and the
 * CPU-baseline leg of bench.py).  It follows the reference loop by loop so
 * that its arithmetic order matches the Fortran (compiled without FP
 * contraction); every function cites the reference routine it restates.
 *
 * Reference files (paths relative to the HERMESHD repository):
 *   src/basis_funcs.f90   tables          src/flux.f90        fluxes, HLLC, limiter
 *   src/initialize.f90    flags, cbasis   src/sources.f90     relaxation sources
 *   src/integrator.f90    RK, update      src/prepare_step.f90, src/boundary.f90  traces, BCs
 *   src/hermeshd.f90      step, get_min_dt
 *   src/ces/10_moment/dg_3D_hydro_lqc_SEAN.f90  flux_roe (legacy, only Roe solver)
 *
 * Deliberate differences from the shipped reference (each also applies to
 * the product path):
 *   - FP64 throughout (the reference must be built with -fdefault-real-8).
 *   - x-high outflow copies the interior trace before editing it (the shipped
 *     xhibc_outflow, src/boundary.f90:246-282, leaves the exterior at 0).
 *   - periodic boundaries wrap on the device grid (the reference gets the
 *     wrap from MPI_CART_CREATE periods, src/initialize.f90:254-266).
 *   - optional limiter-off switch and body force (not in the reference).
 *   - noise: Philox-4x32-10 stream, one draw per face point shared by both
 *     sides and one per volume point shared by the three directions.
 */
#include "hd_oracle.h"
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NBMAX 10
#define QMAX 3
#define NFMAX (QMAX * QMAX)
#define NPGMAX (QMAX * QMAX * QMAX)
#define NQMAX 13

/* field indices, 0-based (src/params.f90:42-47) */
enum { RH = 0, MX = 1, MY = 2, MZ = 3, EN = 4, PXX = 5, PYY = 6, PZZ = 7, PXY = 8, PXZ = 9, PYZ = 10 };
/* 13-moment fields (no counterpart in the reference): traceless stress
   without pzz (= -pxx - pyy) and the heat flux */
enum { T_PXX = 5, T_PYY = 6, T_PXY = 7, T_PXZ = 8, T_PYZ = 9, T_QX = 10, T_QY = 11, T_QZ = 12 };
/* mode indices, 0-based (src/initialize.f90:126-194 with ibitri = 0 or nbasis = 8) */
enum { K1 = 0, KX = 1, KY = 2, KZ = 3, KYZ = 4, KZX = 5, KXY = 6, KXYZ = 7, KXX = 7, KYY = 8, KZZ = 9 };

static __thread char g_err[256];
static __thread int64_t g_bad_cell = -1;

static int set_err(int code, const char *msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return code;
}
const char *orc_last_error(void) { return g_err; }

/* threads of the cell and face loops (OpenMP builds; 1 otherwise): results do
   not depend on it, since every element is computed by the same code */
int orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
    return n > 0 ? n : 1;
#else
    (void)n;
    return 1;
#endif
}
int64_t orc_last_bad_cell(void) { return g_bad_cell; }

struct orc_ctx {
    orc_config c;
    int q, nb, nQ, nface, npg, npge, nfe, ibitri;
    long ncell;
    double xq[QMAX], w1[QMAX], w2[NFMAX], w3[NPGMAX];
    double cb[NBMAX];
    double bint[NPGMAX][NBMAX];
    double bxp[NFMAX][NBMAX], bxm[NFMAX][NBMAX], byp[NFMAX][NBMAX], bym[NFMAX][NBMAX];
    double bzp[NFMAX][NBMAX], bzm[NFMAX][NBMAX];
    double bff[6 * NFMAX + NPGMAX][NBMAX];
    double bwint[NPGMAX][NBMAX];
    double dx, dy, dz, dxi, dyi, dzi;
    double aindex, aindm1, P_floor;
    uint64_t step_index;
    double noise_dt;
    /* test hook: a unit stress impulse replaces the drawn noise field */
    int imp_on, imp_comp;
    uint32_t imp_stream;
    uint64_t imp_id;
    double imp_val;
    /* boundary traces: [face][a][b][nface][nQ] flattened per side */
    double *ext[6], *intr[6];
    /* face fluxes flux_x(nface,nx+1,ny,nz,nQ) etc. (src/flux.f90:11-13) */
    double *fx, *fy, *fz;
    double *glf, *src, *integ;
    double *inflow;         /* x-low inflow exterior state (ny,nz,nface,nQ), or NULL */
    uint8_t *mask;          /* immersed-solid cells, or NULL */
    double rho_mask;
};

static inline long qidx(const orc_ctx *c, int i, int j, int k, int f, int r) {
    const orc_config *g = &c->c;
    return (long)i + (long)g->nx * ((long)j + (long)g->ny * ((long)k + (long)g->nz * ((long)f + (long)c->nQ * r)));
}

/* ------------------------------------------------------------------------ */
/* tables: src/basis_funcs.f90:265-627, src/initialize.f90:88-122,199-220    */
/* ------------------------------------------------------------------------ */
static double p2(double s) { return 1.5 * (s * s) - 0.5; }

static void fill_products(double *v, int nb) {
    /* v[KX..KZ] set; products as set_internal_vals_3D / set_face_vals_3D */
    v[KYZ] = v[KY] * v[KZ];
    v[KZX] = v[KZ] * v[KX];
    v[KXY] = v[KX] * v[KY];
    if (nb == 8) v[KXYZ] = v[KX] * v[KY] * v[KZ];
    if (nb == 10) {
        v[KXX] = p2(v[KX]);
        v[KYY] = p2(v[KY]);
        v[KZZ] = p2(v[KZ]);
    }
}

static void build_tables(orc_ctx *c) {
    int q = c->q, nb = c->nb;
    if (q == 2) {
        double s = 1. / sqrt(3.);
        c->xq[0] = -s; c->xq[1] = s;
        c->w1[0] = c->w1[1] = 1.;
    } else {
        double s = sqrt(15.) / 5.;
        c->xq[0] = -s; c->xq[1] = 0.; c->xq[2] = s;
        c->w1[0] = 5. / 9.; c->w1[1] = 8. / 9.; c->w1[2] = 5. / 9.;
    }
    /* weights: set_weights_3D (src/basis_funcs.f90:570-627) */
    for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++) c->w2[i * q + j] = (q == 2) ? 1. : c->w1[j] * c->w1[i];
    for (int i = 0; i < q; i++)
        for (int p = 0; p < q * q; p++) c->w3[i * q * q + p] = (q == 2) ? 1. : c->w2[p] * c->w1[i];
    /* cbasis: set_cbasis_3D (src/initialize.f90:199-220) */
    c->cb[K1] = 1.;
    c->cb[KX] = c->cb[KY] = c->cb[KZ] = 3.;
    if (nb > 4) c->cb[KYZ] = c->cb[KZX] = c->cb[KXY] = 9.;
    if (nb == 8) c->cb[KXYZ] = 27.;
    if (nb == 10) c->cb[KXX] = c->cb[KYY] = c->cb[KZZ] = 5.;
    /* interior: point (i,j,k) -> x=xq[i], y=xq[j], z=xq[k], z fastest */
    for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++)
            for (int k = 0; k < q; k++) {
                double *v = c->bint[(i * q + j) * q + k];
                memset(v, 0, sizeof(double) * NBMAX);
                v[K1] = 1.; v[KX] = c->xq[i]; v[KY] = c->xq[j]; v[KZ] = c->xq[k];
                fill_products(v, nb);
            }
    /* faces: set_face_vals_3D.  z-face point (i,j): x=xq[j], y=xq[i];
       y-face point (i,j): x=xq[i], z=xq[j]; x-face point (i,j): y=xq[j], z=xq[i] */
    for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++) {
            int p = i * q + j;
            double *v;
            v = c->bzp[p]; memset(v, 0, sizeof(double) * NBMAX);
            v[K1] = 1.; v[KX] = c->xq[j]; v[KY] = c->xq[i]; v[KZ] = 1.; fill_products(v, nb);
            v = c->byp[p]; memset(v, 0, sizeof(double) * NBMAX);
            v[K1] = 1.; v[KX] = c->xq[i]; v[KY] = 1.; v[KZ] = c->xq[j]; fill_products(v, nb);
            v = c->bxp[p]; memset(v, 0, sizeof(double) * NBMAX);
            v[K1] = 1.; v[KX] = 1.; v[KY] = c->xq[j]; v[KZ] = c->xq[i]; fill_products(v, nb);
            /* minus faces: sign flips of modes odd in the normal coordinate */
            memcpy(c->bxm[p], c->bxp[p], sizeof(double) * NBMAX);
            c->bxm[p][KX] = -c->bxp[p][KX]; c->bxm[p][KZX] = -c->bxp[p][KZX]; c->bxm[p][KXY] = -c->bxp[p][KXY];
            if (nb == 8) c->bxm[p][KXYZ] = -c->bxp[p][KXYZ];
            memcpy(c->bym[p], c->byp[p], sizeof(double) * NBMAX);
            c->bym[p][KY] = -c->byp[p][KY]; c->bym[p][KYZ] = -c->byp[p][KYZ]; c->bym[p][KXY] = -c->byp[p][KXY];
            if (nb == 8) c->bym[p][KXYZ] = -c->byp[p][KXYZ];
            memcpy(c->bzm[p], c->bzp[p], sizeof(double) * NBMAX);
            c->bzm[p][KZ] = -c->bzp[p][KZ]; c->bzm[p][KYZ] = -c->bzp[p][KYZ]; c->bzm[p][KZX] = -c->bzp[p][KZX];
            if (nb == 8) c->bzm[p][KXYZ] = -c->bzp[p][KXYZ];
        }
    /* bf_faces rows: xm, xp, ym, yp, zm, zp, interior (src/basis_funcs.f90:527-546) */
    int nf = q * q;
    for (int p = 0; p < nf; p++) {
        double (*tabs[6])[NBMAX] = {c->bxm, c->bxp, c->bym, c->byp, c->bzm, c->bzp};
        for (int s = 0; s < 6; s++) {
            double *v = c->bff[s * nf + p];
            memset(v, 0, sizeof(double) * NBMAX);
            v[K1] = 1.; v[KX] = tabs[s][p][KX]; v[KY] = tabs[s][p][KY]; v[KZ] = tabs[s][p][KZ];
            fill_products(v, nb);
        }
    }
    for (int g = 0; g < q * q * q; g++) {
        memcpy(c->bff[6 * nf + g], c->bint[g], sizeof(double) * NBMAX);
        for (int r = 0; r < nb; r++) c->bwint[g][r] = c->w3[g] * c->bint[g][r]; /* init_bf_weights */
    }
}

orc_ctx *orc_create(const orc_config *cfg) {
    int ok = (cfg->nbasis == 4 && cfg->iquad == 2) || (cfg->nbasis == 8 && cfg->iquad == 2) ||
             (cfg->nbasis == 10 && cfg->iquad == 3);
    if (!ok) { set_err(-1, "unsupported (nbasis, iquad)"); return NULL; }
    if (cfg->nQ != 5 && cfg->nQ != 11 && cfg->nQ != 13) { set_err(-1, "nQ must be 5, 11 or 13"); return NULL; }
    if (cfg->ivis < 0 || cfg->ivis > 2) { set_err(-1, "ivis must be 0, 1 or 2"); return NULL; }
    if (cfg->ivis >= 1 && cfg->nQ == 5) { set_err(-1, "ivis=1,2 need the stress fields"); return NULL; }
    if (cfg->nQ == 13 && (cfg->ivis == 2 || cfg->noise || cfg->ieos != 1)) {
        set_err(-1, "the 13-moment model runs with ivis 0 or 1, an ideal gas and no noise");
        return NULL;
    }
    if (cfg->ivis == 2 && cfg->noise) { set_err(-1, "ivis=2 has no stochastic stress"); return NULL; }
    if (cfg->noise && cfg->nQ != 11) { set_err(-1, "noise needs nQ=11"); return NULL; }
    if (cfg->ieos != 1 && cfg->ieos != 2) { set_err(-1, "ieos must be 1 or 2"); return NULL; }
    if (cfg->ieos == 2 && cfg->riemann == ORC_FLUX_ROE) { set_err(-1, "ieos=2 has no Roe solver"); return NULL; }
    for (int s = 0; s < 6; s++) {
        if (cfg->bc[s] < 0 || cfg->bc[s] > ORC_BC_INFLOW) { set_err(-1, "unknown boundary kind"); return NULL; }
        if (cfg->bc[s] == ORC_BC_INFLOW && s != 0) { set_err(-1, "inflow is an x-low boundary"); return NULL; }
    }
    for (int d = 0; d < 3; d++) {
        int lo = cfg->bc[2 * d], hi = cfg->bc[2 * d + 1];
        if ((lo == ORC_BC_PERIODIC) != (hi == ORC_BC_PERIODIC)) {
            set_err(-1, "periodic boundaries must be periodic on both sides");
            return NULL;
        }
    }
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->c = *cfg;
    c->q = cfg->iquad; c->nb = cfg->nbasis; c->nQ = cfg->nQ;
    c->nface = c->q * c->q; c->npg = c->nface * c->q; c->npge = 6 * c->nface; c->nfe = 2 * c->nface;
    c->ibitri = (cfg->nbasis == 8) ? 1 : 0;
    c->ncell = (long)cfg->nx * cfg->ny * cfg->nz;
    build_tables(c);
    /* init_spatial_params (src/initialize.f90:298-327), one rank */
    double lxu = cfg->lx / 2.0, lxd = -(cfg->lx / 2.0);
    double lyu = cfg->ly / 2.0, lyd = -(cfg->ly / 2.0);
    double lzu = cfg->lz / 2.0, lzd = -(cfg->lz / 2.0);
    c->dxi = (double)cfg->nx / (lxu - lxd);
    c->dyi = (double)cfg->ny / (lyu - lyd);
    c->dzi = (double)cfg->nz / (lzu - lzd);
    c->dx = 1. / c->dxi; c->dy = 1. / c->dyi; c->dz = 1. / c->dzi;
    c->aindex = cfg->gamma;
    c->aindm1 = cfg->gamma - 1.;
    c->P_floor = cfg->T_floor * cfg->rh_floor;
    c->noise_dt = 1.0;
    int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz, nf = c->nface, nQ = c->nQ;
    long sz[6] = {(long)ny * nz, (long)ny * nz, (long)nx * nz, (long)nx * nz, (long)nx * ny, (long)nx * ny};
    for (int s = 0; s < 6; s++) {
        c->ext[s] = (double *)calloc((size_t)sz[s] * nf * nQ, sizeof(double));
        c->intr[s] = (double *)calloc((size_t)sz[s] * nf * nQ, sizeof(double));
    }
    c->fx = (double *)calloc((size_t)nf * (nx + 1) * ny * nz * nQ, sizeof(double));
    c->fy = (double *)calloc((size_t)nf * nx * (ny + 1) * nz * nQ, sizeof(double));
    c->fz = (double *)calloc((size_t)nf * nx * ny * (nz + 1) * nQ, sizeof(double));
    c->glf = (double *)calloc((size_t)c->ncell * nQ * c->nb, sizeof(double));
    c->src = (double *)calloc((size_t)c->ncell * nQ * c->nb, sizeof(double));
    c->integ = (double *)calloc((size_t)c->ncell * nQ * c->nb, sizeof(double));
    return c;
}

void orc_destroy(orc_ctx *c) {
    if (!c) return;
    for (int s = 0; s < 6; s++) { free(c->ext[s]); free(c->intr[s]); }
    free(c->fx); free(c->fy); free(c->fz); free(c->glf); free(c->src); free(c->integ);
    free(c->inflow); free(c->mask);
    free(c);
}

int orc_set_inflow(orc_ctx *c, const double *def) {
    size_t n = (size_t)c->c.ny * c->c.nz * c->nface * c->nQ;
    if (!c->inflow) c->inflow = (double *)malloc(n * sizeof(double));
    memcpy(c->inflow, def, n * sizeof(double));
    return 0;
}

int orc_set_mask(orc_ctx *c, const uint8_t *mask, double rho_mask) {
    free(c->mask);
    c->mask = NULL;
    if (!mask) return 0;
    c->mask = (uint8_t *)malloc((size_t)c->ncell);
    memcpy(c->mask, mask, (size_t)c->ncell);
    c->rho_mask = rho_mask;
    return 0;
}

/* immersed solid (apply_cyl_in_2d_pipe_boundaries, src/boundary.f90:686-700,
   with its intent: the mode loop there runs over nQ): in masked cells the
   density is rho_mask (mode 1) with no slope, and every momentum mode is 0 */
int orc_apply_mask(orc_ctx *c, double *q) {
    if (!c->mask) return 0;
    const orc_config *g = &c->c;
    for (int k = 0; k < g->nz; k++)
        for (int j = 0; j < g->ny; j++)
            for (int i = 0; i < g->nx; i++) {
                long e = i + (long)g->nx * (j + (long)g->ny * k);
                if (!c->mask[e]) continue;
                for (int r = 0; r < c->nb; r++) {
                    q[qidx(c, i, j, k, RH, r)] = r == 0 ? c->rho_mask : 0.0;
                    q[qidx(c, i, j, k, MX, r)] = 0.0;
                    q[qidx(c, i, j, k, MY, r)] = 0.0;
                    q[qidx(c, i, j, k, MZ, r)] = 0.0;
                }
            }
    return 0;
}

int orc_set_step_index(orc_ctx *c, uint64_t step) { c->step_index = step; return 0; }
int orc_set_noise_dt(orc_ctx *c, double dt) { c->noise_dt = dt; return 0; }
int orc_set_noise_impulse(orc_ctx *c, int on, uint32_t stream, uint64_t id, int comp, double value) {
    if (on && (stream < 1 || stream > 4 || comp < 0 || comp > 5)) return set_err(-1, "impulse stream or component out of range");
    c->imp_on = on; c->imp_stream = stream; c->imp_id = id; c->imp_comp = comp; c->imp_val = value;
    return 0;
}

int orc_tables(const orc_ctx *c, double *xq, double *w1, double *w2, double *w3, double *cbasis,
               double *bint, double *bxp, double *bxm, double *byp, double *bym, double *bzp,
               double *bzm, double *bfaces) {
    int q = c->q, nb = c->nb, nf = c->nface, npg = c->npg;
    for (int i = 0; i < q; i++) { xq[i] = c->xq[i]; w1[i] = c->w1[i]; }
    for (int p = 0; p < nf; p++) w2[p] = c->w2[p];
    for (int g = 0; g < npg; g++) w3[g] = c->w3[g];
    for (int r = 0; r < nb; r++) cbasis[r] = c->cb[r];
    for (int g = 0; g < npg; g++)
        for (int r = 0; r < nb; r++) bint[g * nb + r] = c->bint[g][r];
    for (int p = 0; p < nf; p++)
        for (int r = 0; r < nb; r++) {
            bxp[p * nb + r] = c->bxp[p][r]; bxm[p * nb + r] = c->bxm[p][r];
            byp[p * nb + r] = c->byp[p][r]; bym[p * nb + r] = c->bym[p][r];
            bzp[p * nb + r] = c->bzp[p][r]; bzm[p * nb + r] = c->bzm[p][r];
        }
    for (int p = 0; p < 6 * nf + npg; p++)
        for (int r = 0; r < nb; r++) bfaces[p * nb + r] = c->bff[p][r];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller          */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += W0; k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Single-precision Box-Muller (the reference draws its Gaussians in single
   precision, vsRngGaussian, src/random.f90:61,87), from correctly rounded
   FP32 operations only, in a fixed order with no contraction (this file is
   built with -ffp-contract=off), so that every IEEE FP32 implementation of
   the same steps gives the same bits:
     log a = e ln2 + 2 atanh(s), a = m 2^e, m in [sqrt(1/2), sqrt(2)),
             s = (m-1)/(m+1), atanh by its odd series to s^9 (|s| < 0.172);
     sin/cos(pi x): x = n/2 + r, |r| <= 1/4, Taylor series of sin and cos of
             pi r to degree 9 and 10, then the signs of quadrant n. */
static float bm_logf(float a) {
    uint32_t bits;
    memcpy(&bits, &a, 4);
    int e = (int)((bits >> 23) & 0xffu) - 127;
    bits = (bits & 0x007fffffu) | 0x3f800000u;
    float m;
    memcpy(&m, &bits, 4);
    if (m > 0x1.6a09e6p+0f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float s = (m - 1.0f) / (m + 1.0f);
    const float s2 = s * s;
    float p = 0x1.c71c72p-4f;
    p = p * s2 + 0x1.24924ap-3f;
    p = p * s2 + 0x1.99999ap-3f;
    p = p * s2 + 0x1.555556p-2f;
    p = p * s2 + 1.0f;
    const float lm = 2.0f * s * p;
    return (float)e * 0x1.62e430p-1f + lm;
}

static void bm_sincospi(float x, float *sv, float *cv) {
    const float nf = floorf(2.0f * x + 0.5f);
    const int n = (int)nf;
    const float r = x - nf * 0.5f;
    const float z = r * 0x1.921fb6p+1f;
    const float z2 = z * z;
    float sp = 0x1.71de3ap-19f;
    sp = sp * z2 - 0x1.a01a02p-13f;
    sp = sp * z2 + 0x1.111112p-7f;
    sp = sp * z2 - 0x1.555556p-3f;
    sp = sp * z2 + 1.0f;
    const float sn = z * sp;
    float cp = -0x1.27e4fcp-22f;
    cp = cp * z2 + 0x1.a01a02p-16f;
    cp = cp * z2 - 0x1.6c16c2p-10f;
    cp = cp * z2 + 0x1.555556p-5f;
    cp = cp * z2 - 0.5f;
    const float cs = cp * z2 + 1.0f;
    const int qd = n & 3;
    *sv = qd == 0 ? sn : (qd == 1 ? cs : (qd == 2 ? -sn : -cs));
    *cv = qd == 0 ? cs : (qd == 1 ? -sn : (qd == 2 ? -cs : sn));
}

/* Four N(0,1) numbers for one counter value.  Counter layout:
   ctr = (index lo, index hi | stream<<24, step lo, step hi), key = seed ^ (realization mix).
   Uniforms in (0,1]: (u + 0.5) * 2^-32 in FP32; Box-Muller on pairs. */
void orc_noise_normals(uint64_t seed, uint32_t realization, uint64_t step, uint32_t stream,
                       uint64_t index, double out[4]) {
    uint32_t ctr[4] = {(uint32_t)index, (uint32_t)(index >> 32) | (stream << 24), (uint32_t)step,
                       (uint32_t)(step >> 32)};
    uint32_t key[2] = {(uint32_t)seed ^ (realization * 0x9E3779B9u), (uint32_t)(seed >> 32) ^ realization};
    uint32_t u[4];
    orc_philox4x32_10(ctr, key, u);
    for (int h = 0; h < 2; h++) {
        const float a = ((float)u[2 * h] + 0.5f) * 0x1p-32f;
        const float b = ((float)u[2 * h + 1] + 0.5f) * 0x1p-32f;
        const float rad = sqrtf(-2.0f * bm_logf(a));
        float sv, cv;
        bm_sincospi(2.0f * b, &sv, &cv);
        out[2 * h] = (double)(rad * cv);
        out[2 * h + 1] = (double)(rad * sv);
    }
}

/* ------------------------------------------------------------------------ */
/* physical flux at points: flux_calc_pnts_r (src/flux.f90:23-193)           */
/* S: optional symmetric stochastic stress per point, 6 comps xx,yy,zz,xy,xz,yz */
/* ------------------------------------------------------------------------ */
static void flux_points_s(const orc_ctx *c, const double *qp, double *fp, int dir, int npnts,
                          const double *S) {
    const int nQ = c->nQ;
    const double colvis = c->c.colvis;
    const double c2d3cv = (2. / 3.) * colvis, c4d3cv = (4. / 3.) * colvis;
    for (int p = 0; p < npnts; p++) {
        const double *Q = qp + (long)p * nQ;
        double *F = fp + (long)p * nQ;
        double dn = Q[RH], M_x = Q[MX], M_y = Q[MY], M_z = Q[MZ], Ener = Q[EN];
        double P_xx = 0, P_yy = 0, P_zz = 0, P_xy = 0, P_xz = 0, P_yz = 0;
        if (nQ == 11) { P_xx = Q[PXX]; P_yy = Q[PYY]; P_zz = Q[PZZ]; P_xy = Q[PXY]; P_xz = Q[PXZ]; P_yz = Q[PYZ]; }
        double Sxx = 0, Syy = 0, Szz = 0, Sxy = 0, Sxz = 0, Syz = 0;
        if (S) { Sxx = S[6 * p + 0]; Syy = S[6 * p + 1]; Szz = S[6 * p + 2]; Sxy = S[6 * p + 3]; Sxz = S[6 * p + 4]; Syz = S[6 * p + 5]; }
        double dni = 1. / dn;
        double vx = M_x * dni, vy = M_y * dni, vz = M_z * dni;
        double smsq = M_x * M_x + M_y * M_y + M_z * M_z;
        double P = c->aindm1 * (Ener - 0.5 * dni * smsq);
        /* Murnaghan-Tait water (src/flux.f90:64, src/params.f90:97-105) */
        if (c->c.ieos == 2) P = c->c.eos_p1 * (pow(dn, 7.2) - 1.) + c->c.eos_pbase + P;
        if (P < c->P_floor) P = c->P_floor;
        if (nQ == 13) {
            /* 13-moment relaxation model (no counterpart in the reference): the
               stress rows of the ivis 0/1 model (src/flux.f90:101-106,138-143,
               175-180) for the five stored components, pzz = -pxx - pyy, and a
               Maxwell-Cattaneo heat flux, d_t q_i + d_i(col_q T) = -coll_q q_i
               with T = P/rho, whose q_k enters the energy flux */
            const double T_xx = Q[T_PXX], T_yy = Q[T_PYY], T_zz = -(T_xx + T_yy);
            const double T_xy = Q[T_PXY], T_xz = Q[T_PXZ], T_yz = Q[T_PYZ];
            const double ct = c->c.col_q * (P * dni);
            switch (dir) {
            case 1:
                F[RH] = M_x;
                F[MX] = M_x * vx + P + T_xx;
                F[MY] = M_y * vx + T_xy;
                F[MZ] = M_z * vx + T_xz;
                F[EN] = (Ener + P) * vx + T_xx * vx + T_xy * vy + T_xz * vz + Q[T_QX];
                F[T_PXX] = c4d3cv * vx; F[T_PYY] = -c2d3cv * vx;
                F[T_PXY] = colvis * vy; F[T_PXZ] = colvis * vz; F[T_PYZ] = 0;
                F[T_QX] = ct; F[T_QY] = 0; F[T_QZ] = 0;
                break;
            case 2:
                F[RH] = M_y;
                F[MX] = M_x * vy + T_xy;
                F[MY] = M_y * vy + P + T_yy;
                F[MZ] = M_z * vy + T_yz;
                F[EN] = (Ener + P) * vy + T_yy * vy + T_xy * vx + T_yz * vz + Q[T_QY];
                F[T_PXX] = -c2d3cv * vy; F[T_PYY] = c4d3cv * vy;
                F[T_PXY] = colvis * vx; F[T_PXZ] = 0; F[T_PYZ] = colvis * vz;
                F[T_QX] = 0; F[T_QY] = ct; F[T_QZ] = 0;
                break;
            default:
                F[RH] = M_z;
                F[MX] = M_x * vz + T_xz;
                F[MY] = M_y * vz + T_yz;
                F[MZ] = M_z * vz + P + T_zz;
                F[EN] = (Ener + P) * vz + T_zz * vz + T_xz * vx + T_yz * vy + Q[T_QZ];
                F[T_PXX] = -c2d3cv * vz; F[T_PYY] = -c2d3cv * vz;
                F[T_PXY] = 0; F[T_PXZ] = colvis * vx; F[T_PYZ] = colvis * vy;
                F[T_QX] = 0; F[T_QY] = 0; F[T_QZ] = ct;
                break;
            }
            continue;
        }
        if (c->c.ivis == 2) {
            /* full 10-moment model (src/flux.f90:64-66,89-117,134-155,172-191):
               the stress fields hold the second moments E_ij = rho u_i u_j + P_ij;
               the momentum flux swaps the isotropic 10-moment pressure P_10 for
               the equation-of-state pressure P_5 */
            double P_10 = (1. / 3.) * (P_xx + P_yy + P_zz - dni * smsq);
            if (P_10 < c->P_floor) P_10 = c->P_floor;
            const double P_5 = P;
            switch (dir) {
            case 1:
                F[RH] = M_x;
                F[MX] = P_xx - P_10 + P_5; F[MY] = P_xy; F[MZ] = P_xz;
                F[EN] = (Ener + P) * vx + P_xx * vx + P_xy * vy + P_xz * vz;
                F[PXX] = vx * (3 * P_xx - 2 * vx * M_x);
                F[PYY] = 2 * vy * (P_xy - vy * M_x) + vx * P_yy;
                F[PZZ] = 2 * vz * (P_xz - vz * M_x) + vx * P_zz;
                F[PXY] = 2 * vx * (P_xy - vy * M_x) + vy * P_xx;
                F[PXZ] = 2 * vx * (P_xz - vz * M_x) + vz * P_xx;
                F[PYZ] = vx * P_yz + vy * P_xz + vz * P_xy - 2 * vy * vz * M_x;
                break;
            case 2:
                F[RH] = M_y;
                F[MX] = P_xy; F[MY] = P_yy - P_10 + P_5; F[MZ] = P_yz;
                F[EN] = (Ener + P) * vy + P_yy * vy + P_xy * vx + P_yz * vz;
                F[PXX] = 2 * vx * (P_xy - vx * M_y) + vy * P_xx;
                F[PYY] = vy * (3 * P_yy - 2 * vy * M_y);
                F[PZZ] = 2 * vz * (P_yz - vz * M_y) + vy * P_zz;
                F[PXY] = 2 * vy * (P_xy - vx * M_y) + vx * P_yy;
                F[PXZ] = vx * P_yz + vy * P_xz + vz * P_xy - 2 * vx * vz * M_y;
                F[PYZ] = 2 * vy * (P_yz - vz * M_y) + vz * P_yy;
                break;
            default:
                F[RH] = M_z;
                F[MX] = P_xz; F[MY] = P_yz; F[MZ] = P_zz - P_10 + P_5;
                F[EN] = (Ener + P) * vz + P_zz * vz + P_xz * vx + P_yz * vy;
                F[PXX] = 2 * vx * (P_xz - vx * M_z) + vz * P_xx;
                F[PYY] = 2 * vy * (P_yz - vy * M_z) + vz * P_yy;
                F[PZZ] = vz * (3 * P_zz - 2 * vz * M_z);
                F[PXY] = vx * P_yz + vy * P_xz + vz * P_xy - 2 * vx * vy * M_z;
                F[PXZ] = 2 * vz * (P_xz - vx * M_z) + vx * P_zz;
                F[PYZ] = 2 * vz * (P_yz - vy * M_z) + vy * P_zz;
                break;
            }
            continue;
        }
        switch (dir) {
        case 1:
            F[RH] = M_x;
            F[MX] = M_x * vx + P + P_xx - Sxx;
            F[MY] = M_y * vx + P_xy - Sxy;
            F[MZ] = M_z * vx + P_xz - Sxz;
            F[EN] = (Ener + P) * vx + (P_xx - Sxx) * vx + (P_xy - Sxy) * vy + (P_xz - Sxz) * vz;
            if (nQ >= 11) {
                F[PXX] = c4d3cv * vx; F[PYY] = -c2d3cv * vx; F[PZZ] = -c2d3cv * vx;
                F[PXY] = colvis * vy; F[PXZ] = colvis * vz; F[PYZ] = 0;
            }
            break;
        case 2:
            F[RH] = M_y;
            F[MX] = M_x * vy + P_xy - Sxy;
            F[MY] = M_y * vy + P + P_yy - Syy;
            F[MZ] = M_z * vy + P_yz - Syz;
            F[EN] = (Ener + P) * vy + (P_yy - Syy) * vy + (P_xy - Sxy) * vx + (P_yz - Syz) * vz;
            if (nQ >= 11) {
                F[PXX] = -c2d3cv * vy; F[PYY] = c4d3cv * vy; F[PZZ] = -c2d3cv * vy;
                F[PXY] = colvis * vx; F[PXZ] = 0; F[PYZ] = colvis * vz;
            }
            break;
        default:
            F[RH] = M_z;
            F[MX] = M_x * vz + P_xz - Sxz;
            F[MY] = M_y * vz + P_yz - Syz;
            F[MZ] = M_z * vz + P + P_zz - Szz;
            F[EN] = (Ener + P) * vz + (P_zz - Szz) * vz + (P_xz - Sxz) * vx + (P_yz - Syz) * vy;
            if (nQ >= 11) {
                F[PXX] = -c2d3cv * vz; F[PYY] = -c2d3cv * vz; F[PZZ] = c4d3cv * vz;
                F[PXY] = 0; F[PXZ] = colvis * vx; F[PYZ] = colvis * vy;
            }
            break;
        }
    }
}

int orc_flux_points(const orc_ctx *c, const double *qp, double *fp, int dir, int npnts) {
    flux_points_s(c, qp, fp, dir, npnts, NULL);
    return 0;
}

/* cfcal (src/flux.f90:1029-1072), ieos = 1; ivis 2 takes the pressure from
   the trace of the second moments */
double orc_cfcal(const orc_ctx *c, const double *Q, int dir) {
    double dn = Q[RH], dni = 1. / dn;
    double vx = Q[MX] * dni, vy = Q[MY] * dni, vz = Q[MZ] * dni;
    double P = c->aindm1 * (Q[EN] - 0.5 * dn * (vx * vx + vy * vy + vz * vz));
    if (c->c.ivis == 2) P = (1. / 3.) * (Q[PXX] + Q[PYY] + Q[PZZ] - dn * (vx * vx + vy * vy + vz * vz));
    double cs = c->c.ieos == 2 ? sqrt(7.2 * c->c.eos_p1 * pow(dn, 6.2)) : sqrt(c->aindex * P * dni);
    double v = (dir == 1) ? vx : (dir == 2) ? vy : vz;
    return fabs(v) + cs;
}

static void dir_fields(int dir, int *ipar, int *ip1, int *ip2) {
    /* iparr/iperp1/iperp2 (src/flux.f90:491-506) */
    if (dir == 1) { *ipar = MX; *ip1 = MY; *ip2 = MZ; }
    else if (dir == 2) { *ipar = MY; *ip1 = MZ; *ip2 = MX; }
    else { *ipar = MZ; *ip1 = MX; *ip2 = MY; }
}

/* flux_hllc (src/flux.f90:466-647) with ibatten = 0, ieos 1 or 2.
   qlr, flr: (nfe x nQ) rows 0..nface-1 left, nface..nfe-1 right; fout: (nface x 5) */
static void hllc_face(const orc_ctx *c, const double *qlr, const double *flr, double *fout, int dir) {
    const int nf = c->nface, nfe = c->nfe, nQ = c->nQ;
    int ipar, ip1, ip2;
    dir_fields(dir, &ipar, &ip1, &ip2);
    int ivar[5] = {RH, ipar, ip1, ip2, EN};
    double rhov[2 * NFMAX], v1[2 * NFMAX], plr[2 * NFMAX];
    double cslr[2 * NFMAX], qslr[2 * NFMAX], s_lr[2 * NFMAX];
    double sm_num[NFMAX], sm_den[NFMAX], s_m[NFMAX], pstar[NFMAX];
    const double aindex = c->aindex, aindm1 = c->aindm1;
    for (int k = 0; k < nfe; k++) {
        const double *Q = qlr + (long)k * nQ;
        rhov[k] = Q[RH];
        double rho_i = 1.0 / rhov[k];
        double a = Q[ipar] * rho_i, b = Q[ip1] * rho_i, d = Q[ip2] * rho_i;
        v1[k] = a;
        double qsq = a * a + b * b + d * d;
        plr[k] = aindm1 * (Q[EN] - 0.5 * rhov[k] * qsq);
        if (c->c.ieos == 2) plr[k] = c->c.eos_p1 * (pow(rhov[k], 7.2) - 1.) + c->c.eos_pbase + plr[k];
    }
    if (c->c.ieos == 2) {
        /* the water-EOS wave speeds use rho_i as the loop above leaves it:
           1/rho of the last right state, for every point (src/flux.f90:519-534,569-571) */
        const double rho_i = 1.0 / rhov[nfe - 1], p1 = c->c.eos_p1;
        for (int k = 0; k < nf; k++) {
            int k2 = k + nf;
            cslr[k] = v1[k] - sqrt(7.2 * p1 * pow(rhov[k], 6.2) + plr[k] * rho_i);
            cslr[k2] = v1[k2] + sqrt(7.2 * p1 * pow(rhov[k2], 6.2) + plr[k2] * rho_i);
            qslr[k] = v1[k2] - sqrt(7.2 * p1 * pow(rhov[k2], 6.2) + plr[k2] * rho_i);
            qslr[k2] = v1[k] + sqrt(7.2 * p1 * pow(rhov[k], 6.2) + plr[k] * rho_i);
        }
    } else {
        for (int k = 0; k < nf; k++) {
            int k2 = k + nf;
            cslr[k] = v1[k] - sqrt(aindex * plr[k] / rhov[k]);
            cslr[k2] = v1[k2] + sqrt(aindex * plr[k2] / rhov[k2]);
            qslr[k] = v1[k2] - sqrt(aindex * plr[k2] / rhov[k2]);
            qslr[k2] = v1[k] + sqrt(aindex * plr[k] / rhov[k]);
        }
    }
    for (int k = 0; k < nf; k++) {
        int k2 = k + nf;
        s_lr[k] = fmin(cslr[k], qslr[k]);
        s_lr[k2] = fmax(cslr[k2], qslr[k2]);
        sm_num[k] = rhov[k2] * v1[k2] * (s_lr[k2] - v1[k2]) - rhov[k] * v1[k] * (s_lr[k] - v1[k]);
        sm_num[k] = sm_num[k] + plr[k] - plr[k2];
        sm_den[k] = rhov[k2] * (s_lr[k2] - v1[k2]) - rhov[k] * (s_lr[k] - v1[k]);
    }
    for (int k = 0; k < nf; k++)
        if (sm_den[k] == 0.0) sm_den[k] = c->c.rh_floor;
    for (int k = 0; k < nf; k++) {
        s_m[k] = sm_num[k] / sm_den[k];
        pstar[k] = rhov[k] * (v1[k] - s_lr[k]) * (v1[k] - s_m[k]) + plr[k];
    }
    double fstar[2 * NFMAX][5];
    for (int k = 0; k < nfe; k++) {
        int i4 = (k < nf) ? k : k - nf;
        const double *Q = qlr + (long)k * nQ;
        const double *F = flr + (long)k * nQ;
        double den = s_lr[k] - s_m[i4];
        if (den == 0.0) den = c->c.rh_floor;
        double slrm_i = 1.0 / den;
        double sq = s_lr[k] - v1[k];
        double Qs[5];
        Qs[0] = rhov[k] * sq * slrm_i;
        Qs[1] = (sq * Q[ipar] + pstar[i4] - plr[k]) * slrm_i;
        Qs[2] = sq * Q[ip1] * slrm_i;
        Qs[3] = sq * Q[ip2] * slrm_i;
        Qs[4] = (sq * Q[EN] - plr[k] * v1[k] + pstar[i4] * s_m[i4]) * slrm_i;
        for (int e = 0; e < 5; e++) fstar[k][e] = F[ivar[e]] + s_lr[k] * (Qs[e] - Q[ivar[e]]);
    }
    for (int i4 = 0; i4 < nf; i4++) {
        double *out = fout + (long)i4 * 5;
        /* Eq. (26) of Batten 1997, in the reference's sequential-if order */
        if (s_lr[i4] > 0.0)
            for (int e = 0; e < 5; e++) out[ivar[e]] = flr[(long)i4 * nQ + ivar[e]];
        if (s_lr[i4] <= 0.0 && 0.0 < s_m[i4])
            for (int e = 0; e < 5; e++) out[ivar[e]] = fstar[i4][e];
        if (s_m[i4] <= 0.0 && 0.0 <= s_lr[i4 + nf])
            for (int e = 0; e < 5; e++) out[ivar[e]] = fstar[i4 + nf][e];
        if (s_lr[i4 + nf] < 0.0)
            for (int e = 0; e < 5; e++) out[ivar[e]] = flr[(long)(i4 + nf) * nQ + ivar[e]];
    }
}

/* flux_roe (src/ces/10_moment/dg_3D_hydro_lqc_SEAN.f90:1614-1789), Eulderink &
   Mellema.  kroe[p] = 0 where the Roe sound speed squared is not positive. */
static void roe_face(const orc_ctx *c, const double *qlr, const double *flr, double *fout, int dir,
                     int *kroe) {
    const int nf = c->nface, nQ = c->nQ;
    int ipar, ip1, ip2;
    dir_fields(dir, &ipar, &ip1, &ip2);
    const double aindm1 = c->aindm1, cp = c->aindex / aindm1;
    double skr[2][NFMAX], swr[2][NFMAX][4];
    for (int i9 = 0; i9 < nf; i9++) {
        kroe[i9] = 1;
        for (int j9 = 0; j9 < 2; j9++) {
            int ilr = (j9 == 0) ? i9 : i9 + nf;
            const double *Q = qlr + (long)ilr * nQ;
            double dnii = 1. / Q[RH];
            double vel0 = Q[ipar] * dnii, vel1 = Q[ip1] * dnii, vel2 = Q[ip2] * dnii;
            double vsq = vel0 * vel0 + vel1 * vel1 + vel2 * vel2;
            double Pri = aindm1 * (Q[EN] - 0.5 * Q[RH] * vsq);
            skr[j9][i9] = sqrt(Q[RH]);
            swr[j9][i9][0] = skr[j9][i9] * vel0;
            swr[j9][i9][1] = skr[j9][i9] * vel1;
            swr[j9][i9][2] = skr[j9][i9] * vel2;
            swr[j9][i9][3] = 0.5 * skr[j9][i9] * (vsq + cp * Pri / Q[RH]);
        }
    }
    for (int i9 = 0; i9 < nf; i9++) {
        const double *QL = qlr + (long)i9 * nQ, *QR = qlr + (long)(i9 + nf) * nQ;
        double Qri = 1. / QL[RH];
        double delrho = QR[RH] * Qri - 1.;
        double dele = (QR[EN] - QL[EN]) * Qri;
        double dm1 = (QR[ipar] - QL[ipar]) * Qri;
        double dm2 = (QR[ip1] - QL[ip1]) * Qri;
        double dm3 = (QR[ip2] - QL[ip2]) * Qri;
        double dwr[4];
        for (int j9 = 0; j9 < 4; j9++) dwr[j9] = 0.5 * (swr[1][i9][j9] + swr[0][i9][j9]);
        double dsk = 0.5 * (skr[1][i9] + skr[0][i9]);
        double vc0 = dwr[0] / dsk, vc1 = dwr[1] / dsk, vc2 = dwr[2] / dsk, hc = dwr[3] / dsk;
        double vcsq = vc0 * vc0 + vc1 * vc1 + vc2 * vc2;
        double asq = aindm1 * (hc - 0.5 * vcsq);
        if (asq <= 0.0) kroe[i9] = 0;
        double sc9 = sqrt(asq);
        double lam[5];
        lam[0] = fabs(vc0 - sc9); lam[1] = fabs(vc0 + sc9); lam[2] = fabs(vc0); lam[3] = lam[2]; lam[4] = lam[2];
        /* eigenvectors: rows (rho, energy, parallel, perp1, perp2) */
        double ev[5][5] = {{0}};
        ev[0][0] = 1.; ev[0][1] = 1.; ev[0][2] = 1.; ev[3][3] = 1.; ev[4][4] = 1.;
        ev[1][0] = hc - sc9 * vc0; ev[2][0] = vc0 - sc9; ev[3][0] = vc1; ev[4][0] = vc2;
        ev[1][1] = hc + sc9 * vc0; ev[2][1] = vc0 + sc9; ev[3][1] = vc1; ev[4][1] = vc2;
        ev[1][2] = 0.5 * vcsq; ev[2][2] = vc0; ev[3][2] = vc1; ev[4][2] = vc2;
        ev[1][3] = vc1; ev[1][4] = vc2;
        double ea2i = aindm1 / asq;
        double vcdel = vc0 * dm1 + vc1 * dm2 + vc2 * dm3 - dele;
        double a9e = 0.5 * ea2i * (0.5 * vcsq * delrho - vcdel);
        double sk9 = 0.5 * (dm1 - vc0 * delrho) / sc9;
        double a9[5];
        a9[0] = a9e - sk9; a9[1] = a9e + sk9;
        a9[2] = ea2i * ((hc - vcsq) * delrho + vcdel);
        a9[3] = dm2 - vc1 * delrho; a9[4] = dm3 - vc2 * delrho;
        double df[5] = {0, 0, 0, 0, 0};
        for (int j9 = 0; j9 < 5; j9++)
            for (int r = 0; r < 5; r++) df[r] = df[r] + a9[j9] * lam[j9] * ev[r][j9];
        for (int r = 0; r < 5; r++) df[r] = df[r] * QL[RH];
        double *out = fout + (long)i9 * 5;
        const double *FL = flr + (long)i9 * nQ, *FR = flr + (long)(i9 + nf) * nQ;
        if (kroe[i9] > 0) {
            out[RH] = 0.5 * (FL[RH] + FR[RH] - df[0]);
            out[EN] = 0.5 * (FL[EN] + FR[EN] - df[1]);
            out[ipar] = 0.5 * (FL[ipar] + FR[ipar] - df[2]);
            out[ip1] = 0.5 * (FL[ip1] + FR[ip1] - df[3]);
            out[ip2] = 0.5 * (FL[ip2] + FR[ip2] - df[4]);
        }
    }
}

/* numerical flux at one face: central average (src/flux.f90:257-263 with
   cfr = 0 because ivis /= 2), overridden for rh..en by HLLC or Roe.  For Roe,
   points with kroe = 0 use the local Lax-Friedrichs flux of the legacy
   flux_cal (dg_3D_hydro_lqc_SEAN.f90:1255-1270) for rh..en. */
int orc_riemann_face(const orc_ctx *c, const double *qlr, const double *flr, double *fout, int dir,
                     int kind) {
    const int nf = c->nface, nQ = c->nQ;
    /* cfr: local Lax-Friedrichs speeds, set only when ivis == 2
       (src/flux.f90:243-254); otherwise the module array is zero */
    double cfr[NFMAX];
    for (int p = 0; p < nf; p++)
        cfr[p] = c->c.ivis == 2 ? fmax(orc_cfcal(c, qlr + (long)p * nQ, dir), orc_cfcal(c, qlr + (long)(p + nf) * nQ, dir))
                               : 0.0;
    for (int e = 0; e < nQ; e++)
        for (int p = 0; p < nf; p++) {
            int pp = p + nf;
            fout[(long)p * nQ + e] = 0.5 * (flr[(long)p * nQ + e] + flr[(long)pp * nQ + e]) -
                                     0.5 * cfr[p] * (qlr[(long)pp * nQ + e] - qlr[(long)p * nQ + e]);
        }
    if (kind == ORC_FLUX_HLLC) {
        double fh[NFMAX * 5];
        for (int e = 0; e < 5; e++)
            for (int p = 0; p < nf; p++) fh[p * 5 + e] = fout[(long)p * nQ + e];
        hllc_face(c, qlr, flr, fh, dir);
        for (int e = 0; e < 5; e++)
            for (int p = 0; p < nf; p++) fout[(long)p * nQ + e] = fh[p * 5 + e];
    } else if (kind == ORC_FLUX_ROE) {
        double fr[NFMAX * 5];
        int kroe[NFMAX];
        roe_face(c, qlr, flr, fr, dir, kroe);
        for (int p = 0; p < nf; p++) {
            if (kroe[p] > 0) {
                for (int e = 0; e < 5; e++) fout[(long)p * nQ + e] = fr[p * 5 + e];
            } else {
                double cw = fmax(orc_cfcal(c, qlr + (long)p * nQ, dir), orc_cfcal(c, qlr + (long)(p + nf) * nQ, dir));
                for (int e = 0; e < 5; e++)
                    fout[(long)p * nQ + e] = 0.5 * (flr[(long)p * nQ + e] + flr[(long)(p + nf) * nQ + e]) -
                                             0.5 * cw * (qlr[(long)(p + nf) * nQ + e] - qlr[(long)p * nQ + e]);
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* limiter (src/flux.f90:1077-1164)                                          */
/* ------------------------------------------------------------------------ */
int orc_limiter(orc_ctx *c, double *q) {
    const orc_config *g = &c->c;
    const int nb = c->nb, npge = c->npge;
    const double epsi = g->rh_floor, epsiP = g->rh_floor * g->T_floor, aindm1 = c->aindm1;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < g->nz; k++)
        for (int j = 0; j < g->ny; j++)
            for (int i = 0; i < g->nx; i++) {
#define QQ(f, r) q[qidx(c, i, j, k, (f), (r))]
                if (QQ(RH, 0) < g->rh_floor) {
                    for (int r = 1; r < nb; r++)
                        for (int f = RH; f <= EN; f++) QQ(f, r) = 0.0;
                    QQ(RH, 0) = g->rh_floor;
                } else {
                    double Qrhmin = 0;
                    for (int p = 0; p < npge; p++) {
                        double s = 0;
                        for (int r = 0; r < nb; r++) s += c->bff[p][r] * QQ(RH, r);
                        if (p == 0 || s < Qrhmin) Qrhmin = s;
                    }
                    if (Qrhmin < epsi) {
                        double theta = (epsi - QQ(RH, 0)) / (Qrhmin - QQ(RH, 0));
                        if (theta > 1.) theta = 1.;
                        if (theta < 0) theta = 0.;
                        for (int r = 1; r < nb; r++) QQ(RH, r) = fabs(theta) * QQ(RH, r);
                    }
                    double Pave = aindm1 * (QQ(EN, 0) - 0.5 * (QQ(MX, 0) * QQ(MX, 0) + QQ(MY, 0) * QQ(MY, 0) +
                                                               QQ(MZ, 0) * QQ(MZ, 0)) / QQ(RH, 0));
                    if (Pave < epsiP) {
                        for (int r = 1; r < nb; r++)
                            for (int f = RH; f <= EN; f++) QQ(f, r) = 0.0;
                    } else {
                        double theta = 1.;
                        for (int p = 0; p < npge; p++) {
                            double e[5];
                            for (int f = RH; f <= EN; f++) {
                                double s = 0;
                                for (int r = 0; r < nb; r++) s += c->bff[p][r] * QQ(f, r);
                                e[f] = s;
                            }
                            double dn = e[RH], dni = 1. / dn;
                            double P = aindm1 * (e[EN] - 0.5 * dni * (e[MX] * e[MX] + e[MY] * e[MY] + e[MZ] * e[MZ]));
                            if (P < epsiP) {
                                if (Pave != P) {
                                    double thetaj = (Pave - epsiP) / (Pave - P);
                                    theta = fmin(theta, thetaj);
                                }
                            }
                        }
                        if (theta > 1.) theta = 1.;
                        if (theta < 0.) theta = 0.;
                        for (int r = 1; r < nb; r++)
                            for (int f = RH; f <= EN; f++) QQ(f, r) = theta * QQ(f, r);
                    }
                }
#undef QQ
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* output_vtk (src/output.f90:14-318) with set_vtk_vals_3D                   */
/* (src/basis_funcs.f90:90-200)                                              */
/* ------------------------------------------------------------------------ */
int orc_sample(const orc_ctx *c, const double *q, int nvtk, uint32_t fields, const double *scales, float *out) {
    const orc_config *g = &c->c;
    const int nb = c->nb, nQ = c->nQ, nv3 = nvtk * nvtk * nvtk;
    if (nvtk < 1 || nvtk > 4) return set_err(-1, "nvtk must be 1..4");
    double xg[4];
    for (int i = 0; i < nvtk; i++) xg[i] = (double)(2 * i + 1 - nvtk) / nvtk;
    /* bfvtk, bfvtk_dx, bfvtk_dy: sub-point igrid = nvtk^2*(ib) + nvtk*(jb) + kb */
    double bf[64][NBMAX], bdx[64][NBMAX], bdy[64][NBMAX];
    for (int ib = 0; ib < nvtk; ib++)
        for (int jb = 0; jb < nvtk; jb++)
            for (int kb = 0; kb < nvtk; kb++) {
                int ig = nvtk * nvtk * ib + nvtk * jb + kb;
                double *v = bf[ig];
                double x = xg[ib], y = xg[jb], z = xg[kb];
                for (int r = 0; r < NBMAX; r++) v[r] = bdx[ig][r] = bdy[ig][r] = 0.;
                v[K1] = 1.; v[KX] = x; v[KY] = y; v[KZ] = z;
                fill_products(v, nb);
                bdx[ig][KX] = 1.; bdy[ig][KY] = 1.;
                if (nb > 4) {
                    bdx[ig][KZX] = z; bdx[ig][KXY] = y;
                    bdy[ig][KYZ] = z; bdy[ig][KXY] = x;
                }
                if (nb == 8) { bdx[ig][KXYZ] = y * z; bdy[ig][KXYZ] = z * x; }
                if (nb == 10) { bdx[ig][KXX] = 3. * x; bdy[ig][KYY] = 3. * y; }
            }
    const int nnx = g->nx * nvtk, nny = g->ny * nvtk, nnz = g->nz * nvtk;
    const long npts = (long)nnx * nny * nnz;
    int sel[14], nsel = 0;
    for (int f = 0; f < 14; f++)
        if (fields & (1u << f)) sel[nsel++] = f;
    const double gm1 = c->aindex - 1.;
    for (int k = 0; k < nnz; k++)
        for (int j = 0; j < nny; j++)
            for (int i = 0; i < nnx; i++) {
                int ir = i / nvtk, jr = j / nvtk, kr = k / nvtk;
                int ig = nvtk * nvtk * (i % nvtk) + nvtk * (j % nvtk) + (k % nvtk);
                double v[NQMAX];
                for (int f = 0; f < nQ; f++) {
                    double s = 0.;
                    for (int r = 0; r < nb; r++) s += bf[ig][r] * q[qidx(c, ir, jr, kr, f, r)];
                    v[f] = s;
                }
                double dxrh = 0., dyrh = 0., dxmy = 0., dymx = 0.;
                for (int r = 0; r < nb; r++) {
                    dxrh += bdx[ig][r] * q[qidx(c, ir, jr, kr, RH, r)];
                    dyrh += bdy[ig][r] * q[qidx(c, ir, jr, kr, RH, r)];
                    dxmy += bdx[ig][r] * q[qidx(c, ir, jr, kr, MY, r)];
                    dymx += bdy[ig][r] * q[qidx(c, ir, jr, kr, MX, r)];
                }
                double dxvy = (v[RH] * dxmy - v[MY] * dxrh) / (v[RH] * v[RH]);
                double dyvx = (v[RH] * dymx - v[MX] * dyrh) / (v[RH] * v[RH]);
                double dn = v[RH], dni = 1. / dn;
                double vx = v[MX] * dni, vy = v[MY] * dni, vz = v[MZ] * dni;
                double U = v[EN] - 0.5 * dn * (vx * vx + vy * vy + vz * vz);
                double P = gm1 * U;
                long id = i + (long)nnx * (j + (long)nny * k);
                for (int s = 0; s < nsel; s++) {
                    int f = sel[s];
                    double sc = scales ? scales[f] : 1.0, val = 0.;
                    switch (f) {
                    case 0: val = dn * sc; break;
                    case 1: val = log(dn * sc) / log(10.); break;
                    case 2: val = sc * v[MX] / v[RH]; break;
                    case 3: val = sc * v[MY] / v[RH]; break;
                    case 4: val = sc * v[MZ] / v[RH]; break;
                    case 5: val = P * sc * dni; break;
                    case 6: val = U * sc; break;
                    case 7: val = v[EN] * sc; break;
                    case 8: val = P * pow(dni, gm1) * sc; break;
                    case 9: val = P * sc; break;
                    case 10: val = nQ > PXX ? v[PXX] * sc : 0.; break;
                    case 11: val = nQ > PYY ? v[PYY] * sc : 0.; break;
                    case 12:
                        /* the 13-moment model stores the traceless stress without pzz */
                        val = nQ == 13 ? -(v[T_PXX] + v[T_PYY]) * sc : (nQ > PZZ ? v[PZZ] * sc : 0.);
                        break;
                    case 13: val = -2. * (dxvy - dyvx) * c->dxi * sc; break;
                    }
                    out[(long)s * npts + id] = (float)val;
                }
            }
    (void)nv3;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* exchange_flux + apply_boundaries (src/prepare_step.f90:20-179,            */
/* src/boundary.f90:40-338,748-782), one device, periodic wrap local         */
/* ------------------------------------------------------------------------ */
static double trace_sum(const orc_ctx *c, const double *q, const double *bf, int i, int j, int k, int f) {
    double s = 0;
    for (int r = 0; r < c->nb; r++) s += bf[r] * q[qidx(c, i, j, k, f, r)];
    return s;
}

/* Q*_int / Q*_ext element (a,b,p,f) with a,b the two tangential cell indices */
#define BIDX(na, nb_, a, b, p, f) ((long)(a) + (long)(na) * ((long)(b) + (long)(nb_) * ((long)(p) + (long)nface * (f))))

static void exchange_and_bc(orc_ctx *c, const double *q) {
    const orc_config *g = &c->c;
    const int nx = g->nx, ny = g->ny, nz = g->nz, nface = c->nface, nQ = c->nQ;
    for (int f = 0; f < nQ; f++)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int p = 0; p < nface; p++) {
                    c->intr[0][BIDX(ny, nz, j, k, p, f)] = trace_sum(c, q, c->bxm[p], 0, j, k, f);
                    c->intr[1][BIDX(ny, nz, j, k, p, f)] = trace_sum(c, q, c->bxp[p], nx - 1, j, k, f);
                }
    for (int f = 0; f < nQ; f++)
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++)
                for (int p = 0; p < nface; p++) {
                    c->intr[2][BIDX(nx, nz, i, k, p, f)] = trace_sum(c, q, c->bym[p], i, 0, k, f);
                    c->intr[3][BIDX(nx, nz, i, k, p, f)] = trace_sum(c, q, c->byp[p], i, ny - 1, k, f);
                }
    for (int f = 0; f < nQ; f++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                for (int p = 0; p < nface; p++) {
                    c->intr[4][BIDX(nx, ny, i, j, p, f)] = trace_sum(c, q, c->bzm[p], i, j, 0, f);
                    c->intr[5][BIDX(nx, ny, i, j, p, f)] = trace_sum(c, q, c->bzp[p], i, j, nz - 1, f);
                }
    int na[6] = {ny, ny, nx, nx, nx, nx}, nbb[6] = {nz, nz, nz, nz, ny, ny};
    for (int s = 0; s < 6; s++) {
        long n = (long)na[s] * nbb[s] * nface * nQ;
        int bc = g->bc[s];
        int normal = MX + s / 2; /* mx, my, mz */
        if (bc == ORC_BC_PERIODIC) {
            /* periodic neighbour is self: lo ext <- hi int, hi ext <- lo int */
            memcpy(c->ext[s], c->intr[s ^ 1], sizeof(double) * n);
            continue;
        }
        memcpy(c->ext[s], c->intr[s], sizeof(double) * n);
        for (int b = 0; b < nbb[s]; b++)
            for (int a = 0; a < na[s]; a++) {
#define EX(p, f) c->ext[s][BIDX(na[s], nbb[s], a, b, p, f)]
#define IN(p, f) c->intr[s][BIDX(na[s], nbb[s], a, b, p, f)]
                if (bc == ORC_BC_INFLOW) {
                    /* x_bc_inflow (src/boundary.f90:571-592): interior copy, then
                       rh, mx, en from the prescribed state and no transverse momentum */
                    for (int p = 0; p < nface; p++) {
                        EX(p, RH) = c->inflow[BIDX(na[s], nbb[s], a, b, p, RH)];
                        EX(p, MX) = c->inflow[BIDX(na[s], nbb[s], a, b, p, MX)];
                        EX(p, MY) = 0.0;
                        EX(p, MZ) = 0.0;
                        EX(p, EN) = c->inflow[BIDX(na[s], nbb[s], a, b, p, EN)];
                    }
                } else if (bc == ORC_BC_NOSLIP) {
                    for (int p = 0; p < nface; p++) EX(p, MX) = EX(p, MY) = EX(p, MZ) = 0.0;
                } else if (bc == ORC_BC_WALL) {
                    for (int p = 0; p < nface; p++) EX(p, normal) = 0.0;
                } else if (bc == ORC_BC_OUTFLOW) {
                    if (s == 1) {
                        /* xhibc_outflow (src/boundary.f90:246-282) after the copy */
                        for (int p = 0; p < nface; p++)
                            if (EX(p, MX) > 0.0) EX(p, MX) = -IN(p, MX);
                    } else if ((s & 1) == 0) {
                        double mxv = EX(0, normal);
                        for (int p = 1; p < nface; p++) mxv = fmax(mxv, EX(p, normal));
                        if (mxv > 0.0)
                            for (int p = 0; p < nface; p++) EX(p, normal) = 0.0;
                    } else {
                        double mnv = EX(0, normal);
                        for (int p = 1; p < nface; p++) mnv = fmin(mnv, EX(p, normal));
                        if (mnv < 0.0)
                            for (int p = 0; p < nface; p++) EX(p, normal) = 0.0;
                    }
                }
#undef EX
#undef IN
            }
    }
}

/* ------------------------------------------------------------------------ */
/* stochastic stress at points (src/random.f90:106-145 intent, FDT-scaled)   */
/* ------------------------------------------------------------------------ */
/* Symmetric stochastic stress with the Landau-Lifshitz covariance
   <S_ij S_kl> = ae^2 (d_ik d_jl + d_il d_jk - 2/3 d_ij d_kl) + 3 az^2 d_ij d_kl.
   random_stresses_pnts_r symmetrises a 3x3 Gaussian matrix W: off-diagonals
   (W_ij+W_ji)/sqrt2 are N(0,1), so one normal each is equal in law; the
   diagonal here is sqrt2*W_ii - trace/3 (variance 4/3, the deviatoric value
   the covariance above needs; the reference uses W_ii - trace(W)/3), plus the
   bulk part az*tr(W).  Six normals per point from two Philox blocks:
   g0..g2 diagonal, g3..g5 the xy, xz, yz off-diagonals.
   stream 1 = x faces, 2 = y faces, 3 = z faces, 4 = volume points;
   id = point*N + linear cell/face id, counters 2*id and 2*id+1. */
static void noise_stress(const orc_ctx *c, uint32_t stream, uint64_t id, double amp_eta, double amp_zeta,
                         double S[6]) {
    if (c->imp_on) {
        for (int e = 0; e < 6; e++) S[e] = 0.0;
        if (stream == c->imp_stream && id == c->imp_id) S[c->imp_comp] = c->imp_val;
        return;
    }
    double g[8];
    orc_noise_normals(c->c.seed, c->c.realization, c->step_index, stream, 2 * id, g);
    orc_noise_normals(c->c.seed, c->c.realization, c->step_index, stream, 2 * id + 1, g + 4);
    const double s2 = 1.4142135623730951;
    double dxx = s2 * g[0], dyy = s2 * g[1], dzz = s2 * g[2];
    double tr = (dxx + dyy + dzz) / 3.0;
    double trW = g[0] + g[1] + g[2];
    S[0] = amp_eta * (dxx - tr) + amp_zeta * trW;
    S[1] = amp_eta * (dyy - tr) + amp_zeta * trW;
    S[2] = amp_eta * (dzz - tr) + amp_zeta * trW;
    S[3] = amp_eta * g[3];
    S[4] = amp_eta * g[4];
    S[5] = amp_eta * g[5];
    /* the field is stored in single precision, the reference's precision
       for its Gaussian draws (vsRngGaussian, src/random.f90:61,87) */
    for (int e = 0; e < 6; e++) S[e] = (double)(float)S[e];
}

int orc_noise_stress(const orc_ctx *c, uint32_t stream, uint64_t id, double amp_eta, double amp_zeta, double S[6]) {
    noise_stress(c, stream, id, amp_eta, amp_zeta, S);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* calc_flux_x/y/z (src/flux.f90:197-450)                                    */
/* ------------------------------------------------------------------------ */
static void calc_face_fluxes(orc_ctx *c, const double *q, double noise_amp_eta, double noise_amp_zeta) {
    const orc_config *g = &c->c;
    const int nx = g->nx, ny = g->ny, nz = g->nz, nface = c->nface, nQ = c->nQ, nfe = c->nfe;
    double Qf[2 * NFMAX * NQMAX], Ff[2 * NFMAX * NQMAX], Fo[NFMAX * NQMAX], S[2 * NFMAX * 6];
    for (int dir = 1; dir <= 3; dir++) {
        int ni = nx + (dir == 1), nj = ny + (dir == 2), nk = nz + (dir == 3);
        double *out = dir == 1 ? c->fx : dir == 2 ? c->fy : c->fz;
#pragma omp parallel for schedule(static) private(Qf, Ff, Fo, S)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                for (int i = 0; i < ni; i++) {
                    int lo_edge = (dir == 1) ? (i == 0) : (dir == 2) ? (j == 0) : (k == 0);
                    int hi_edge = (dir == 1) ? (i == nx) : (dir == 2) ? (j == ny) : (k == nz);
                    double(*bp)[NBMAX] = dir == 1 ? c->bxp : dir == 2 ? c->byp : c->bzp;
                    double(*bm)[NBMAX] = dir == 1 ? c->bxm : dir == 2 ? c->bym : c->bzm;
                    int il = i - (dir == 1), jl = j - (dir == 2), kl = k - (dir == 3);
                    for (int f = 0; f < nQ; f++)
                        for (int p = 0; p < nface; p++) {
                            double vl, vr;
                            if (!lo_edge) vl = trace_sum(c, q, bp[p], il, jl, kl, f);
                            else if (dir == 1) vl = c->ext[0][BIDX(ny, nz, j, k, p, f)];
                            else if (dir == 2) vl = c->ext[2][BIDX(nx, nz, i, k, p, f)];
                            else vl = c->ext[4][BIDX(nx, ny, i, j, p, f)];
                            if (!hi_edge) vr = trace_sum(c, q, bm[p], i, j, k, f);
                            else if (dir == 1) vr = c->ext[1][BIDX(ny, nz, j, k, p, f)];
                            else if (dir == 2) vr = c->ext[3][BIDX(nx, nz, i, k, p, f)];
                            else vr = c->ext[5][BIDX(nx, ny, i, j, p, f)];
                            Qf[p * nQ + f] = vl;
                            Qf[(p + nface) * nQ + f] = vr;
                        }
                    const double *Sp = NULL;
                    if (g->noise) {
                        /* one draw per face point, shared by both sides; the
                           periodic wrap face uses the id of the low face */
                        int fi = (dir == 1 && hi_edge && g->bc[0] == ORC_BC_PERIODIC) ? 0 : i;
                        int fj = (dir == 2 && hi_edge && g->bc[2] == ORC_BC_PERIODIC) ? 0 : j;
                        int fk = (dir == 3 && hi_edge && g->bc[4] == ORC_BC_PERIODIC) ? 0 : k;
                        uint64_t fid = (uint64_t)fi + (uint64_t)(nx + 1) * ((uint64_t)fj + (uint64_t)(ny + 1) * fk);
                        uint64_t nfaces = (uint64_t)(nx + 1) * (ny + 1) * (nz + 1);
                        for (int p = 0; p < nface; p++) {
                            noise_stress(c, (uint32_t)dir, (uint64_t)p * nfaces + fid, noise_amp_eta, noise_amp_zeta, S + 6 * p);
                            for (int e = 0; e < 6; e++) S[6 * (p + nface) + e] = S[6 * p + e];
                        }
                        Sp = S;
                    }
                    flux_points_s(c, Qf, Ff, dir, nfe, Sp);
                    orc_riemann_face(c, Qf, Ff, Fo, dir, g->riemann);
                    for (int f = 0; f < nQ; f++)
                        for (int p = 0; p < nface; p++) {
                            long o = (long)p + (long)nface * ((long)i + (long)ni * ((long)j + (long)nj * ((long)k + (long)nk * f)));
                            out[o] = Fo[p * nQ + f];
                        }
                }
    }
}

static inline long fxi(const orc_ctx *c, int p, int i, int j, int k, int f, int dir) {
    const orc_config *g = &c->c;
    int ni = g->nx + (dir == 1), nj = g->ny + (dir == 2), nk = g->nz + (dir == 3);
    return (long)p + (long)c->nface * ((long)i + (long)ni * ((long)j + (long)nj * ((long)k + (long)nk * f)));
}

/* innerintegral2 (src/flux.f90:751-928), nbasis in {4,8,10} */
static void inner_integral(orc_ctx *c, const double *q, double noise_amp_eta, double noise_amp_zeta) {
    const orc_config *g = &c->c;
    const int nb = c->nb, npg = c->npg, nQ = c->nQ;
    double Qin[NPGMAX * NQMAX], fx[NPGMAX * NQMAX], fy[NPGMAX * NQMAX], fz[NPGMAX * NQMAX], S[NPGMAX * 6];
    const double dxi = c->dxi, dyi = c->dyi, dzi = c->dzi;
    const double *w = c->w3;
#pragma omp parallel for schedule(static) private(Qin, fx, fy, fz, S)
    for (int k = 0; k < g->nz; k++)
        for (int j = 0; j < g->ny; j++)
            for (int i = 0; i < g->nx; i++) {
                for (int f = 0; f < nQ; f++)
                    for (int p = 0; p < npg; p++) {
                        double s = 0;
                        for (int r = 0; r < nb; r++) s += c->bint[p][r] * q[qidx(c, i, j, k, f, r)];
                        Qin[p * nQ + f] = s;
                    }
                const double *Sp = NULL;
                if (g->noise) {
                    uint64_t cid = (uint64_t)i + (uint64_t)g->nx * ((uint64_t)j + (uint64_t)g->ny * k);
                    for (int p = 0; p < npg; p++)
                        noise_stress(c, 4u, (uint64_t)p * (uint64_t)c->ncell + cid, noise_amp_eta, noise_amp_zeta, S + 6 * p);
                    Sp = S;
                }
                flux_points_s(c, Qin, fx, 1, npg, Sp);
                flux_points_s(c, Qin, fy, 2, npg, Sp);
                flux_points_s(c, Qin, fz, 3, npg, Sp);
                double int_r[NBMAX];
                for (int f = 0; f < nQ; f++) {
                    for (int r = 0; r < nb; r++) int_r[r] = 0;
                    double s1 = 0, s2 = 0, s3 = 0;
                    for (int p = 0; p < npg; p++) {
                        s1 = s1 + w[p] * fx[p * nQ + f];
                        s2 = s2 + w[p] * fy[p * nQ + f];
                        s3 = s3 + w[p] * fz[p * nQ + f];
                    }
                    int_r[KX] = 0.25 * c->cb[KX] * dxi * s1;
                    int_r[KY] = 0.25 * c->cb[KY] * dyi * s2;
                    int_r[KZ] = 0.25 * c->cb[KZ] * dzi * s3;
                    if (nb > 4) {
                        /* ibitri == 1 .or. iquad > 2 */
                        double t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0;
                        for (int p = 0; p < npg; p++) {
                            t1 = t1 + w[p] * c->bint[p][KZ] * fy[p * nQ + f];
                            t2 = t2 + w[p] * c->bint[p][KY] * fz[p * nQ + f];
                            t3 = t3 + w[p] * c->bint[p][KZ] * fx[p * nQ + f];
                            t4 = t4 + w[p] * c->bint[p][KX] * fz[p * nQ + f];
                            t5 = t5 + w[p] * c->bint[p][KY] * fx[p * nQ + f];
                            t6 = t6 + w[p] * c->bint[p][KX] * fy[p * nQ + f];
                        }
                        int_r[KYZ] = 0.25 * c->cb[KYZ] * (dyi * t1 + dzi * t2);
                        int_r[KZX] = 0.25 * c->cb[KZX] * (dxi * t3 + dzi * t4);
                        int_r[KXY] = 0.25 * c->cb[KXY] * (dxi * t5 + dyi * t6);
                        if (c->q > 2) {
                            double u1 = 0, u2 = 0, u3 = 0;
                            for (int p = 0; p < npg; p++) {
                                u1 = u1 + w[p] * 3. * c->bint[p][KX] * fx[p * nQ + f];
                                u2 = u2 + w[p] * 3. * c->bint[p][KY] * fy[p * nQ + f];
                                u3 = u3 + w[p] * 3. * c->bint[p][KZ] * fz[p * nQ + f];
                            }
                            int_r[KXX] = 0.25 * c->cb[KXX] * (dxi * u1);
                            int_r[KYY] = 0.25 * c->cb[KYY] * (dyi * u2);
                            int_r[KZZ] = 0.25 * c->cb[KZZ] * (dzi * u3);
                        }
                        if (c->ibitri == 1) {
                            double u7 = 0, u8 = 0, u9 = 0;
                            for (int p = 0; p < npg; p++) {
                                u7 = u7 + w[p] * c->bint[p][KYZ] * fx[p * nQ + f];
                                u8 = u8 + w[p] * c->bint[p][KZX] * fy[p * nQ + f];
                                u9 = u9 + w[p] * c->bint[p][KXY] * fz[p * nQ + f];
                            }
                            int_r[KXYZ] = 0.25 * c->cb[KXYZ] * (dxi * u7 + dyi * u8 + dzi * u9);
                        }
                    }
                    for (int r = 0; r < nb; r++) c->integ[qidx(c, i, j, k, f, r)] = int_r[r];
                }
            }
}

/* Point amplitudes of the stochastic stress.  sqrt(dVi/dt) is sqrt_dVdt_i of
   src/old_mod/perseushd_main.f90:128-131.  A Gauss-weighted average of
   independent point values has variance var*sum_p (w_p/W)^2, so the point
   amplitude carries kappa = 1/sqrt(sum_p (w_p/W)^2) for the face- (or
   cell-) averaged stress to have the Landau-Lifshitz variance
   eta_sd^2/(dV dt); face points use the 2D weights, volume points the 3D. */
static void noise_amps(const orc_ctx *c, double dt, int vol, double *ae, double *az) {
    *ae = 0; *az = 0;
    if (!c->c.noise) return;
    double dVi = c->dxi * c->dyi * c->dzi;
    double sq = sqrt(dVi / dt);
    int n = vol ? c->npg : c->nface;
    const double *w = vol ? c->w3 : c->w2;
    double W = 0, s2 = 0;
    for (int p = 0; p < n; p++) W += w[p];
    for (int p = 0; p < n; p++) s2 += (w[p] / W) * (w[p] / W);
    double kappa = 1.0 / sqrt(s2);
    *ae = c->c.eta_sd * sq * kappa;
    *az = c->c.zeta_sd * sq * kappa;
}

/* glflux (src/flux.f90:932-1022) */
int orc_glflux(orc_ctx *c, const double *q, double *glflux) {
    const orc_config *g = &c->c;
    const int nb = c->nb, nface = c->nface, nQ = c->nQ;
    double ae, az, aev, azv;
    noise_amps(c, c->noise_dt, 0, &ae, &az);
    noise_amps(c, c->noise_dt, 1, &aev, &azv);
    if (g->bc[0] == ORC_BC_INFLOW && !c->inflow) return set_err(-1, "inflow boundary without its state (orc_set_inflow)");
    exchange_and_bc(c, q);
    calc_face_fluxes(c, q, ae, az);
    inner_integral(c, q, aev, azv);
    const double *w2 = c->w2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int f = 0; f < nQ; f++)
        for (int k = 0; k < g->nz; k++)
            for (int j = 0; j < g->ny; j++)
                for (int i = 0; i < g->nx; i++) {
                    double sx = 0, sy = 0, sz = 0;
                    for (int p = 0; p < nface; p++) sx = sx + w2[p] * (c->fx[fxi(c, p, i + 1, j, k, f, 1)] - c->fx[fxi(c, p, i, j, k, f, 1)]);
                    for (int p = 0; p < nface; p++) sy = sy + w2[p] * (c->fy[fxi(c, p, i, j + 1, k, f, 2)] - c->fy[fxi(c, p, i, j, k, f, 2)]);
                    for (int p = 0; p < nface; p++) sz = sz + w2[p] * (c->fz[fxi(c, p, i, j, k + 1, f, 3)] - c->fz[fxi(c, p, i, j, k, f, 3)]);
                    glflux[qidx(c, i, j, k, f, 0)] = 0.25 * (c->dxi * sx + c->dyi * sy + c->dzi * sz);
                    for (int r = 1; r < nb; r++) {
                        sx = 0; sy = 0; sz = 0;
                        for (int p = 0; p < nface; p++)
                            sx = sx + w2[p] * (c->bxp[p][r] * c->fx[fxi(c, p, i + 1, j, k, f, 1)] - c->bxm[p][r] * c->fx[fxi(c, p, i, j, k, f, 1)]);
                        for (int p = 0; p < nface; p++)
                            sy = sy + w2[p] * (c->byp[p][r] * c->fy[fxi(c, p, i, j + 1, k, f, 2)] - c->bym[p][r] * c->fy[fxi(c, p, i, j, k, f, 2)]);
                        for (int p = 0; p < nface; p++)
                            sz = sz + w2[p] * (c->bzp[p][r] * c->fz[fxi(c, p, i, j, k + 1, f, 3)] - c->bzm[p][r] * c->fz[fxi(c, p, i, j, k, f, 3)]);
                        glflux[qidx(c, i, j, k, f, r)] = 0.25 * c->cb[r] * (c->dxi * sx + c->dyi * sy + c->dzi * sz) -
                                                         c->integ[qidx(c, i, j, k, f, r)];
                    }
                }
    return 0;
}

/* source_calc (src/sources.f90:12-84) plus the body force */
int orc_source(orc_ctx *c, const double *q, double *source) {
    const orc_config *g = &c->c;
    const int nb = c->nb, npg = c->npg, nQ = c->nQ;
    double src[NPGMAX * NQMAX];
#pragma omp parallel for schedule(static) private(src)
    for (int k = 0; k < g->nz; k++)
        for (int j = 0; j < g->ny; j++)
            for (int i = 0; i < g->nx; i++) {
                for (int p = 0; p < npg; p++) {
                    double Qin[NQMAX];
                    for (int f = 0; f < nQ; f++) {
                        double s = 0;
                        for (int r = 0; r < nb; r++) s += c->bint[p][r] * q[qidx(c, i, j, k, f, r)];
                        Qin[f] = s;
                    }
                    for (int f = 0; f < nQ; f++) src[p * nQ + f] = 0;
                    if (nQ == 11 && g->ivis == 0)
                        for (int f = PXX; f <= PYZ; f++) src[p * nQ + f] = -g->coll * Qin[f];
                    if (nQ == 13 && g->ivis == 0) {
                        /* explicit relaxation of the stress and the heat flux */
                        for (int f = T_PXX; f <= T_PYZ; f++) src[p * nQ + f] = -g->coll * Qin[f];
                        for (int f = T_QX; f <= T_QZ; f++) src[p * nQ + f] = -g->coll_q * Qin[f];
                    }
                    if (nQ >= 11 && g->ivis == 2) {
                        /* explicit 10-moment relaxation source (src/sources.f90:55-73) */
                        double dn = Qin[RH], dni = 1. / dn;
                        double vx = Qin[MX] * dni, vy = Qin[MY] * dni, vz = Qin[MZ] * dni;
                        double vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
                        double colldn = g->coll * dn;
                        src[p * nQ + PXX] = colldn * (2 * vx2 - vy2 - vz2) * (1. / 3.);
                        src[p * nQ + PYY] = colldn * (2 * vy2 - vz2 - vx2) * (1. / 3.);
                        src[p * nQ + PZZ] = colldn * (2 * vz2 - vx2 - vy2) * (1. / 3.);
                        src[p * nQ + PXY] = colldn * vx * vy;
                        src[p * nQ + PXZ] = colldn * vx * vz;
                        src[p * nQ + PYZ] = colldn * vy * vz;
                    }
                }
                for (int r = 0; r < nb; r++)
                    for (int f = 0; f < nQ; f++) {
                        double s = 0;
                        for (int p = 0; p < npg; p++) s += c->bwint[p][r] * src[p * nQ + f];
                        double v = 0.125 * c->cb[r] * s;
                        /* body force: rho*g in momentum, m.g in energy, projected
                           exactly onto the modes (not in the reference) */
                        if (f >= MX && f <= MZ) v += g->body_force[f - MX] * q[qidx(c, i, j, k, RH, r)];
                        if (f == EN)
                            v += g->body_force[0] * q[qidx(c, i, j, k, MX, r)] + g->body_force[1] * q[qidx(c, i, j, k, MY, r)] +
                                 g->body_force[2] * q[qidx(c, i, j, k, MZ, r)];
                        source[qidx(c, i, j, k, f, r)] = v;
                    }
            }
    return 0;
}

int orc_rhs(orc_ctx *c, const double *q, double *rhs) {
    long n = c->ncell * c->nQ * c->nb;
    int rc = orc_glflux(c, q, c->glf);
    if (rc) return rc;
    orc_source(c, q, c->src);
    for (long e = 0; e < n; e++) rhs[e] = -(c->glf[e] - c->src[e]);
    return 0;
}

/* euler_step (src/integrator.f90:93-122) + advance_time_level (125-252) */
int orc_euler_step(orc_ctx *c, double *q_in, double *q_out, double dt) {
    const orc_config *g = &c->c;
    const int nb = c->nb, nQ = c->nQ;
    if (g->limiter) orc_limiter(c, q_in);
    const double keep_dt = c->noise_dt; /* the stage's dt sets the noise amplitude */
    c->noise_dt = dt;
    int rc = orc_glflux(c, q_in, c->glf);
    c->noise_dt = keep_dt;
    if (rc) return rc;
    orc_source(c, q_in, c->src);
    double faci = 1. / (1. + dt * g->coll);
    double faciq = 1. / (1. + dt * g->coll_q);
    long n = c->ncell;
#pragma omp parallel for collapse(2) schedule(static)
    for (int r = 0; r < nb; r++)
        for (int f = 0; f < nQ; f++)
            for (long e = 0; e < n; e++) {
                long o = e + n * (f + (long)nQ * r);
                double v = q_in[o] - dt * (c->glf[o] - c->src[o]);
                /* 13-moment: the heat flux relaxes at its own rate */
                if (g->ivis == 1 && nQ == 13 && f >= T_QX) v = faciq * v;
                else if (g->ivis == 1 && f >= PXX) v = faci * v;
                q_out[o] = v;
            }
    if (g->ivis == 2) {
        /* backward-Euler relaxation of the second moments towards their
           trace (src/integrator.f90:181-228) */
        for (int r = 0; r < nb; r++)
            for (long e = 0; e < n; e++) {
#define O(f) (e + n * ((f) + (long)nQ * r))
                double Q_sum = g->coll * dt * (q_in[O(PXX)] + q_in[O(PYY)] + q_in[O(PZZ)]) * (1. / 3.);
                double glf_sum = g->coll * dt * dt * (c->glf[O(PXX)] + c->glf[O(PYY)] + c->glf[O(PZZ)]) * (1. / 3.);
                double src_sum = g->coll * dt * dt * (c->src[O(PXX)] + c->src[O(PYY)] + c->src[O(PZZ)]) * (1. / 3.);
                for (int f = PXX; f <= PZZ; f++)
                    q_out[O(f)] = faci * (q_in[O(f)] - dt * c->glf[O(f)] + dt * c->src[O(f)] + Q_sum - glf_sum + src_sum);
                for (int f = PXY; f <= PYZ; f++)
                    q_out[O(f)] = (q_in[O(f)] - dt * c->glf[O(f)] + dt * c->src[O(f)]) * faci;
#undef O
            }
    }
    for (int f = 0; f < nQ; f++)
        for (long e = 0; e < n; e++) {
            double v = q_out[e + n * f];
            if (v != v) {
                g_bad_cell = e;
                return set_err(-3, "NaN in mode 1 of the stage output");
            }
        }
    return 0;
}

/* get_min_dt (src/hermeshd.f90:145-198), one rank */
double orc_min_dt(orc_ctx *c, const double *q) {
    const orc_config *g = &c->c;
    double vmag = 0., cs = 0.;
    for (int k = 0; k < g->nz; k++)
        for (int j = 0; j < g->ny; j++)
            for (int i = 0; i < g->nx; i++) {
                double dn = q[qidx(c, i, j, k, RH, 0)];
                double dni = 1. / dn;
                double vx = q[qidx(c, i, j, k, MX, 0)] * dni;
                double vy = q[qidx(c, i, j, k, MY, 0)] * dni;
                double vz = q[qidx(c, i, j, k, MZ, 0)] * dni;
                if (g->ieos == 2)
                    cs = sqrt(7.2 * g->eos_p1 * pow(dn, 6.2) + g->T_floor);
                else
                    cs = sqrt(c->aindex * (q[qidx(c, i, j, k, EN, 0)] * dni - 0.5 * (vx * vx + vy * vy + vz * vz)));
                double vmag0 = fmax(fmax(fabs(vx) + cs, fabs(vy) + cs), fabs(vz) + cs);
                if (vmag0 > vmag && dn > 1.01 * g->rh_floor) vmag = vmag0;
            }
    double vmax = (vmag + cs) * c->dxi;
    return g->cflm / vmax;
}

/* RK2 / RK3 (src/integrator.f90:62-86) */
int orc_update(orc_ctx *c, double *q_io, double *q1, double *q2, double dt) {
    long n = c->ncell * c->nQ * c->nb;
    int rc;
    if (c->c.integrator == 3) {
        if ((rc = orc_euler_step(c, q_io, q1, dt))) return rc;
        if ((rc = orc_euler_step(c, q1, q2, dt))) return rc;
        for (long e = 0; e < n; e++) q1[e] = 0.75 * q_io[e] + 0.25 * q2[e];
        if ((rc = orc_euler_step(c, q1, q2, dt))) return rc;
        for (long e = 0; e < n; e++) q_io[e] = (1. / 3.) * q_io[e] + (2. / 3.) * q2[e];
    } else {
        if ((rc = orc_euler_step(c, q_io, q1, dt))) return rc;
        if ((rc = orc_euler_step(c, q1, q2, dt))) return rc;
        for (long e = 0; e < n; e++) q_io[e] = 0.5 * (q_io[e] + q2[e]);
    }
    orc_apply_mask(c, q_io);
    c->step_index++; /* the noise of the next step is a fresh draw */
    return 0;
}

/* step (src/hermeshd.f90:53-62) */
int orc_step(orc_ctx *c, double *q_io, double *q1, double *q2, double *t, double *dt) {
    *dt = orc_min_dt(c, q_io);
    int rc = orc_update(c, q_io, q1, q2, *dt);
    *t = *t + *dt;
    return rc;
}
