// hd_tables.h -- Legendre-product basis and Gauss tables for one cell.
//
// Mode order (src/initialize.f90:126-194): 1, x, y, z, yz, zx, xy, then
// xyz (nbasis 8) or P2(x), P2(y), P2(z) (nbasis 10).  Point order:
//   interior (src/basis_funcs.f90:265-343): x slowest, z fastest;
//   x faces: y fastest then z;  y faces: z fastest then x;  z faces: x fastest then y
//   (src/basis_funcs.f90:346-567).
// Minus-face tables flip modes odd in the normal coordinate.  Weights follow
// set_weights_3D (src/basis_funcs.f90:570-627); cbasis set_cbasis_3D
// (src/initialize.f90:199-220).  Values are computed with the same double
// operations as the reference (e.g. 1/sqrt(3), sqrt(15)/5, 1.5*s*s-0.5).
#pragma once
#include <cmath>
#include <cstring>

namespace hd {

enum : int { K1 = 0, KX = 1, KY = 2, KZ = 3, KYZ = 4, KZX = 5, KXY = 6, KXYZ = 7, KXX = 7, KYY = 8, KZZ = 9 };
enum : int { F_XM = 0, F_XP = 1, F_YM = 2, F_YP = 3, F_ZM = 4, F_ZP = 5 };

struct Tables {
    int q, nb, nface, npg;
    double xq[3], w1[3], w2[9], w3[27], cb[10];
    double bint[27][10];
    double bfc[6][9][10];  // faces in bf_faces order: xm, xp, ym, yp, zm, zp
    double bwint[27][10];  // w3 * bint (init_bf_weights)
};

inline double p2val(double s) { return 1.5 * (s * s) - 0.5; }

inline void fill_products(double* v, int nb) {
    v[KYZ] = v[KY] * v[KZ];
    v[KZX] = v[KZ] * v[KX];
    v[KXY] = v[KX] * v[KY];
    if (nb == 8) v[KXYZ] = v[KX] * v[KY] * v[KZ];
    if (nb == 10) { v[KXX] = p2val(v[KX]); v[KYY] = p2val(v[KY]); v[KZZ] = p2val(v[KZ]); }
}

inline void build_tables(int nb, int q, Tables* t) {
    std::memset(t, 0, sizeof(Tables));
    t->q = q; t->nb = nb; t->nface = q * q; t->npg = q * q * q;
    if (q == 2) {
        const double s = 1. / std::sqrt(3.);
        t->xq[0] = -s; t->xq[1] = s; t->w1[0] = t->w1[1] = 1.;
    } else {
        const double s = std::sqrt(15.) / 5.;
        t->xq[0] = -s; t->xq[1] = 0.; t->xq[2] = s;
        t->w1[0] = 5. / 9.; t->w1[1] = 8. / 9.; t->w1[2] = 5. / 9.;
    }
    for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++) t->w2[i * q + j] = (q == 2) ? 1. : t->w1[j] * t->w1[i];
    for (int i = 0; i < q; i++)
        for (int p = 0; p < q * q; p++) t->w3[i * q * q + p] = (q == 2) ? 1. : t->w2[p] * t->w1[i];
    t->cb[K1] = 1.; t->cb[KX] = t->cb[KY] = t->cb[KZ] = 3.;
    if (nb > 4) t->cb[KYZ] = t->cb[KZX] = t->cb[KXY] = 9.;
    if (nb == 8) t->cb[KXYZ] = 27.;
    if (nb == 10) t->cb[KXX] = t->cb[KYY] = t->cb[KZZ] = 5.;
    for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++)
            for (int k = 0; k < q; k++) {
                double* v = t->bint[(i * q + j) * q + k];
                v[K1] = 1.; v[KX] = t->xq[i]; v[KY] = t->xq[j]; v[KZ] = t->xq[k];
                fill_products(v, nb);
            }
    for (int i = 0; i < q; i++)
        for (int j = 0; j < q; j++) {
            const int p = i * q + j;
            double* xp = t->bfc[F_XP][p];
            xp[K1] = 1.; xp[KX] = 1.; xp[KY] = t->xq[j]; xp[KZ] = t->xq[i]; fill_products(xp, nb);
            double* yp = t->bfc[F_YP][p];
            yp[K1] = 1.; yp[KX] = t->xq[i]; yp[KY] = 1.; yp[KZ] = t->xq[j]; fill_products(yp, nb);
            double* zp = t->bfc[F_ZP][p];
            zp[K1] = 1.; zp[KX] = t->xq[j]; zp[KY] = t->xq[i]; zp[KZ] = 1.; fill_products(zp, nb);
            double* xm = t->bfc[F_XM][p];
            double* ym = t->bfc[F_YM][p];
            double* zm = t->bfc[F_ZM][p];
            std::memcpy(xm, xp, sizeof(double) * 10);
            std::memcpy(ym, yp, sizeof(double) * 10);
            std::memcpy(zm, zp, sizeof(double) * 10);
            xm[KX] = -xp[KX]; xm[KZX] = -xp[KZX]; xm[KXY] = -xp[KXY];
            ym[KY] = -yp[KY]; ym[KYZ] = -yp[KYZ]; ym[KXY] = -yp[KXY];
            zm[KZ] = -zp[KZ]; zm[KYZ] = -zp[KYZ]; zm[KZX] = -zp[KZX];
            if (nb == 8) { xm[KXYZ] = -xp[KXYZ]; ym[KXYZ] = -yp[KXYZ]; zm[KXYZ] = -zp[KXYZ]; }
        }
    for (int g = 0; g < q * q * q; g++)
        for (int r = 0; r < nb; r++) t->bwint[g][r] = t->w3[g] * t->bint[g][r];
}

}  // namespace hd
