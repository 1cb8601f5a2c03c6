/*
 * charm_geom.c -- FP64 element geometry of the cube octree: the Taylor-basis
 * mass matrix of a level-l hexahedron by the 8-point Gauss rule and its
 * inverse by Gauss-Jordan elimination with partial pivoting
 * (charm_geom_quad_calc, src/charm_geom.c:496-530; charm_matr_inv,
 * src/charm_globals.c:285-394).
 *
 * On an axis-aligned octree the matrix depends on the level only (every
 * element of a level is a translate of one cube), so one FP64 inversion per
 * level serves all elements; the kernels take the inverse's diagonal.
 */
#include <math.h>
#include <string.h>
#include "charm_internal.h"

/* Gauss-Jordan inverse of a 4x4 matrix, partial pivoting on |a[r][i]| with
 * ties to the later row; returns 0 on a zero pivot */
int charm_gj_inverse4(const double src[4][4], double inv[4][4])
{
    double a[4][4];
    int i, j, r;
    memcpy(a, src, sizeof(a));
    for (i = 0; i < 4; i++)
        for (j = 0; j < 4; j++) inv[i][j] = (i == j) ? 1.0 : 0.0;
    for (i = 0; i < 4; i++) {
        int p = i;
        double best = fabs(a[i][i]);
        for (r = i + 1; r < 4; r++)
            if (best <= fabs(a[r][i])) { best = fabs(a[r][i]); p = r; }
        if (a[p][i] == 0.0) return 0;
        if (p != i)
            for (j = 0; j < 4; j++) {
                double t = a[i][j]; a[i][j] = a[p][j]; a[p][j] = t;
                t = inv[i][j]; inv[i][j] = inv[p][j]; inv[p][j] = t;
            }
        {
            const double piv = a[i][i];
            for (j = 0; j < 4; j++) { a[i][j] /= piv; inv[i][j] /= piv; }
        }
        for (r = 0; r < 4; r++) {
            if (r == i) continue;
            {
                const double f = a[r][i];
                for (j = 0; j < 4; j++) { a[r][j] -= f * a[i][j]; inv[r][j] -= f * inv[i][j]; }
            }
        }
    }
    return 1;
}

/* mass matrix a_ij = sum_gp wJ phi_i phi_j of the level-l cube with the
 * Taylor basis (1, (x-xc)/h, (y-yc)/h, (z-zc)/h) and Gauss points at
 * +-1/sqrt(3) of the reference cube (wJ = h^3/8), inverted in FP64 */
int charm_level_mass_inverse(int level, double a[4][4], double ainv[4][4])
{
    const double h = ldexp(1.0, -level), s3 = 1.0 / sqrt(3.0), wj = h * h * h / 8.0;
    int g, i, j;
    memset(a, 0, 16 * sizeof(double));
    for (g = 0; g < 8; g++) {
        double phi[4];
        phi[0] = 1.0;
        phi[1] = ((g & 1) ? s3 : -s3) * 0.5;
        phi[2] = ((g & 2) ? s3 : -s3) * 0.5;
        phi[3] = ((g & 4) ? s3 : -s3) * 0.5;
        for (i = 0; i < 4; i++)
            for (j = 0; j < 4; j++) a[i][j] += wj * phi[i] * phi[j];
    }
    if (!charm_gj_inverse4((const double (*)[4]) a, ainv)) {
        charm_set_error("mass matrix of level %d is singular", level);
        return CHARM_ERR_ARG;
    }
    return CHARM_OK;
}
