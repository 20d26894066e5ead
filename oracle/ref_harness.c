/*
 * Harness around the reference's own C particle_particle
 * (yukawa_p3m.c:182-416), compiled together with that file where it lies
 * under the reference tree (see oracle/Makefile, target "ref").  The
 * reference keeps positions and accelerations as arrays of row pointers and
 * takes the box length and Ewald parameters as scalars; this wrapper exposes
 * flat (n, 3) row-major arrays for ctypes.  It is a checker only.
 */
#include <stdlib.h>

void particle_particle(double **pos, double **a, double *Z, int n, double L,
                       double kappa, double G, double rc);

int ref_particle_particle(const double *pos, const double *Z, int n, double L,
                          double kappa, double G, double rc, double *acc)
{
    double **p = malloc((size_t)n * sizeof(double *));
    double **a = malloc((size_t)n * sizeof(double *));
    double *pbuf = malloc((size_t)n * 3 * sizeof(double));
    double *abuf = malloc((size_t)n * 3 * sizeof(double));
    double *zbuf = malloc((size_t)n * sizeof(double));
    if (!p || !a || !pbuf || !abuf || !zbuf) {
        free(p); free(a); free(pbuf); free(abuf); free(zbuf);
        return -1;
    }
    for (int i = 0; i < n; ++i) {
        p[i] = pbuf + 3 * i;
        a[i] = abuf + 3 * i;
        for (int d = 0; d < 3; ++d) pbuf[3 * i + d] = pos[3 * i + d];
        zbuf[i] = Z[i];
    }
    particle_particle(p, a, zbuf, n, L, kappa, G, rc);
    for (int i = 0; i < 3 * n; ++i) acc[i] = abuf[i];
    free(p); free(a); free(pbuf); free(abuf); free(zbuf);
    return 0;
}
