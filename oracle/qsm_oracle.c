/*
 * qsm_oracle.c -- CPU restatement of the QSM.m primal-dual TV/TGV hot path
 * and its finite-difference operators.  TEST INFRASTRUCTURE ONLY: the checker
 * for the HIP library, never part of it (see qsm_oracle.h for the pins).
 *
 * Build like the reference MEX (src/inversion/mex/make.m:7-8):
 *   gcc -std=c11 -O2 -fopenmp -fPIC -shared  (no -march: no FMA contraction)
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qsm_oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif
#include <time.h>

/* Seconds spent in the iteration loop of the last whole solve (the
 * reference's loop only: allocation and mask set-up excluded). */
static double orc_loop_seconds_ = 0.0;

static double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double orc_last_loop_seconds(void) { return orc_loop_seconds_; }

/* Thread count of the OpenMP loops (for the one-thread CPU baseline). */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* shared pieces                                                            */
/* ------------------------------------------------------------------------ */

/* pd_tgv_mex.c:272-278 and pd_tv_mex.c:255-261: closed-form step sizes from
 * the operator-norm bound ||K||^2 <= gnrm^2 + wnrm^2. */
void orc_step_sizes(const double h[3], double *tau, double *sigma)
{
    const double gnrm = 4.0 * (1.0 / (h[0] * h[0]) + 1.0 / (h[1] * h[1])
                               + 1.0 / (h[2] * h[2]));
    const double wnrm = 6.0 * (1.0 / (h[0] * h[0]) + 1.0 / (h[1] * h[1])
                               + 2.0 / (h[2] * h[2])) / 3.0;
    const double K2 = gnrm * gnrm + wnrm * wnrm;
    const double t = 1 / sqrt(K2);
    *tau = t;
    *sigma = 1 / (t * K2);
}

/* pd_tgv_mex.c:1049-1121.  The reference builds a 13-bit neighbour code per
 * voxel and sets mi where the "both neighbours present" bit of all three axes
 * is set; that is: all six face neighbours exist and are in m.  The voxel's
 * own m value is not consulted. */
void orc_init_mask(uint8_t *mi, const uint8_t *m, const size_t sz[3])
{
    const size_t nx = sz[0], ny = sz[1], nz = sz[2];
    const size_t sy = nx, sk = nx * ny;
    for (size_t k = 0; k < nz; ++k)
        for (size_t j = 0; j < ny; ++j)
            for (size_t i = 0; i < nx; ++i) {
                const size_t l = i + sy * j + sk * k;
                const int ix = (i > 0 && m[l - 1]) && (i + 1 < nx && m[l + 1]);
                const int iy = (j > 0 && m[l - sy]) && (j + 1 < ny && m[l + sy]);
                const int iz = (k > 0 && m[l - sk]) && (k + 1 < nz && m[l + sk]);
                mi[l] = (uint8_t)(ix && iy && iz);
            }
}

/* pd_tgv_mex.c:980-1011 (identical in pd_tv_mex.c:730-761). */
uint8_t orc_convergence_d(double *nr1, double *nr2,
                          const double *chi, const double *chi0,
                          const double *eta, const double *eta0,
                          double tol, size_t N)
{
    double n1 = 0.0, n2 = 0.0, d1 = 0.0, d2 = 0.0;
    long l;
    #pragma omp parallel for reduction(+: n1, n2, d1, d2) schedule(static)
    for (l = 0; l < (long)N; ++l) {
        d1 += chi[l] * chi[l];
        d2 += eta[l] * eta[l];
        n1 += (chi[l] - chi0[l]) * (chi[l] - chi0[l]);
        n2 += (eta[l] - eta0[l]) * (eta[l] - eta0[l]);
    }
    n1 = sqrt(n1 / d1);
    n2 = sqrt(n2 / d2);
    *nr1 = n1;
    *nr2 = n2;
    return (uint8_t)((n1 < tol) && (n2 < 2.0 * tol));
}

/* pd_tgv_mex.c:1014-1045: float state, double accumulators.  The second
 * factor of each difference term subtracts in double (the first in float),
 * as the reference writes it. */
uint8_t orc_convergence_f(double *nr1, double *nr2,
                          const float *chi, const float *chi0,
                          const float *eta, const float *eta0,
                          double tol, size_t N)
{
    double n1 = 0.0, n2 = 0.0, d1 = 0.0, d2 = 0.0;
    long l;
    #pragma omp parallel for reduction(+: n1, n2, d1, d2) schedule(static)
    for (l = 0; l < (long)N; ++l) {
        d1 += ((double)chi[l]) * ((double)chi[l]);
        d2 += ((double)eta[l]) * ((double)eta[l]);
        n1 += ((double)(chi[l] - chi0[l])) * ((double)chi[l] - chi0[l]);
        n2 += ((double)(eta[l] - eta0[l])) * ((double)eta[l] - eta0[l]);
    }
    n1 = sqrt(n1 / d1);
    n2 = sqrt(n2 / d2);
    *nr1 = n1;
    *nr2 = n2;
    return (uint8_t)((n1 < tol) && (n2 < 2.0 * tol));
}

/* ------------------------------------------------------------------------ */
/* precision-generic sweeps and solves                                      */
/* ------------------------------------------------------------------------ */

#define R double
#define SQRTR sqrt
#define C_(x) ((double)(x))
#define FN(name) name##_d
#include "pd_body.inc"
#undef R
#undef SQRTR
#undef C_
#undef FN

#define R float
#define SQRTR sqrtf
#define C_(x) ((float)(x))
#define FN(name) name##_f
#include "pd_body.inc"
#undef R
#undef SQRTR
#undef C_
#undef FN

int orc_pd_tgv(int dtype, const void *d2f, const uint8_t *mask,
               const size_t sz[3], const double h[3],
               double lam, double alpha, double tol,
               uint32_t maxiter, uint32_t verbose,
               void *chi_out, uint32_t *iters_out, double *nr_out)
{
    if (maxiter == 0)
        return -1;
    if (dtype == 0)
        run_tgv_d((const double *)d2f, mask, sz, h, lam, alpha, tol, maxiter,
                  verbose, (double *)chi_out, iters_out, nr_out);
    else
        run_tgv_f((const float *)d2f, mask, sz, h, lam, alpha, tol, maxiter,
                  verbose, (float *)chi_out, iters_out, nr_out);
    return 0;
}

int orc_pd_tv(int dtype, const void *d2f, const uint8_t *mask,
              const size_t sz[3], const double h[3],
              double lam, double tol,
              uint32_t maxiter, uint32_t verbose,
              void *chi_out, uint32_t *iters_out, double *nr_out)
{
    if (maxiter == 0)
        return -1;
    if (dtype == 0)
        run_tv_d((const double *)d2f, mask, sz, h, lam, tol, maxiter, verbose,
                 (double *)chi_out, iters_out, nr_out);
    else
        run_tv_f((const float *)d2f, mask, sz, h, lam, tol, maxiter, verbose,
                 (float *)chi_out, iters_out, nr_out);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* finite-difference operators                                              */
/* ------------------------------------------------------------------------ */

#define R double
#define C_(x) ((double)(x))
#define FMODR fmod
#define FN(name) name##_d
#include "fd_body.inc"
#undef R
#undef C_
#undef FMODR
#undef FN

#define R float
#define C_(x) ((float)(x))
#define FMODR fmodf
#define FN(name) name##_f
#include "fd_body.inc"
#undef R
#undef C_
#undef FMODR
#undef FN

void orc_gradfp(int dtype, void *dx, void *dy, void *dz, const void *u,
                const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradfp_d(dx, dy, dz, u, sz, h);
    else gradfp_f(dx, dy, dz, u, sz, h);
}

void orc_gradfp_adj(int dtype, void *du, const void *x, const void *y,
                    const void *z, const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradfp_adj_d(du, x, y, z, sz, h);
    else gradfp_adj_f(du, x, y, z, sz, h);
}

void orc_gradf(int dtype, void *dx, void *dy, void *dz, const void *u,
               const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradf_d(dx, dy, dz, u, sz, h);
    else gradf_f(dx, dy, dz, u, sz, h);
}

void orc_gradfm(int dtype, void *dx, void *dy, void *dz, const void *u,
                const uint8_t *G, const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradfm_d(dx, dy, dz, u, G, sz, h);
    else gradfm_f(dx, dy, dz, u, G, sz, h);
}

void orc_gradf_adj(int dtype, void *du, const void *x, const void *y,
                   const void *z, const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradf_adj_d(du, x, y, z, sz, h);
    else gradf_adj_f(du, x, y, z, sz, h);
}

void orc_gradb(int dtype, void *dx, void *dy, void *dz, const void *u,
               const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradb_d(dx, dy, dz, u, sz, h);
    else gradb_f(dx, dy, dz, u, sz, h);
}

void orc_gradbm(int dtype, void *dx, void *dy, void *dz, const void *u,
                const uint8_t *G, const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradbm_d(dx, dy, dz, u, G, sz, h);
    else gradbm_f(dx, dy, dz, u, G, sz, h);
}

void orc_gradb_adj(int dtype, void *du, const void *x, const void *y,
                   const void *z, const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradb_adj_d(du, x, y, z, sz, h);
    else gradb_adj_f(du, x, y, z, sz, h);
}

void orc_gradbm_adj(int dtype, void *du, const void *x, const void *y,
                    const void *z, const uint8_t *G, const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradbm_adj_d(du, x, y, z, G, sz, h);
    else gradbm_adj_f(du, x, y, z, G, sz, h);
}

void orc_lap1(int dtype, void *du, const void *u, const uint8_t *G, const size_t sz[3],
              const double h[3])
{
    if (dtype == 0) lap1_d(du, u, G, sz, h);
    else lap1_f(du, u, G, sz, h);
}

void orc_gradc(int dtype, void *dx, void *dy, void *dz, const void *u,
               const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradc_d(dx, dy, dz, u, sz, h);
    else gradc_f(dx, dy, dz, u, sz, h);
}

void orc_gradcm(int dtype, void *dx, void *dy, void *dz, const void *u,
                const uint8_t *G, const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradcm_d(dx, dy, dz, u, G, sz, h);
    else gradcm_f(dx, dy, dz, u, G, sz, h);
}

void orc_gradfm_adj(int dtype, void *du, const void *x, const void *y,
                    const void *z, const uint8_t *G,
                    const size_t sz[3], const double h[3])
{
    if (dtype == 0) gradfm_adj_d(du, x, y, z, G, sz, h);
    else gradfm_adj_f(du, x, y, z, G, sz, h);
}

void orc_lapw(int dtype, void *d2u, const void *u,
              const size_t sz[3], const double h[3])
{
    if (dtype == 0) lapw_d(d2u, u, sz, h);
    else lapw_f(d2u, u, sz, h);
}
