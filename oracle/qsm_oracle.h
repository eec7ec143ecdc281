/*
 * qsm_oracle.h -- CPU restatement of the QSM.m primal-dual TV/TGV hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against; it is never linked into, loaded by or
 * called from the product library (libqsm_hip.so) or the MEX wrappers.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every function restates one reference routine (file:line given per
 * function) in plain C11 + OpenMP, keeping the reference's association order,
 * its coefficient rounding (double coefficients cast to float on the single
 * path) and its interior/mask semantics, and is built like the reference
 * (-std=c11 -O2 -fopenmp, no -march, hence no FMA contraction).
 *
 * Pinning (what checks this restatement, since the reference C-MEX cannot be
 * built here -- it needs MATLAB's mex.h):
 *   - tests/test_oracle_pins.py: closed forms on polynomials, adjointness of
 *     the dual/primal linear parts on masked domains, projection invariants,
 *     mask semantics against an independent bit-field formulation of the
 *     reference's init_bitmask, step-size identities, zero-data fixed point,
 *     convergence of the full loop to a feasible point, an independent numpy
 *     restatement on tiny volumes, and the committed golden fixtures.
 *   - tests/test_oracle_fd.py: fd operators on impulses and ramps, periodic
 *     transpose pair, lapw closed forms and wrap invariance.
 *
 * Layout: column-major, x fastest, l = i + nx*j + nx*ny*k.  Vector fields
 * (w, p: 3 components; q: 6 components) are stored component-major, i.e.
 * component c of voxel l at index c*N + l, as the reference allocates them.
 */
#ifndef QSM_ORACLE_H
#define QSM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- primal-dual building blocks (pd_tgv_mex.c / pd_tv_mex.c) ---------- */

/* Step sizes: pd_tgv_mex.c:272-278 (identical in pd_tv_mex.c:255-261). */
void orc_step_sizes(const double h[3], double *tau, double *sigma);

/* Interior mask mi from m: pd_tgv_mex.c:1049-1121 (init_bitmask). */
void orc_init_mask(uint8_t *mi, const uint8_t *m, const size_t sz[3]);

/* TGV dual update: pd_tgv_mex.c:493-646 (double), :331-485 (float). */
void orc_tgv_dual_d(double *nu, double *p, double *q,
                    const double *chi_, const double *eta_, const double *w_,
                    const double *f, const uint8_t *mi,
                    const size_t sz[3], const double h[3],
                    double alpha0, double alpha1, double sigma);
void orc_tgv_dual_f(float *nu, float *p, float *q,
                    const float *chi_, const float *eta_, const float *w_,
                    const float *f, const uint8_t *mi,
                    const size_t sz[3], const double h[3],
                    double alpha0, double alpha1, double sigma);

/* TGV primal update: pd_tgv_mex.c:824-977 (double), :658-812 (float). */
void orc_tgv_primal_d(double *chi, double *eta, double *w,
                      double *chi_, double *eta_, double *w_,
                      const double *nu, const double *p, const double *q,
                      const uint8_t *m, const size_t sz[3], const double h[3],
                      double tau);
void orc_tgv_primal_f(float *chi, float *eta, float *w,
                      float *chi_, float *eta_, float *w_,
                      const float *nu, const float *p, const float *q,
                      const uint8_t *m, const size_t sz[3], const double h[3],
                      double tau);

/* TV dual update: pd_tv_mex.c:417-513 (double), :313-410 (float). */
void orc_tv_dual_d(double *nu, double *p,
                   const double *chi_, const double *eta_,
                   const double *f, const uint8_t *mi,
                   const size_t sz[3], const double h[3],
                   double lam, double sigma);
void orc_tv_dual_f(float *nu, float *p,
                   const float *chi_, const float *eta_,
                   const float *f, const uint8_t *mi,
                   const size_t sz[3], const double h[3],
                   double lam, double sigma);

/* TV primal update: pd_tv_mex.c:630-727 (double), :523-620 (float). */
void orc_tv_primal_d(double *chi, double *eta, double *chi_, double *eta_,
                     const double *nu, const double *p,
                     const uint8_t *m, const size_t sz[3], const double h[3],
                     double tau);
void orc_tv_primal_f(float *chi, float *eta, float *chi_, float *eta_,
                     const float *nu, const float *p,
                     const uint8_t *m, const size_t sz[3], const double h[3],
                     double tau);

/* Convergence check: pd_tgv_mex.c:980-1011 (double), :1014-1045 (float). */
uint8_t orc_convergence_d(double *nr1, double *nr2,
                          const double *chi, const double *chi0,
                          const double *eta, const double *eta0,
                          double tol, size_t N);
uint8_t orc_convergence_f(double *nr1, double *nr2,
                          const float *chi, const float *chi0,
                          const float *eta, const float *eta0,
                          double tol, size_t N);

/* ---- whole solves (mexFunction + mx_init + mx_pdd/mx_pdf) --------------
 * dtype: 0 = double, 1 = single.  d2f is the solver input (the MEX's
 * prhs[0]); it is not modified.  chi_out receives chi (N values).
 * iters_out (may be NULL) receives the number of iterations run; nr_out
 * (may be NULL, 2 doubles) receives the last convergence norms.
 * verbose > 0 prints the reference's progress table to stdout.
 * Returns 0, or -1 on bad arguments (maxiter == 0: the reference divides by
 * zero computing its print period, pd_tgv_mex.c:290).
 * pd_tgv_mex.c:104-135, 232-323 (double), 138-229 (float), 1124-1164. */
int orc_pd_tgv(int dtype, const void *d2f, const uint8_t *mask,
               const size_t sz[3], const double h[3],
               double lam, double alpha, double tol,
               uint32_t maxiter, uint32_t verbose,
               void *chi_out, uint32_t *iters_out, double *nr_out);

/* pd_tv_mex.c:100-130, 221-306 (double), 133-218 (float), 874-909. */
int orc_pd_tv(int dtype, const void *d2f, const uint8_t *mask,
              const size_t sz[3], const double h[3],
              double lam, double tol,
              uint32_t maxiter, uint32_t verbose,
              void *chi_out, uint32_t *iters_out, double *nr_out);

/* Seconds the last orc_pd_tgv/orc_pd_tv call spent in its iteration loop
 * (clock_gettime(CLOCK_MONOTONIC); allocation and set-up excluded). */
double orc_last_loop_seconds(void);

/* Sets the OpenMP thread count of the restatement's loops (n > 0). */
void orc_set_threads(int n);

/* ---- finite-difference operators (outputs overwritten in place) -------- */

/* gradfp_mex.c:147-242 / :50-144  periodic forward gradient. */
void orc_gradfp(int dtype, void *dx, void *dy, void *dz, const void *u,
                const size_t sz[3], const double h[3]);
/* gradfp_adj_mex.c:156-257 / :52-153  -(periodic backward divergence). */
void orc_gradfp_adj(int dtype, void *du, const void *x, const void *y,
                    const void *z, const size_t sz[3], const double h[3]);
/* gradf_mex.c:149-245 / :50-146  forward gradient, last face backward. */
void orc_gradf(int dtype, void *dx, void *dy, void *dz, const void *u,
               const size_t sz[3], const double h[3]);
/* gradfm_mex.c:105-154 / :53-102  masked forward/backward gradient. */
void orc_gradfm(int dtype, void *dx, void *dy, void *dz, const void *u,
                const uint8_t *G, const size_t sz[3], const double h[3]);
/* gradf_adj_mex.c:152-249 / :52-149  -backward divergence, first face fwd. */
void orc_gradf_adj(int dtype, void *du, const void *x, const void *y,
                   const void *z, const size_t sz[3], const double h[3]);
/* gradfm_adj_mex.c:109-162 / :53-106  masked variant. */
/* gradb_mex.c, gradbm_mex.c, gradb_adj_mex.c, gradbm_adj_mex.c, gradc_mex.c,
 * gradcm_mex.c (src/utils/fd/mex): backward and central differences. */
void orc_gradb(int dtype, void *dx, void *dy, void *dz, const void *u,
               const size_t sz[3], const double h[3]);
void orc_gradbm(int dtype, void *dx, void *dy, void *dz, const void *u,
                const uint8_t *G, const size_t sz[3], const double h[3]);
void orc_gradb_adj(int dtype, void *du, const void *x, const void *y,
                   const void *z, const size_t sz[3], const double h[3]);
void orc_gradbm_adj(int dtype, void *du, const void *x, const void *y,
                    const void *z, const uint8_t *G, const size_t sz[3], const double h[3]);
/* lap1_mex.c (src/utils/fd/mex): the masked Laplacian of lap.m. */
void orc_lap1(int dtype, void *du, const void *u, const uint8_t *G, const size_t sz[3],
              const double h[3]);
void orc_gradc(int dtype, void *dx, void *dy, void *dz, const void *u,
               const size_t sz[3], const double h[3]);
void orc_gradcm(int dtype, void *dx, void *dy, void *dz, const void *u,
                const uint8_t *G, const size_t sz[3], const double h[3]);

void orc_gradfm_adj(int dtype, void *du, const void *x, const void *y,
                    const void *z, const uint8_t *G,
                    const size_t sz[3], const double h[3]);
/* lapw_mex.c:112-187 / :54-109  Laplacian of wrapped phase (interior). */
void orc_lapw(int dtype, void *d2u, const void *u,
              const size_t sz[3], const double h[3]);

#ifdef __cplusplus
}
#endif

#endif /* QSM_ORACLE_H */
