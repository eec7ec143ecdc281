/*
 * qsm_hip.h -- C ABI of libqsm_hip.so, the MI355X (gfx950) implementation of
 * the QSM.m primal-dual TV/TGV dipole-inversion path.
 *
 * The entry points replace the reference's C-MEX compute routines one for
 * one; the MEX wrappers in qsm.m_amd/mex/ keep the reference's mexFunction
 * argument layouts and call these.  Signatures use plain pointers and sizes
 * only.  Conventions shared by every entry point:
 *
 *   dtype    QSM_F64 (MATLAB double) or QSM_F32 (MATLAB single); the MEX
 *            layer picks it from mxIsSingle(prhs[0]) like the reference.
 *   sz[3]    nx, ny, nz; column-major, x fastest, l = i + nx*j + nx*ny*k.
 *   h[3]     voxel size (always double, as the reference reads prhs[k]).
 *   masks    uint8 (MATLAB logical), non-zero = inside.
 *   vector fields (w, p, q) are component-major: component c of voxel l is
 *            at c*N + l.
 *   return   QSM_OK or a QSM_E* code; qsm_last_error() describes the last
 *            failure of the calling thread.
 *
 * Host-pointer entry points are synchronous from the caller's view (they
 * return after the outputs are in host memory), like a MEX call.  The
 * *_dev entry points take device pointers and a hipStream_t passed as
 * void* (NULL = the null stream) and are asynchronous on that stream.
 */
#ifndef QSM_HIP_H
#define QSM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSM_OK       0
#define QSM_EARG     1   /* invalid argument (sizes, dtype, null pointer) */
#define QSM_EHIP     2   /* HIP runtime error */
#define QSM_ENOMEM   3   /* device allocation failed */
#define QSM_ENODEV   4   /* no usable HIP device */
#define QSM_EFULL    5   /* admission refused: over the HBM budget */

#define QSM_F64 0
#define QSM_F32 1

#define QSM_TV  0
#define QSM_TGV 1

/* Receives the progress lines the reference prints with mexPrintf
 * (pd_tgv_mex.c:292-293, 313-315).  NULL prints to stdout. */
typedef void (*qsm_print_fn)(const char *line, void *ctx);

/* Library version (major*10000 + minor*100 + patch). */
int qsm_version(void);

/* Text of the calling thread's last error (never NULL). */
const char *qsm_last_error(void);

/* ------------------------------------------------------------------------
 * Primal-dual dipole inversion (drop-in, host pointers, synchronous)
 * ------------------------------------------------------------------------ */

/* Replaces pd_tgv_mex.c:104-135 (mexFunction -> mx_init -> mx_pdd/mx_pdf ->
 * mxDuplicateArray(chi)).  d2f: solver input (the MEX's prhs[0], i.e.
 * -lapw(phase)); mask: prhs[1]; lam, alpha, tol, maxiter, verbose: prhs[3..7]
 * (alpha0 = alpha*lam, alpha1 = lam).  chi: N outputs.  maxiter == 0 is
 * rejected with QSM_EARG (the reference divides by zero there). */
int qsm_pd_tgv(int dtype, const void *d2f, const uint8_t *mask,
               const size_t sz[3], const double h[3],
               double lam, double alpha, double tol,
               uint32_t maxiter, uint32_t verbose,
               void *chi, qsm_print_fn print, void *print_ctx);

/* Replaces pd_tv_mex.c:100-130 (prhs: d2f, mask, vsz, lam, tol, maxiter,
 * verbose). */
int qsm_pd_tv(int dtype, const void *d2f, const uint8_t *mask,
              const size_t sz[3], const double h[3],
              double lam, double tol, uint32_t maxiter, uint32_t verbose,
              void *chi, qsm_print_fn print, void *print_ctx);

/* Replaces the per-echo body of sstgv.m:54-89 (sstgv_): crop phas (nechoes
 * volumes of sz, column-major) and mask to the mask's bounding box +/-2
 * voxels (sstgv.m:57-66), d2u = lapw(u, vsz) on the crop, then
 * pd_tgv_mex(-d2u, mask, vsz, lam, alpha, tol, maxit, verbose) when alpha >
 * eps, pd_tv_mex(-d2u, mask, vsz, lam, tol, maxit, verbose) otherwise
 * (sstgv.m:76-87), uncropped into zeros: x holds nechoes volumes of sz.  The
 * argument checks of sstgv.m:92-136 stay with the caller.  Every step after
 * the bounding box runs on the device; the same chi as those calls. */
int qsm_sstgv(int dtype, const void *phas, const uint8_t *mask, const size_t sz[3],
              size_t nechoes, const double vsz[3], double lam, double alpha,
              double tol, uint32_t maxit, uint32_t verbose, void *x,
              qsm_print_fn print, void *print_ctx);

/* ------------------------------------------------------------------------
 * Resident solver handle: the state of one inversion kept in HBM between
 * calls (used by the drop-in entry points above, by the benchmark, and by the
 * batching/serving runtime).  Replaces the mx_init/mx_pdd/mx_cleanup life
 * cycle (pd_tgv_mex.c:1124-1231) with an explicit one.
 * ------------------------------------------------------------------------ */
typedef struct qsm_pd qsm_pd;

/* kind: QSM_TV or QSM_TGV.  alpha is ignored for TV.  stream: hipStream_t
 * or NULL.  Allocates the state (TGV: 31 words/voxel, TV: 17 words/voxel,
 * plus one mask byte). */
int qsm_pd_create(qsm_pd **out, int kind, int dtype, const size_t sz[3],
                  const double h[3], double lam, double alpha, void *stream);

/* Bytes of device memory qsm_pd_create would allocate for this problem. */
size_t qsm_pd_footprint(int kind, int dtype, const size_t sz[3]);

/* Loads the input (N values of dtype) and the mask (N bytes), zeroes the
 * iterates and scales f by sigma (pd_tgv_mex.c:280-283).  on_device != 0:
 * d2f and mask are device pointers (device-to-device copies on the stream).
 * Returns after the handle's stream has drained: the work list of active
 * tiles (qsm_pd_set_active_items) is built from the loaded mask. */
int qsm_pd_load(qsm_pd *s, const void *d2f, const uint8_t *mask,
                int on_device);

/* Enqueues niter further iterations (the reference's loop body,
 * pd_tgv_mex.c:296-320, including the every-10th-iteration convergence
 * check against tol; tol <= 0 never stops).  Asynchronous: nothing is
 * waited for.  Iterations after a stop are no-ops on the device. */
int qsm_pd_iterate(qsm_pd *s, uint32_t niter, double tol);

/* Waits for the stream and reports: iterations done (the reference's
 * iteration count at exit), whether the tolerance test stopped the loop,
 * and the last convergence norms nr[2] (any pointer may be NULL). */
int qsm_pd_status(qsm_pd *s, uint32_t *iters, int *stopped, double nr[2]);

/* Convergence norms of every check so far: nr[2*c], nr[2*c+1] for check c
 * (iteration 10*(c+1)).  Returns the number of checks written (<= cap). */
int qsm_pd_norm_history(qsm_pd *s, double *nr, int cap);

/* Copies the current chi (N values) out; on_device != 0: chi is a device
 * pointer (copy enqueued on the stream, not waited for). */
int qsm_pd_get_chi(qsm_pd *s, void *chi, int on_device);

/* Frees the handle and its device memory. */
void qsm_pd_destroy(qsm_pd *s);

/* Selects the sweep schedule of a handle: 0 = fused passes that advance
 * two iterations each, with one-iteration passes where a convergence check
 * or the call's iteration count needs them (default); 1 = the reference's
 * two sweeps per iteration (dual kernel then primal kernel); 2 = fused
 * one-iteration passes only.  All give identical results. */
int qsm_pd_set_schedule(qsm_pd *s, int schedule);

/* Tile shape of the one-iteration fused pass: tx x ty voxels per block
 * (supported: 64x4, 64x8, 64x16, 32x8, 32x16, 32x32) marching kz planes.
 * Results do not depend on it; speed does (see the autotuner).  Any time
 * between calls. */
int qsm_pd_set_tiles(qsm_pd *s, int tx, int ty, int kz);
int qsm_pd_get_tiles(qsm_pd *s, int *tx, int *ty, int *kz);

/* z-chunk depth (planes per block) of the two-iteration fused pass, whose
 * tiles are fixed per precision (32x16 voxels FP32, 16x16 FP64); default:
 * at most one block per CU.  Results do not depend on it. */
int qsm_pd_set_kz2(qsm_pd *s, int kz);

/* Work list of the two-iteration pass for masked volumes.  At each load
 * the handle finds, per tile, the planes holding active voxels and, when
 * marching only those (in chunks of near-equal length) takes fewer plane
 * steps per CU than the uniform grid, launches one block per item instead.
 * Results are bitwise the same either way.  on: 0 the uniform grid, 1 the
 * list where the cost model favours it (default), 2 the list whenever a
 * voxel is active.  Takes effect at once on a loaded handle (the list is
 * rebuilt synchronously).  No counterpart in the reference (its loop runs
 * over every voxel, pd_tgv_mex.c:331-977). */
int qsm_pd_set_active_items(qsm_pd *s, int on);
/* Items of the current list (0: the uniform grid) and their chunk depth. */
int qsm_pd_active_items(qsm_pd *s, int *nitems, int *kz);

/* Kernel launches the handle has enqueued since it was created:
 * counts[0] two-iteration passes, [1] one-iteration fused passes, [2]
 * iterations of the two-sweep schedule, [3] convergence checks.  Launches
 * enqueued after a stop are counted (they exit at once on the device).  A
 * measurement aid with no counterpart in the reference. */
int qsm_pd_launch_counts(qsm_pd *s, uint64_t counts[4]);

/* ------------------------------------------------------------------------
 * Allocation-free solve
 * ------------------------------------------------------------------------ */

/* Bytes of the library's static device work space (per GPU, held from the
 * library's first use of that GPU until the process ends). */
#define QSM_PD_STATIC_BYTES (64u << 20)

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The solve of qsm_pd_tgv (kind QSM_TGV, alpha0 =
 * alpha*lam, alpha1 = lam) or qsm_pd_tv (kind QSM_TV, alpha ignored) with
 * bitwise the same chi, but on DEVICE pointers d2f, mask and chi of the
 * current device, asynchronous on stream (NULL: the null stream), without
 * verbose output, and with all solver state -- both banks, f, the packed
 * mask, control block, partial sums -- in the static work space instead of
 * an allocation: no hipMalloc, hipFree or memset / memcpy node is issued,
 * only kernels.  Volumes whose state exceeds QSM_PD_STATIC_BYTES are
 * refused with QSM_EARG (cubes up to about 81^3 voxels for FP32 TGV,
 * 64^3 for FP64 TGV, 108^3 for FP32 TV and 86^3 for FP64 TV).  Every
 * call on one GPU shares its work space and leaves its contents there:
 * calls on one stream run one after another, calls that overlap on one
 * GPU (other streams, other threads) race and must not be issued. */
int qsm_pd_solve_static(int kind, int dtype, const void *d2f, const uint8_t *mask,
                        const size_t sz[3], const double h[3], double lam, double alpha,
                        double tol, uint32_t maxiter, void *chi, void *stream);

/* ------------------------------------------------------------------------
 * Primal-dual single sweeps on device arrays (per-operator access; the
 * layouts are the reference's, asynchronous on stream)
 * ------------------------------------------------------------------------ */

/* pd_tgv_mex.c:493-646 / :331-485 (update_duald / update_dualf). */
int qsm_tgv_dual_dev(int dtype, void *nu, void *p, void *q,
                     const void *chi_, const void *eta_, const void *w_,
                     const void *f, const uint8_t *mi,
                     const size_t sz[3], const double h[3],
                     double alpha0, double alpha1, double sigma, void *stream);

/* pd_tgv_mex.c:824-977 / :658-812 (update_primald / update_primalf). */
int qsm_tgv_primal_dev(int dtype, void *chi, void *eta, void *w,
                       void *chi_, void *eta_, void *w_,
                       const void *nu, const void *p, const void *q,
                       const uint8_t *m, const size_t sz[3], const double h[3],
                       double tau, void *stream);

/* pd_tv_mex.c:417-513 / :313-410. */
int qsm_tv_dual_dev(int dtype, void *nu, void *p,
                    const void *chi_, const void *eta_,
                    const void *f, const uint8_t *mi,
                    const size_t sz[3], const double h[3],
                    double lam, double sigma, void *stream);

/* pd_tv_mex.c:630-727 / :523-620. */
int qsm_tv_primal_dev(int dtype, void *chi, void *eta,
                      void *chi_, void *eta_,
                      const void *nu, const void *p,
                      const uint8_t *m, const size_t sz[3], const double h[3],
                      double tau, void *stream);

/* pd_tgv_mex.c:1049-1121 (init_bitmask): mi from m. */
int qsm_init_mask_dev(uint8_t *mi, const uint8_t *m, const size_t sz[3],
                      void *stream);

/* pd_tgv_mex.c:980-1045 (convergence_checkd/f): writes out[0..1] = n1, n2
 * (host pointer; waits for the stream).  Returns the stop flag in *flag. */
int qsm_convergence_dev(int dtype, const void *chi, const void *chi0,
                        const void *eta, const void *eta0, size_t N,
                        double tol, double out[2], int *flag, void *stream);

/* pd_tgv_mex.c:272-278: tau and sigma for voxel size h. */
void qsm_step_sizes(const double h[3], double *tau, double *sigma);

/* ------------------------------------------------------------------------
 * Finite-difference operators (drop-in, host pointers, synchronous).
 * Outputs are overwritten in place, as the reference MEX overwrite their
 * caller-allocated prhs arrays.
 * ------------------------------------------------------------------------ */

/* gradfp_mex.c:13-47 -> gradfpd/gradfpf (:147-242 / :50-144). */
int qsm_gradfp(int dtype, void *dx, void *dy, void *dz, const void *u,
               const size_t sz[3], const double h[3]);
/* gradfp_adj_mex.c:15-49 -> gradfp_adjd/f (:156-257 / :52-153). */
int qsm_gradfp_adj(int dtype, void *du, const void *x, const void *y,
                   const void *z, const size_t sz[3], const double h[3]);
/* gradf_mex.c:13-47 -> gradfd/gradff (:149-245 / :50-146).  Needs every
 * dimension >= 2 (the reference reads out of bounds otherwise). */
int qsm_gradf(int dtype, void *dx, void *dy, void *dz, const void *u,
              const size_t sz[3], const double h[3]);
/* gradfm_mex.c:15-50 -> gradfmd/gradfmf (:105-154 / :53-102). */
int qsm_gradfm(int dtype, void *dx, void *dy, void *dz, const void *u,
               const uint8_t *G, const size_t sz[3], const double h[3]);
/* gradf_adj_mex.c:15-49 -> gradf_adjd/f (:152-249 / :52-149).  Needs every
 * dimension >= 2. */
int qsm_gradf_adj(int dtype, void *du, const void *x, const void *y,
                  const void *z, const size_t sz[3], const double h[3]);
/* gradfm_adj_mex.c:15-50 -> gradfm_adjd/f (:109-162 / :53-106). */
int qsm_gradfm_adj(int dtype, void *du, const void *x, const void *y,
                   const void *z, const uint8_t *G,
                   const size_t sz[3], const double h[3]);

/* Replace src/utils/fd/mex/gradb_mex.c:142-226 (backward differences,
 * forward on the first face), gradbm_mex.c:106-154 (masked: backward if the
 * -1 neighbour is in G, else forward, else 0; outside G untouched),
 * gradb_adj_mex.c:159-251 (-1/h forward differences summed (X + Y) + Z,
 * backward on the last face; (nx-1, ny-1, nz-1) is never written, as in the
 * reference), gradbm_adj_mex.c:110-162, gradc_mex.c:309-554 (central
 * differences, one-sided on the faces; the reference's k = 0 loop that
 * steps j by 1 rewrites l in [nx-1+nx, nx-1+nx(ny-1)) with dx backward, dy
 * central, dz forward, and (nx-1, ny-1, nz-1) is never written; both kept)
 * and gradcm_mex.c:122-186 (masked central differences).  Caller-allocated
 * outputs, overwritten where the reference writes; the unmasked forms need
 * every dimension >= 2.  Synchronous. */
int qsm_gradb(int dtype, void *dx, void *dy, void *dz, const void *u,
              const size_t sz[3], const double h[3]);
int qsm_gradbm(int dtype, void *dx, void *dy, void *dz, const void *u, const uint8_t *G,
               const size_t sz[3], const double h[3]);
int qsm_gradb_adj(int dtype, void *du, const void *x, const void *y, const void *z,
                  const size_t sz[3], const double h[3]);
int qsm_gradbm_adj(int dtype, void *du, const void *x, const void *y, const void *z,
                   const uint8_t *G, const size_t sz[3], const double h[3]);
int qsm_gradc(int dtype, void *dx, void *dy, void *dz, const void *u,
              const size_t sz[3], const double h[3]);
int qsm_gradcm(int dtype, void *dx, void *dy, void *dz, const void *u, const uint8_t *G,
               const size_t sz[3], const double h[3]);

/* lap1_mex.c:29-60 -> lapd/lapf (:136-205 / :63-133), the kernel of lap.m:
 * on G, d2u = sum over axes of 1/h^2 times a second difference chosen by
 * which neighbours lie in G (central, or one-sided of third, second or
 * first order); voxels outside G keep the caller's d2u.  Every extent must
 * be >= 2 (the reference reads outside the array on an axis of extent 1). */
int qsm_lap1(int dtype, void *d2u, const void *u, const uint8_t *G,
             const size_t sz[3], const double h[3]);

/* lapw_mex.c:21-51 -> lapwd/lapwf (:112-187 / :54-109).  Interior voxels
 * are written; the faces of d2u are left as the caller passed them. */
int qsm_lapw(int dtype, void *d2u, const void *u,
             const size_t sz[3], const double h[3]);

/* Device forms of the backward and central differences above. */
int qsm_gradb_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                  const size_t sz[3], const double h[3], void *stream);
int qsm_gradbm_dev(int dtype, void *dx, void *dy, void *dz, const void *u, const uint8_t *G,
                   const size_t sz[3], const double h[3], void *stream);
int qsm_gradb_adj_dev(int dtype, void *du, const void *x, const void *y, const void *z,
                      const size_t sz[3], const double h[3], void *stream);
int qsm_gradbm_adj_dev(int dtype, void *du, const void *x, const void *y, const void *z,
                       const uint8_t *G, const size_t sz[3], const double h[3], void *stream);
int qsm_gradc_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                  const size_t sz[3], const double h[3], void *stream);
int qsm_gradcm_dev(int dtype, void *dx, void *dy, void *dz, const void *u, const uint8_t *G,
                   const size_t sz[3], const double h[3], void *stream);

/* Device-pointer forms of the operators above (asynchronous on stream). */
int qsm_gradfp_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                   const size_t sz[3], const double h[3], void *stream);
int qsm_gradfp_adj_dev(int dtype, void *du, const void *x, const void *y,
                       const void *z, const size_t sz[3], const double h[3],
                       void *stream);
int qsm_gradf_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                  const size_t sz[3], const double h[3], void *stream);
int qsm_gradfm_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                   const uint8_t *G, const size_t sz[3], const double h[3],
                   void *stream);
int qsm_gradf_adj_dev(int dtype, void *du, const void *x, const void *y,
                      const void *z, const size_t sz[3], const double h[3],
                      void *stream);
int qsm_gradfm_adj_dev(int dtype, void *du, const void *x, const void *y,
                       const void *z, const uint8_t *G, const size_t sz[3],
                       const double h[3], void *stream);
int qsm_lapw_dev(int dtype, void *d2u, const void *u, const size_t sz[3],
                 const double h[3], void *stream);
int qsm_lap1_dev(int dtype, void *d2u, const void *u, const uint8_t *G,
                 const size_t sz[3], const double h[3], void *stream);

/* ------------------------------------------------------------------------
 * Batched inversions: many uneven volumes advanced together (no reference
 * counterpart -- QSM.m solves one volume per MEX call; this is the batching
 * runtime built on the same per-voxel arithmetic, so every volume's result
 * is bitwise the single-volume solve's).
 * ------------------------------------------------------------------------ */
typedef struct qsm_batch qsm_batch;

/* A batch of one dtype whose fused passes use tx x ty tiles.  hbm_budget:
 * bytes of volume state the batch may hold (admission limit).  stream:
 * the HIP stream the batch runs on; NULL makes the batch create (and
 * destroy) a stream of its own, since iterations are captured into graphs,
 * which the legacy null stream cannot hold. */
int qsm_batch_create(qsm_batch **out, int dtype, int tx, int ty,
                     size_t hbm_budget, int max_volumes, void *stream);

/* Device bytes one admitted volume holds. */
size_t qsm_batch_volume_bytes(int kind, int dtype, const size_t sz[3], int tx,
                              int ty, int kz, uint32_t maxiter);

/* Admits one inversion (kind QSM_TV/QSM_TGV, the arguments of pd_t*v_mex):
 * loads it, applies the dual of iteration 1 and makes it live; *slot
 * receives its slot.  QSM_EFULL when the HBM budget or the slots are
 * exhausted.  Waits for the stream. */
int qsm_batch_add(qsm_batch *b, int kind, const size_t sz[3], const double h[3],
                  double lam, double alpha, const void *d2f, const uint8_t *mask,
                  int on_device, uint32_t maxiter, double tol, int kz, int *slot);

/* Enqueues `chunk` iterations of every live volume (two launches per
 * iteration, replayed from a captured hipGraph).  Asynchronous. */
int qsm_batch_step(qsm_batch *b, uint32_t chunk);

/* Waits for the stream and refreshes the per-slot state. */
int qsm_batch_sync(qsm_batch *b);

/* Slot state after the last sync: used, live (still iterating), iterations. */
int qsm_batch_state(qsm_batch *b, int slot, int *used, int *live, uint32_t *iter);

/* Result of a slot: chi (host, N values; may be NULL), iterations run,
 * whether the tolerance stopped it, last norms.  Waits for the stream. */
int qsm_batch_result(qsm_batch *b, int slot, void *chi, uint32_t *iters,
                     int *stopped, double nr[2]);

/* Frees a slot and its HBM. */
int qsm_batch_remove(qsm_batch *b, int slot);

/* 1 (default): replay captured graphs; 0: launch directly. */
int qsm_batch_set_graph(qsm_batch *b, int use_graph);

size_t qsm_batch_used_bytes(qsm_batch *b);
void qsm_batch_destroy(qsm_batch *b);

/* ------------------------------------------------------------------------
 * FFT and the FFT dipole convolution
 * ------------------------------------------------------------------------ */

/* Smallest 2,3,5,7-smooth integer >= n: third_party/nextfastfft.m:59-123
 * (inputs above 2^20 are halved until in range, looked up, scaled back). */
size_t qsm_nextfastfft(size_t n);

/* In-place 3-D complex transform of an interleaved (re, im) array of
 * M[0] x M[1] x M[2] (x fastest): fft3.m:5 (inverse = 0) or ifft3.m:5
 * (inverse = 1, including 1/M per axis).  Lengths must have no prime
 * factor above 13 and be at most 2048 (double) / 4096 (single).  Asynchronous on the stream
 * (plans are built once per device and length, then cached). */
int qsm_fft3_dev(int dtype, void *data, const size_t M[3], int inverse,
                 void *stream);

/* Padded size and leading offset per axis of padfastfft(u, pad) ('both',
 * padfastfft.m:34-51): M = nextfastfft(sz + 2*pad) (sz where pad == 0). */
int qsm_dipole_pad_size(const size_t sz[3], const size_t pad[3], size_t M[3],
                        size_t off[3]);

/* Bytes of complex work space qsm_dipole_conv_dev needs. */
size_t qsm_dipole_work_bytes(int dtype, const size_t sz[3], const size_t pad[3]);

/* field = real(ifft3(D .* fft3(padfastfft(chi, pad)))) cropped back to sz:
 * the forward model the toolbox writes with dipoleKernel.m:56-66 (k-space
 * kernel D = 1/3 - (k.b)^2/|k|^2, D(0) = 0, bdir normalised) and
 * fft3/ifft3 (e.g. ndi.m:100-101, medi.m:151).  D is generated on the fly.
 * Device pointers; work: qsm_dipole_work_bytes bytes (no clearing
 * needed).  Asynchronous on the stream; work must stay allocated until the
 * stream has run the call. */
int qsm_dipole_conv_dev(int dtype, const void *chi, const size_t sz[3],
                        const double h[3], const double bdir[3],
                        const size_t pad[3], void *field, void *work,
                        void *stream);

/* Host-pointer form of qsm_dipole_conv_dev (synchronous). */
int qsm_dipole_conv(int dtype, const void *chi, const size_t sz[3],
                    const double h[3], const double bdir[3],
                    const size_t pad[3], void *field);

/* ------------------------------------------------------------------------
 * TV-regularised dipole inversion by ADMM (src/inversion/tv.m)
 * ------------------------------------------------------------------------ */

/* The opts fields of tv.m:38-86, same names and meaning. */
typedef struct qsm_tv_opts {
    double mu;      /* regularisation weight of the data term (7.5e3) */
    int tv;         /* 1 anisotropic, 2 isotropic total variation (1) */
    double rtol;    /* relative stopping tolerance (1e-1) */
    double atol;    /* absolute stopping tolerance (1e-4) */
    int maxit;      /* maximum ADMM iterations (20) */
    int verbose;    /* print the residual table to stdout (1) */
    double rho;     /* initial ADMM penalty (9) */
    double alpha;   /* over-relaxation (1.5) */
    double rtau;    /* penalty update: primal reduction ratio (0.7) */
    double stau;    /* penalty update: primal/dual balance (1e2) */
    double gamma;   /* penalty update factor (3) */
} qsm_tv_opts;

/* Fills *o with tv.m's defaults (tv.m:38-86). */
void qsm_tv_opts_default(qsm_tv_opts *o);

/* Bytes of device work space qsm_tv_dev needs for one sz[0] x sz[1] x sz[2]
 * echo. */
size_t qsm_tv_work_bytes(int dtype, const size_t sz[3]);

/* Replaces tv.m:1-104 for one echo (tv_, tv.m:107-310): x = mask .* u
 * after at most opts->maxit ADMM iterations on the local field f, with the
 * dipole kernel of dipoleKernel.m:56-66 on the unpadded grid (bdir
 * normalised) and periodic gradients (gradfp_mex / gradfp_adj_mex).  Every
 * length must have no prime factor above 13.  trace (may be NULL): 5
 * doubles per iteration done -- ||r||, eps_pri, ||s||, eps_dual (the
 * verbose table) and rho after the iteration's update; *iters (may be
 * NULL): iterations done.  Device pointers; work: qsm_tv_work_bytes bytes.
 * The stopping rule and the penalty update are taken on the device after
 * each iteration; the call enqueues opts->maxit iterations (those after the
 * rule stops return at once) and synchronises the stream once, at the end
 * (verbose: the residual table is printed then).  opts->maxit <= 65536. */
int qsm_tv_dev(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
               const double vsz[3], const double bdir[3], const qsm_tv_opts *opts,
               void *x, void *work, double *trace, int *iters, void *stream);

/* Host-pointer form of qsm_tv_dev (synchronous). */
int qsm_tv(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
           const double vsz[3], const double bdir[3], const qsm_tv_opts *opts,
           void *x, double *trace, int *iters);

/* ------------------------------------------------------------------------
 * Rapid two-step dipole inversion (src/inversion/rts.m)
 * ------------------------------------------------------------------------ */

/* The opts fields of rts.m:50-102, same names and meaning. */
typedef struct qsm_rts_opts {
    int tolls;      /* LSMR iterations of step 1 (4) */
    double mu;      /* weight of the well-conditioned k-space data (1e5) */
    double delta;   /* |D| threshold of the well-conditioned region (0.15) */
    int tv;         /* 1 anisotropic, 2 isotropic total variation (1) */
    double rtol;    /* relative stopping tolerance (1e-1) */
    double atol;    /* absolute stopping tolerance (1e-2) */
    int maxit;      /* maximum ADMM iterations (20) */
    int verbose;    /* print the residual table to stdout (1) */
    double rho;     /* initial ADMM penalty (10) */
    double alpha;   /* over-relaxation (1.5) */
    double rtau;    /* penalty update: primal reduction ratio (0.7) */
    double stau;    /* penalty update: primal/dual balance (1e2) */
    double gamma;   /* penalty update factor (3) */
} qsm_rts_opts;

/* Fills *o with rts.m's defaults (rts.m:50-102). */
void qsm_rts_opts_default(qsm_rts_opts *o);

/* Bytes of device work space qsm_rts_dev needs for one echo. */
size_t qsm_rts_work_bytes(int dtype, const size_t sz[3]);

/* Replaces rts.m:1-118 for one echo (rts_, rts.m:121-308): step 1 x1 =
 * mask .* real(ifft3(lsmr(D, fft3(f), 0, eps, eps, [], tolls))) (the LSMR
 * of third_party/lsmr/lsmr.m with the diagonal k-space operator D of
 * dipoleKernel.m:56-66), step 2 the ADMM of tv.m with the data term
 * M = mu (|D| > delta) on fft3(x1), starting from x1; x = mask .* x.  trace,
 * iters and the synchronisation as qsm_tv_dev.  Every length must have no
 * prime factor above 13. */
int qsm_rts_dev(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
                const double vsz[3], const double bdir[3], const qsm_rts_opts *opts,
                void *x, void *work, double *trace, int *iters, void *stream);

/* Host-pointer form of qsm_rts_dev (synchronous). */
int qsm_rts(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
            const double vsz[3], const double bdir[3], const qsm_rts_opts *opts,
            void *x, double *trace, int *iters);

/* ------------------------------------------------------------------------
 * Background field removal (src/bgremove/sharp.m, vsharp.m)
 * ------------------------------------------------------------------------ */

/* Bytes of device work space qsm_sharp_dev needs for a cropped volume szc,
 * voxel size vsz and radius rmax (the padding ceil(rmax/min(vsz)) of
 * sharp.m:53).  0 for arguments no call accepts: a bad dtype, an empty
 * crop, a non-finite or non-positive vsz or rmax. */
size_t qsm_sharp_work_bytes(int dtype, const size_t szc[3], const double vsz[3], double rmax);

/* The same for qsm_vsharp_dev with the largest radius rmax (vsharp.m:56)
 * and nechoes echoes: V-SHARP also keeps the spectrum of the mask and, for
 * one echo, that of f, reused for every radius.  0 on bad arguments. */
size_t qsm_vsharp_work_bytes(int dtype, const size_t szc[3], const double vsz[3], double rmax,
                             size_t nechoes);

/* Replaces sharp.m:47-84 after its crop: f and mask already cropped to the
 * mask's non-empty slices (cropIndices.m), szc that crop, nechoes echoes
 * stored one after another.  padfastfft(., ceil(r/min(vsz))) is implicit;
 * writes the local field flc (same layout) and the eroded mask mask_out.
 * Device pointers, asynchronous on the stream. */
int qsm_sharp_dev(int dtype, const void *fc, const uint8_t *mc, const size_t szc[3],
                  size_t nechoes, const double vsz[3], double r, double thr, void *flc,
                  uint8_t *mask_out, void *work, void *stream);

/* Replaces vsharp.m:47-153 after its crop, as qsm_sharp_dev, with the radii
 * rs[0..nrs-1] (any order, at least two; vsharp.m's default is
 * 9:-2*max(vsz):2*max(vsz)).  mask_out is the reference's returned mask:
 * the eroded mask of the first radius of the last pair (second smallest
 * radius for an even count) or of the last radius (odd count). */
int qsm_vsharp_dev(int dtype, const void *fc, const uint8_t *mc, const size_t szc[3],
                   size_t nechoes, const double vsz[3], const double *rs, size_t nrs,
                   double thr, void *flc, uint8_t *mask_out, void *work, void *stream);

/* Host-pointer forms of the whole functions sharp.m / vsharp.m: crop with
 * cropIndices.m, run, uncrop into zeros (fl: sz x nechoes, mask_out: sz;
 * may be NULL).  Synchronous. */
int qsm_sharp(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
              size_t nechoes, const double vsz[3], double r, double thr, void *fl,
              uint8_t *mask_out);
int qsm_vsharp(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
               size_t nechoes, const double vsz[3], const double *rs, size_t nrs,
               double thr, void *fl, uint8_t *mask_out);

/* ------------------------------------------------------------------------
 * Masked Poisson solver and Laplacian unwrapping
 * (src/utils/poisson_solver/mgpcg.m, src/unwrap/unwrapLaplacian.m)
 * ------------------------------------------------------------------------ */

/* mgopts of mgpcg.m:11-52 (nlevels <= 0: max(1, floor(log2(min(sz))) - 3)). */
typedef struct qsm_mg_opts {
    double tol;     /* multigrid tolerance; <= 0: fixed cycles (-1) */
    int maxit;      /* V-cycles per preconditioner application (1) */
    int mu;         /* 1 V-cycle, 2 W-cycle (1) */
    int npre;       /* pre-smoothing sweeps (1) */
    int npost;      /* post-smoothing sweeps (1) */
    int nboundary;  /* extra sweeps on the mask's boundary band (2) */
    int nlevels;    /* grid levels (0: the default) */
} qsm_mg_opts;

void qsm_mg_opts_default(qsm_mg_opts *o);

/* Bytes of device work space of qsm_mgpcg_dev / qsm_unwrap_laplacian_dev. */
size_t qsm_mgpcg_work_bytes(int dtype, const size_t sz[3], int nlevels);

/* Replaces mgpcg.m:1-58 / mgpcg_mex.c: v solves -Lap v = f on the mask
 * (7-point Laplacian of voxel size vsz, v = 0 off the mask and outside the
 * volume) by CG preconditioned with mgopts->maxit multigrid V-cycles,
 * stopping at ||r|| <= tol ||f|| (norms over the whole array) or maxit
 * iterations.  The smoother is red-black Gauss-Seidel (the reference's
 * GAUSS_SEIDEL_RED_BLACK variant); mgopts->tol > 0 stops the cycles of one
 * preconditioner application once ||r - A q|| < mgopts->tol ||r|| (mx_mg);
 * the coarsest level takes a fixed-degree Chebyshev iteration.
 * iters: iterations done; resnorm (may be NULL): maxit doubles, ||r|| per
 * iteration.  Device pointers; synchronises the stream (the CG scalars are
 * read on the host). */
int qsm_mgpcg_dev(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
                  const double vsz[3], double tol, int maxit, const qsm_mg_opts *mgopts,
                  void *v, void *work, int *iters, double *resnorm, void *stream);

/* Host-pointer form of qsm_mgpcg_dev (synchronous). */
int qsm_mgpcg(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
              const double vsz[3], double tol, int maxit, const qsm_mg_opts *mgopts,
              void *v, int *iters, double *resnorm);

/* Replaces mgpcg_mex.c:75-122 (mexFunction, mx_pcg): qsm_mgpcg from the
 * initial guess in v (host; overwritten with the solution): r = f - A v,
 * the rule still ||r|| <= tol ||f||.  As in the reference, the guess's
 * values off the mask stay in v and act as boundary values of the stencil
 * at the mask's edge; maxit = 0 returns v unchanged. */
int qsm_mgpcg_x0(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
                 const double vsz[3], double tol, int maxit, const qsm_mg_opts *mgopts,
                 void *v, int *iters, double *resnorm);

/* Replaces unwrapLaplacian.m:78-98 (unwrap_) for one echo of a volume
 * already cropped to the mask's box +1 voxel: d2u = lapw(u, vsz), uphas =
 * mgpcg(-mask .* d2u, mask, vsz, 1e-8, ceil(sqrt(numel(mask))), mgopts
 * with maxit = npre = npost = nboundary = 2).  Device pointers; work:
 * qsm_mgpcg_work_bytes(dtype, sz, 0). */
int qsm_unwrap_laplacian_dev(int dtype, const void *u, const uint8_t *mask,
                             const size_t sz[3], const double vsz[3], void *uphas, void *d2u,
                             void *work, int *iters, void *stream);

/* Host-pointer form of unwrapLaplacian.m:1-76: crop to the mask's box
 * grown by one voxel, per echo qsm_unwrap_laplacian_dev, uncrop into zeros
 * (uphas and, if not NULL, d2phas: sz x nechoes).  Synchronous. */
int qsm_unwrap_laplacian(int dtype, const void *phas, const uint8_t *mask, const size_t sz[3],
                         size_t nechoes, const double vsz[3], void *uphas, void *d2phas);

/* ------------------------------------------------------------------------
 * Device utilities
 * ------------------------------------------------------------------------ */

/* Device-to-device copy of `bytes` (16-byte aligned pointers) with a
 * bandwidth-saturating vector kernel on stream.  No reference counterpart:
 * moves volumes between resident solves and serves as the plain copy kernel
 * the HBM traffic counters are calibrated against. */
int qsm_copy_dev(void *dst, const void *src, size_t bytes, void *stream);

/* Streams narr (1..16) float arrays of n elements (array a at a*n) from in
 * to out, each read and written once, on stream.  No reference
 * counterpart: with narr = 15 it is the state traffic of one fused
 * primal-dual pass without its arithmetic, the bandwidth ceiling of that
 * access pattern (bench.py reports the fused pass against it). */
int qsm_stream_dev(void *out, const void *in, size_t n, int narr, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* QSM_HIP_H */
