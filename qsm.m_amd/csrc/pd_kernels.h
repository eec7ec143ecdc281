// pd_kernels.h -- device kernels of the primal-dual TV/TGV solver.
//
// Two schedules compute the same iterates bit for bit:
//
//  * two sweeps per iteration (the reference's structure,
//    pd_tgv_mex.c:306-307): dual_sweep then primal_sweep, in place, with
//    the extrapolated primal variables (chi_, eta_, w_) stored;
//
//  * one fused pass per iteration: pass n applies the primal update of
//    iteration n and then the dual update of iteration n+1 to a z-marching
//    column of tiles.  The extrapolated variables are consumed straight from
//    registers/LDS and never stored, so a pass moves 31 state words per
//    voxel (TGV: read chi, eta, w[3], nu, p[3], q[6], f; write the 15 state
//    words) instead of 51.  The dual of a tile needs the primal of its x/y
//    neighbours, which the block recomputes on a one-cell ring; every
//    block therefore reads the state of iteration n from one bank and
//    writes iteration n+1 to the other (ping-pong), so ring reads never
//    race with owners' writes.
#pragma once

#include "pd_math.h"

namespace qsm {

// Device-side loop control of one solve (one per handle / batch volume).
struct Ctl {
    int stop;            // set by the tolerance test; later kernels no-op
    uint32_t stop_iter;  // iteration the test stopped at
    uint32_t checks;     // convergence checks done so far
    uint32_t stop_bank;  // bank holding chi of stop_iter (fused schedules)
    double nr[2];        // last n1, n2
};

// Pointers of one state bank.  TV leaves w*, q* null.
template <typename T>
struct Bank {
    T *chi, *eta, *w1, *w2, *w3;
    T *nu, *p1, *p2, *p3, *q1, *q2, *q3, *q4, *q5, *q6;
};

// Extrapolated primal arrays used by the two-sweep schedule.
template <typename T>
struct Bars {
    T *cb, *eb, *b1, *b2, *b3;
};

struct Geom {
    int nx, ny, nz;
};

// Launchers (return a hipError_t from the launch).
template <typename T, bool TGV>
hipError_t launch_dual_sweep(const Bank<T> &s, const Bars<T> &b, const T *f,
                             const uint8_t *mk, uint8_t bit, Geom g,
                             const DualCoef<T> &c, const int *stop,
                             hipStream_t st);

template <typename T, bool TGV>
hipError_t launch_primal_sweep(const Bank<T> &s, const Bars<T> &b,
                               const uint8_t *mk, uint8_t bit, Geom g,
                               const PrimalCoef<T> &c, const int *stop,
                               double *partials, hipStream_t st);

// Number of partial-sum slots primal_sweep writes (one per block).
int primal_sweep_blocks(Geom g);

// Fused pass tile shapes (TX x TY inner cells, one thread each).
struct FusedShape {
    int tx, ty, kz;
};

// in/out: bank bases, array a of voxel l at a*N + l (TGV: chi, eta, w1-3,
// nu, p1-3, q1-6; TV: chi, eta, nu, p1-3).
template <typename T, bool TGV>
hipError_t launch_fused(const T *in, T *out, size_t N, const T *f,
                        const uint8_t *mk, Geom g, FusedShape sh,
                        const DualCoef<T> &dc, const PrimalCoef<T> &pc,
                        const int *stop, double *partials, hipStream_t st);

int fused_blocks(Geom g, FusedShape sh);
bool fused_shape_supported(int tx, int ty);

// One work item of the two-iteration pass: tile (bx, by), owned planes
// [k0, k1).
struct F2Item {
    int bx, by, k0, k1;
};

// Two-iteration pass (fused2.hip): reads bank `in` (primal n-1, dual n)
// and writes bank `out` (primal n+1, dual n+2); partials (if non-null)
// receive the convergence sums of primal n+1, one slot per block.  Tiles of
// fused2_tile(dtype) marching kz planes, or, with items, one block per
// item (nitems blocks).
template <typename T, bool TGV>
hipError_t launch_fused2(const T *in, T *out, size_t N, const T *f,
                         const uint8_t *mk, Geom g, int kz,
                         const DualCoef<T> &dc, const PrimalCoef<T> &pc,
                         const int *stop, double *partials, const F2Item *items, int nitems,
                         hipStream_t st);
void fused2_tile(int dtype, int &tx, int &ty);
int fused2_blocks(Geom g, int dtype, int kz);
// Per tile of fused2_tile(dtype), row-major: out[2t], out[2t+1] = the plane
// range [lo, hi) in [1, nz - 1) holding a voxel with a mask bit (lo == hi:
// none).
hipError_t launch_tile_zrange(const uint8_t *mask, Geom g, int dtype, int *out, hipStream_t st);
// The fused pass addresses a bank with 32-bit byte offsets (< 4 GiB).
bool fused_size_supported(int kind, int dtype, size_t N);

// Reduces nblk partial sums (4 doubles each) and applies the tolerance test
// for iteration iter, whose chi is in bank `bank` (recorded on a stop).
hipError_t launch_conv_finalize(const double *partials, int nblk, Ctl *ctl,
                                double *hist, int hist_cap, double tol,
                                uint32_t iter, uint32_t bank, hipStream_t st);

// Generic 4-sum over N voxels of (chi, chi0, eta, eta0) into partials.
template <typename T>
hipError_t launch_conv_sums(const T *chi, const T *chi0, const T *eta,
                            const T *eta0, size_t N, double *partials,
                            int nblk, hipStream_t st);

// Packed mask: bit 0 = primal active (interior and m != 0), bit 1 = dual
// active (interior and all six face neighbours in m, pd_tgv_mex.c:1049-1121).
// With raw != 0 the output is the reference's mi array (0/1) instead.
hipError_t launch_mask_init(uint8_t *out, const uint8_t *m, Geom g, int raw,
                            hipStream_t st);

// Resets a Ctl: no stop, no checks, last norms 1.0 (pd_tgv_mex.c:287-288).
hipError_t launch_ctl_reset(Ctl *ctl, int n, hipStream_t st);

// dst = chi of the bank that holds the result: ctl->stop_bank when the
// tolerance test stopped the loop, else cur_bank (fused schedules).
template <typename T>
hipError_t launch_select_copy(T *dst, const T *chi0, const T *chi1,
                              const Ctl *ctl, uint32_t cur_bank, size_t N,
                              hipStream_t st);

// Zeroes `bytes` bytes at p (p 16-byte aligned) with a kernel.
hipError_t launch_zero_bytes(void *p, size_t bytes, hipStream_t st);

// f = sg * d2f (pd_tgv_mex.c:280-283 / :186-189).
template <typename T>
hipError_t launch_scale(T *f, const T *src, T sg, size_t N, hipStream_t st);

}  // namespace qsm
