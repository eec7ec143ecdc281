// poisson.hip -- multigrid-preconditioned CG for the masked Poisson problem
// (src/utils/poisson_solver/mgpcg.m, mex/mgpcg_mex.c) and Laplacian phase
// unwrapping (src/unwrap/unwrapLaplacian.m) on gfx950.
//
// The system (lapmg_mex.c): on the voxels of the mask G, -Lap_h v = f with
// the 7-point Laplacian, v = 0 off the mask and on a one-voxel border added
// around the volume (mgpcg_mex.c: mx_pad_boundary).  PCG (mgpcg_mex.c:
// mx_pcg) stops when ||r|| <= tol ||f|| or after maxit iterations; its
// preconditioner is mgopts.maxit V-cycles (mx_mg, mx_cycle): pre-smoothing
// on G and on the boundary band B (2^l nboundary sweeps on level l, from
// per-colour index lists of the band), the residual, full-weighting
// restriction onto the coarse mask, the coarse cycle, trilinear
// correction, the mirrored post-smoothing; on the
// coarsest level the reference sweeps to a relative residual of 1e-8.
// The grid hierarchy (the voxel size doubles per level), the masks (a coarse
// voxel is active iff its 8 fine children are: coarsen_grid_mex.c), the
// band (boundary_mask_mex.c) and the transfer weights (restrict_mex.c,
// correct_mex.c) are the reference's.  The smoother is red-black
// Gauss-Seidel (the reference's GAUSS_SEIDEL_RED_BLACK variant) with the
// reference's colour order in its reverse sweeps (gauss_seidel_mex.c:
// 141-163: the first colour there is the one of parity nx + 1).  Two
// choices differ from the reference and change only the preconditioner,
// hence how fast PCG converges, not what it converges to: every voxel of a
// colour is updated (the reference's loops skip the row end of the second
// colour on rows with j = k mod 2, which on a mask that touches the
// volume's faces makes the V-cycle a poor preconditioner), and the coarsest
// level takes a fixed-degree Chebyshev iteration on the Jacobi-scaled
// operator instead of rounds of 256 Gauss-Seidel sweeps.  Its degree
// reduces the coarse error by 1e3 over the spectrum bound of the level's
// box (the masked operator is a principal submatrix of the box's, so the
// bound holds for any mask); each step is one stencil pass over the level
// by many workgroups, with no reduction, and a fixed polynomial keeps the
// V-cycle a fixed linear map, as PCG requires of its preconditioner.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

using namespace qsm;

namespace {

constexpr int BX = 64, BY = 4, NT = BX * BY;
constexpr int MAX_LEVELS = 16;    // at most 16 grid levels (mgpcg_mex.c: MAX_DEPTH)
constexpr int CHEB_MAX = 512;     // Chebyshev degree cap on the coarsest level
constexpr double CHEB_EPS = 1e-3; // coarse error reduction it aims for
constexpr int DOT_BLOCKS = 1024;  // fixed grid of the dot products

struct Dim {
    int nx, ny, nz;
    __host__ __device__ size_t n() const { return size_t(nx) * ny * nz; }
};

dim3 grid3(Dim d) { return dim3((d.nx + BX - 1) / BX, (d.ny + BY - 1) / BY, d.nz); }

// one lane per voxel of one colour: (nx + 1) / 2 lanes per row
dim3 grid_half(Dim d) { return dim3(((d.nx + 1) / 2 + BX - 1) / BX, (d.ny + BY - 1) / BY, d.nz); }

__device__ __forceinline__ bool vox(Dim d, int &i, int &j, int &k, size_t &l)
{
    int bx, by, bz;
    xcd_block3(bx, by, bz);  // a slab of planes per XCD
    i = bx * BX + threadIdx.x;
    j = by * BY + threadIdx.y;
    k = bz;
    l = size_t(i) + size_t(d.nx) * (size_t(j) + size_t(d.ny) * k);
    return i < d.nx && j < d.ny;
}

__device__ __forceinline__ bool interior(Dim d, int i, int j, int k)
{
    return i > 0 && j > 0 && k > 0 && i < d.nx - 1 && j < d.ny - 1 && k < d.nz - 1;
}

// Laplacian coefficients as lapmg_mex.c forms them (in double): 1 / h^2 per
// axis, h the voxel size times 2^level (mgpcg_mex.c: mx_init).
struct Lap {
    double hx, hy, hz;
};

Lap lap_at(const double vsz[3], int level)
{
    const double s = std::ldexp(1.0, level);
    const double h[3] = {vsz[0] * s, vsz[1] * s, vsz[2] * s};
    return Lap{1.0 / (h[0] * h[0]), 1.0 / (h[1] * h[1]), 1.0 / (h[2] * h[2])};
}

// ---- padding -------------------------------------------------------------

// out (nx+2, ny+2, nz+2) = in with a zero border; G likewise from a mask.
template <typename T>
__global__ __launch_bounds__(NT) void pad_kernel(T *out, const T *in, uint8_t *G,
                                                 const uint8_t *m, Dim dp, T sgn_mask)
{
    int i, j, k;
    size_t l;
    if (!vox(dp, i, j, k, l)) return;
    const bool in_box = interior(dp, i, j, k);
    const size_t li = size_t(i - 1) + size_t(dp.nx - 2) * (size_t(j - 1) + size_t(dp.ny - 2) * (k - 1));
    if (out) {
        T v = T(0);
        // sgn_mask != 0: out = sgn_mask * mask .* in (unwrapLaplacian's -m.*d2u)
        if (in_box) v = (sgn_mask != T(0)) ? sgn_mask * (T(m[li] ? 1 : 0) * in[li]) : in[li];
        out[l] = v;
    }
    if (G) G[l] = in_box ? (m[li] ? 1 : 0) : 0;
}

template <typename T>
__global__ __launch_bounds__(NT) void unpad_kernel(T *out, const T *in, Dim dp)
{
    int i, j, k;
    size_t l;
    if (!vox(dp, i, j, k, l)) return;
    if (!interior(dp, i, j, k)) return;
    out[size_t(i - 1) + size_t(dp.nx - 2) * (size_t(j - 1) + size_t(dp.ny - 2) * (k - 1))] = in[l];
}

// ---- masks of the hierarchy ----------------------------------------------

// coarsen_grid_mex.c: coarse (i2, j2, k2) is active iff the fine voxel
// (2 i2 - 1, ...) and its +1 neighbours in x, y, z (those inside the
// interior) are all active.
__global__ __launch_bounds__(NT) void coarsen_kernel(uint8_t *G2, const uint8_t *G, Dim d2, Dim d)
{
    int i2, j2, k2;
    size_t l2;
    if (!vox(d2, i2, j2, k2, l2)) return;
    uint8_t g = 0;
    if (interior(d2, i2, j2, k2)) {
        const int i = 2 * i2 - 1, j = 2 * j2 - 1, k = 2 * k2 - 1;
        const size_t sx = 1, sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
        const size_t l = size_t(i) + sy * (size_t(j) + size_t(d.ny) * k);
        const bool bi = !(i < d.nx - 2), bj = !(j < d.ny - 2), bk = !(k < d.nz - 2);
        g = G[l] && (bi || G[l + sx]) && (bj || G[l + sy]) && (bk || G[l + sz]) &&
            (bi || bj || G[l + sx + sy]) && (bi || bk || G[l + sx + sz]) &&
            (bj || bk || G[l + sy + sz]) && (bi || bj || bk || G[l + sx + sy + sz]);
    }
    G2[l2] = g;
}

// boundary_mask_mex.c, first two steps: b = G on the interior's outer
// layer, and on every 2x2x2 block (starting at odd indices) that is not
// fully active.
__global__ __launch_bounds__(NT) void band_seed_kernel(uint8_t *b, const uint8_t *G, Dim d)
{
    int i, j, k;
    size_t l;
    if (!vox(d, i, j, k, l)) return;
    uint8_t v = 0;
    if (interior(d, i, j, k)) {
        if (i == 1 || j == 1 || k == 1 || i == d.nx - 2 || j == d.ny - 2 || k == d.nz - 2)
            v = G[l];
        // the block that contains (i, j, k): its corner at odd indices
        const int ic = i - ((i - 1) & 1), jc = j - ((j - 1) & 1), kc = k - ((k - 1) & 1);
        const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
        const size_t lc = size_t(ic) + sy * (size_t(jc) + size_t(d.ny) * kc);
        // the reference's block loop visits corners up to the interior's last
        // index and reads the +1 neighbours, which may be the zero border
        bool full = G[lc] && G[lc + 1] && G[lc + sy] && G[lc + sz] && G[lc + 1 + sy] &&
                    G[lc + 1 + sz] && G[lc + sy + sz] && G[lc + 1 + sy + sz];
        if (!full) v = G[l];
    }
    b[l] = v;
}

// boundary_mask_mex.c, last step: B = G .* (b or any of its 6 neighbours).
__global__ __launch_bounds__(NT) void band_kernel(uint8_t *B, const uint8_t *b, const uint8_t *G,
                                                  Dim d)
{
    int i, j, k;
    size_t l;
    if (!vox(d, i, j, k, l)) return;
    uint8_t v = 0;
    if (interior(d, i, j, k) && G[l]) {
        const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
        v = b[l - sz] || b[l - sy] || b[l - 1] || b[l] || b[l + 1] || b[l + sy] || b[l + sz];
    }
    B[l] = v;
}

// ---- the band as index lists ---------------------------------------------
//
// The band is a thin shell, but a sweep over its mask visits the whole
// grid.  Each level's band voxels are listed once per solve, per colour, in
// increasing index order: a count per row (one wave per row of the grid),
// an exclusive scan over the rows, a fill per row.  A band sweep then takes
// one lane per listed voxel.  Voxels of one colour do not read each other,
// so the order within a colour does not change the result.

// One wave per row (j, k): the row's band voxels of colour 0 and 1, as
// cnt[2 row + c].
__global__ __launch_bounds__(NT) void band_count_kernel(int *cnt, const uint8_t *B, Dim d)
{
    const int lane = int(threadIdx.x) & 63;
    const int row = int(blockIdx.x) * (NT / 64) + int(threadIdx.x) / 64;
    if (row >= d.ny * d.nz) return;  // whole waves
    const int j = row % d.ny, k = row / d.ny;
    const uint8_t *b = B + size_t(d.nx) * row;
    int c0 = 0, c1 = 0;
    for (int x0 = 0; x0 < d.nx; x0 += 64) {
        const int i = x0 + lane;
        const bool on = i < d.nx && b[i];
        const bool odd = ((i + j + k) & 1) != 0;
        c0 += __popcll(__ballot(on && !odd));
        c1 += __popcll(__ballot(on && odd));
    }
    if (lane == 0) {
        cnt[2 * row] = c0;
        cnt[2 * row + 1] = c1;
    }
}

// One block: cnt[2 row + c] becomes the exclusive prefix over rows of
// colour c; tot[0..1] the totals.
__global__ __launch_bounds__(1024) void band_scan_kernel(int *cnt, int rows, int *tot)
{
    __shared__ int part[2][1024];
    const int t = int(threadIdx.x), per = (rows + 1023) / 1024;
    const int r0 = min(rows, t * per), r1 = min(rows, r0 + per);
    int s0 = 0, s1 = 0;
    for (int r = r0; r < r1; ++r) {
        s0 += cnt[2 * r];
        s1 += cnt[2 * r + 1];
    }
    part[0][t] = s0;
    part[1][t] = s1;
    __syncthreads();
    // inclusive scan of the 1024 thread sums (Hillis-Steele)
    for (int o = 1; o < 1024; o <<= 1) {
        const int a0 = t >= o ? part[0][t - o] : 0, a1 = t >= o ? part[1][t - o] : 0;
        __syncthreads();
        part[0][t] += a0;
        part[1][t] += a1;
        __syncthreads();
    }
    int e0 = part[0][t] - s0, e1 = part[1][t] - s1;
    for (int r = r0; r < r1; ++r) {
        const int n0 = cnt[2 * r], n1 = cnt[2 * r + 1];
        cnt[2 * r] = e0;
        cnt[2 * r + 1] = e1;
        e0 += n0;
        e1 += n1;
    }
    if (t == 1023) {
        tot[0] = part[0][1023];
        tot[1] = part[1][1023];
    }
}

// One wave per row: the row's band voxels of colour c go to list[c] from
// the row's offset, in increasing i.
__global__ __launch_bounds__(NT) void band_fill_kernel(int *list0, int *list1, const int *off,
                                                       const uint8_t *B, Dim d)
{
    const int lane = int(threadIdx.x) & 63;
    const int row = int(blockIdx.x) * (NT / 64) + int(threadIdx.x) / 64;
    if (row >= d.ny * d.nz) return;  // whole waves
    const int j = row % d.ny, k = row / d.ny;
    const size_t base = size_t(d.nx) * row;
    int o0 = off[2 * row], o1 = off[2 * row + 1];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int x0 = 0; x0 < d.nx; x0 += 64) {
        const int i = x0 + lane;
        const bool on = i < d.nx && B[base + i];
        const bool odd = ((i + j + k) & 1) != 0;
        const unsigned long long m0 = __ballot(on && !odd), m1 = __ballot(on && odd);
        if (on && !odd) list0[o0 + __popcll(m0 & below)] = int(base + i);
        if (on && odd) list1[o1 + __popcll(m1 & below)] = int(base + i);
        o0 += __popcll(m0);
        o1 += __popcll(m1);
    }
}

// ---- smoothing, residual, transfers --------------------------------------

// One colour of a red-black Gauss-Seidel sweep (gauss_seidel_mex.c):
// v = (f + hx (v-x + v+x) + hy (...) + hz (...)) / (2 (hx + hy + hz)) on the
// voxels of M with (i + j + k) % 2 == colour.  Lanes map onto the
// colour's voxels only: lane t of a row takes i = 2 t + ((colour + j + k)
// & 1), so no lane idles on the other colour (grid: grid_half).
template <typename T>
__global__ __launch_bounds__(NT) void gs_kernel(T *v, const T *f, const uint8_t *M, Dim d, Lap c,
                                                int colour)
{
    int bx, by, bz;
    xcd_block3(bx, by, bz);  // a slab of planes per XCD: the z neighbours meet in its L2
    const int t = bx * BX + threadIdx.x;
    const int j = by * BY + threadIdx.y;
    const int k = bz;
    if (j >= d.ny) return;
    const int i = 2 * t + ((colour + j + k) & 1);
    if (i >= d.nx || !interior(d, i, j, k)) return;
    const size_t l = size_t(i) + size_t(d.nx) * (size_t(j) + size_t(d.ny) * k);
    if (!M[l]) return;
    const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
    const T hh = T(1.0 / (2.0 * (c.hx + c.hy + c.hz)));
    v[l] = hh * (f[l] + ((T(c.hx) * (v[l - 1] + v[l + 1]) + T(c.hy) * (v[l - sy] + v[l + sy])) +
                         T(c.hz) * (v[l - sz] + v[l + sz])));
}

// gs_kernel on the n voxels of one colour listed in idx (interior voxels of
// the band): the same update, one lane per listed voxel.
template <typename T>
__global__ __launch_bounds__(NT) void gs_list_kernel(T *v, const T *f, const int *idx, int n,
                                                     Dim d, Lap c)
{
    const int t = int(blockIdx.x) * NT + int(threadIdx.x);
    if (t >= n) return;
    const size_t l = size_t(idx[t]);
    const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
    const T hh = T(1.0 / (2.0 * (c.hx + c.hy + c.hz)));
    v[l] = hh * (f[l] + ((T(c.hx) * (v[l - 1] + v[l + 1]) + T(c.hy) * (v[l - sy] + v[l + sy])) +
                         T(c.hz) * (v[l - sz] + v[l + sz])));
}

// residual_mex.c: r = f + Lap_h v on G, 0 elsewhere.
template <typename T>
__global__ __launch_bounds__(NT) void residual_kernel(T *r, const T *f, const T *v,
                                                      const uint8_t *G, Dim d, Lap c)
{
    int i, j, k;
    size_t l;
    if (!vox(d, i, j, k, l)) return;
    T out = T(0);
    if (interior(d, i, j, k) && G[l]) {
        const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
        const T hh = T(-2.0 * (c.hx + c.hy + c.hz));
        out = f[l] + (((hh * v[l] + T(c.hx) * (v[l - 1] + v[l + 1])) +
                       T(c.hy) * (v[l - sy] + v[l + sy])) + T(c.hz) * (v[l - sz] + v[l + sz]));
    }
    r[l] = out;
}

// lapmg_mex.c with the sign of A: q = -Lap_h p on G, 0 elsewhere.
template <typename T>
__global__ __launch_bounds__(NT) void apply_kernel(T *q, const T *p, const uint8_t *G, Dim d,
                                                   Lap c)
{
    int i, j, k;
    size_t l;
    if (!vox(d, i, j, k, l)) return;
    T out = T(0);
    if (interior(d, i, j, k) && G[l]) {
        const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
        const T hh = T(-2.0 * (c.hx + c.hy + c.hz));
        out = -((((hh * p[l] + T(c.hx) * (p[l - 1] + p[l + 1])) + T(c.hy) * (p[l - sy] + p[l + sy])) +
                 T(c.hz) * (p[l - sz] + p[l + sz])));
    }
    q[l] = out;
}

// restrict_mex.c: full weighting (1 2 1 / 2 4 2 / 1 2 1 per plane, x2 in
// the middle plane, / 64) around the fine (2 i2 - 1, ...), onto G2; v2 = 0.
template <typename T>
__global__ __launch_bounds__(NT) void restrict_kernel(T *f2, T *v2, const T *r, const uint8_t *G2,
                                                      Dim d2, Dim d)
{
    int i2, j2, k2;
    size_t l2;
    if (!vox(d2, i2, j2, k2, l2)) return;
    v2[l2] = T(0);
    T out = T(0);
    if (interior(d2, i2, j2, k2) && G2[l2]) {
        const int i = 2 * i2 - 1, j = 2 * j2 - 1, k = 2 * k2 - 1;
        const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
        const size_t l = size_t(i) + sy * (size_t(j) + size_t(d.ny) * k);
        double s = 0.0;
        for (int dk = -1; dk <= 1; ++dk)
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    const int w = (2 - (di < 0 ? -di : di)) * (2 - (dj < 0 ? -dj : dj)) *
                                  (2 - (dk < 0 ? -dk : dk));
                    s += double(w) * double(r[l + di + ptrdiff_t(dj) * ptrdiff_t(sy) +
                                              ptrdiff_t(dk) * ptrdiff_t(sz)]);
                }
        out = T(s / 64.0);
    }
    f2[l2] = out;
}

// correct_mex.c: v += trilinear interpolation of v2 on G (each coarse
// voxel serves the 2x2x2 fine block at (2 i2 - 1, ...)).
template <typename T>
__global__ __launch_bounds__(NT) void correct_kernel(T *v, const T *v2, const uint8_t *G, Dim d,
                                                     Dim d2)
{
    int i2, j2, k2;
    size_t l2;
    if (!vox(d2, i2, j2, k2, l2)) return;
    if (!interior(d2, i2, j2, k2)) return;
    const size_t sy2 = size_t(d2.nx), sz2 = size_t(d2.nx) * d2.ny;
    const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
    const int i = 2 * i2 - 1, j = 2 * j2 - 1, k = 2 * k2 - 1;
    const size_t l = size_t(i) + sy * (size_t(j) + size_t(d.ny) * k);
    for (int dk = 0; dk <= 1; ++dk)
        for (int dj = 0; dj <= 1; ++dj)
            for (int di = 0; di <= 1; ++di) {
                const int fi = i + di, fj = j + dj, fk = k + dk;
                if (!interior(d, fi, fj, fk)) continue;
                const size_t lf = l + di + dj * sy + dk * sz;
                if (!G[lf]) continue;
                // mean of the coarse values at the corners this fine voxel
                // lies between (1, 2, 4 or 8 of them)
                double s = 0.0;
                for (int ek = 0; ek <= dk; ++ek)
                    for (int ej = 0; ej <= dj; ++ej)
                        for (int ei = 0; ei <= di; ++ei)
                            s += double(v2[l2 + ei + ej * sy2 + ek * sz2]);
                const int cnt = (1 + di) * (1 + dj) * (1 + dk);
                v[lf] += T(s / double(cnt));
            }
}

// Coarsest level, Chebyshev iteration on hh A (A = -Lap_h, hh = 1 /
// (2 (hx + hy + hz)), spectrum in [lmin, lmax]; Saad, Iterative Methods for
// Sparse Linear Systems, 2nd ed., Algorithm 12.1).  Start: r = f - A v on
// the active voxels, d = s0 hh r with s0 = 1 / theta.
template <typename T>
__global__ __launch_bounds__(NT) void cheb_start_kernel(T *r, T *dn, const T *v, const T *f,
                                                        const uint8_t *G, Dim d, Lap c, T s0)
{
    int i, j, k;
    size_t l;
    if (!vox(d, i, j, k, l)) return;
    T rl = T(0);
    if (interior(d, i, j, k) && G[l]) {
        const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
        const T hh = T(-2.0 * (c.hx + c.hy + c.hz));
        rl = f[l] + (((hh * v[l] + T(c.hx) * (v[l - 1] + v[l + 1])) + T(c.hy) * (v[l - sy] + v[l + sy])) +
                     T(c.hz) * (v[l - sz] + v[l + sz]));
    }
    r[l] = rl;
    dn[l] = s0 * T(1.0 / (2.0 * (c.hx + c.hy + c.hz))) * rl;
}

// One step: v += d; r -= A d; dn = c1 d + c2 hh r (active voxels; dn = 0
// elsewhere).  d is read with its neighbours, so dn is another buffer.
template <typename T>
__global__ __launch_bounds__(NT) void cheb_step_kernel(T *v, T *r, const T *dd, T *dn,
                                                       const uint8_t *G, Dim d, Lap c, T c1, T c2)
{
    int i, j, k;
    size_t l;
    if (!vox(d, i, j, k, l)) return;
    T out = T(0);
    if (interior(d, i, j, k) && G[l]) {
        const size_t sy = size_t(d.nx), sz = size_t(d.nx) * d.ny;
        const T hh = T(-2.0 * (c.hx + c.hy + c.hz));
        const T ad = -(((hh * dd[l] + T(c.hx) * (dd[l - 1] + dd[l + 1])) + T(c.hy) * (dd[l - sy] + dd[l + sy])) +
                       T(c.hz) * (dd[l - sz] + dd[l + sz]));
        v[l] = v[l] + dd[l];
        const T rl = r[l] - ad;
        r[l] = rl;
        out = c1 * dd[l] + c2 * T(1.0 / (2.0 * (c.hx + c.hy + c.hz))) * rl;
    }
    dn[l] = out;
}

// ---- PCG vector kernels ----------------------------------------------------

template <typename T>
__global__ __launch_bounds__(NT) void zero_kernel(T *a, size_t n)
{
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l < n) a[l] = T(0);
}

// r -= q
template <typename T>
__global__ __launch_bounds__(NT) void sub_kernel(T *r, const T *q, size_t n)
{
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l < n) r[l] = r[l] - q[l];
}

// partial sums of a[l] * b[l] over a fixed grid of DOT_BLOCKS blocks (each
// thread strides the array; fixed order, so the result is reproducible)
template <typename T>
__global__ __launch_bounds__(NT) void dot_kernel(const T *a, const T *b, size_t n, double *part)
{
    __shared__ double red[NT];
    double acc = 0.0;
    for (size_t l = size_t(blockIdx.x) * NT + threadIdx.x; l < n; l += size_t(DOT_BLOCKS) * NT)
        acc += double(a[l]) * double(b[l]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(NT) void sum_kernel(const double *part, int nb, double *out)
{
    __shared__ double red[NT];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += NT) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = NT / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// The PCG loop's scalars on the device (mgpcg_mex.c: mx_pcg), so that an
// iteration reads back only ||r||^2 and the breakdown flag.
struct Pcg {
    double rho, prevrho, pq, rr;
    int brk;  // rho or pq not positive: the reference's loop breaks there
    int pad;
};

// sum_kernel's reduction into the PCG scalars.  WHAT 0: rr; 1: rho, the
// previous rho kept as prevrho, breakdown when !(rho > 0); 2: pq,
// breakdown when !(pq > 0).
template <int WHAT>
__global__ __launch_bounds__(NT) void pcg_sum_kernel(const double *part, int nb, Pcg *g)
{
    __shared__ double red[NT];
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += NT) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = NT / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double v = red[0];
    if (WHAT == 0) {
        g->rr = v;
    } else if (WHAT == 1) {
        g->prevrho = g->rho;
        g->rho = v;
        if (!(v > 0.0)) g->brk = 1;
    } else {
        g->pq = v;
        if (!(v > 0.0)) g->brk = 1;
    }
}

// p = q + beta p with beta = rho / prevrho (skipped after a breakdown)
template <typename T>
__global__ __launch_bounds__(NT) void xpby_pcg_kernel(T *p, const T *q, const Pcg *g, size_t n)
{
    if (g->brk) return;
    const T beta = T(g->rho / g->prevrho);
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l < n) p[l] = beta * p[l] + q[l];
}

// x += a p; r -= a q with a = rho / pq (skipped after a breakdown)
template <typename T>
__global__ __launch_bounds__(NT) void update_pcg_kernel(T *x, T *r, const T *p, const T *q,
                                                        const Pcg *g, size_t n)
{
    if (g->brk) return;
    const T a = T(g->rho / g->pq);
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l < n) {
        x[l] = x[l] + a * p[l];
        r[l] = r[l] - a * q[l];
    }
}

unsigned g1(size_t n) { return unsigned((n + NT - 1) / NT); }

size_t al(size_t n) { return (n + 255) / 256 * 256; }

// ---- the solver --------------------------------------------------------------

template <typename T>
struct Level {
    Dim d;
    Lap c;
    T *v, *f, *r;
    uint8_t *G, *B;
    int *bl;    // band voxels of colour 0 from bl, of colour 1 from bl + (n + 1) / 2
    int bn[2];  // their counts (host copy)
};

template <typename T>
struct Solver {
    int L;
    Level<T> lv[MAX_LEVELS];
    T *x, *p;           // PCG iterate and direction (level-0 size)
    T *cp, *cq;         // coarsest-level CG scratch
    uint8_t *tmp;       // band seed scratch (level-0 size)
    int *rowcnt, *btot; // band list construction: per-row counts, totals per level
    double *part, *sum; // reductions
    Pcg *pcg;           // the PCG scalars
};

// Level sizes: the padded grid, then ((n - 2 + 1) >> 1) + 2 per level
// (mgpcg_mex.c: mx_init).
int level_dims(const size_t sz[3], int L, Dim *d)
{
    d[0] = Dim{int(sz[0]) + 2, int(sz[1]) + 2, int(sz[2]) + 2};
    for (int i = 1; i < L; ++i)
        d[i] = Dim{((d[i - 1].nx - 2 + 1) >> 1) + 2, ((d[i - 1].ny - 2 + 1) >> 1) + 2,
                   ((d[i - 1].nz - 2 + 1) >> 1) + 2};
    return QSM_OK;
}

template <typename T>
size_t solver_layout(const size_t sz[3], const double vsz[3], int L, Solver<T> *s, char *base)
{
    Dim d[MAX_LEVELS];
    level_dims(sz, L, d);
    size_t off = 0;
    auto take = [&](size_t b) { char *p = base ? base + off : nullptr; off += al(b); return p; };
    for (int i = 0; i < L; ++i) {
        const size_t n = d[i].n();
        char *pv = take(n * sizeof(T)), *pf = take(n * sizeof(T)), *pr = take(n * sizeof(T));
        char *pG = take(n), *pB = take(n);
        // band lists on the levels that smooth (not the coarsest)
        char *pl = i < L - 1 ? take((n + 1) * sizeof(int)) : nullptr;
        if (s)
            s->lv[i] = Level<T>{d[i], lap_at(vsz, i), reinterpret_cast<T *>(pv),
                                reinterpret_cast<T *>(pf), reinterpret_cast<T *>(pr),
                                reinterpret_cast<uint8_t *>(pG), reinterpret_cast<uint8_t *>(pB),
                                reinterpret_cast<int *>(pl), {0, 0}};
    }
    const size_t n0 = d[0].n();
    char *px = take(n0 * sizeof(T)), *pp = take(n0 * sizeof(T)), *pt = take(n0);
    char *prc = take(2 * size_t(d[0].ny) * d[0].nz * sizeof(int));
    char *pbt = take(2 * MAX_LEVELS * sizeof(int));
    const size_t nc = d[L - 1].n();
    char *pcp = take(nc * sizeof(T)), *pcq = take(nc * sizeof(T));
    char *ppart = take(DOT_BLOCKS * sizeof(double)), *psum = take(sizeof(double));
    char *ppcg = take(sizeof(Pcg));
    if (s) {
        s->L = L;
        s->x = reinterpret_cast<T *>(px);
        s->p = reinterpret_cast<T *>(pp);
        s->cp = reinterpret_cast<T *>(pcp);
        s->cq = reinterpret_cast<T *>(pcq);
        s->tmp = reinterpret_cast<uint8_t *>(pt);
        s->rowcnt = reinterpret_cast<int *>(prc);
        s->btot = reinterpret_cast<int *>(pbt);
        s->part = reinterpret_cast<double *>(ppart);
        s->sum = reinterpret_cast<double *>(psum);
        s->pcg = reinterpret_cast<Pcg *>(ppcg);
    }
    return off;
}

template <typename T>
int dot(const T *a, const T *b, size_t n, Solver<T> &s, double &out, hipStream_t st)
{
    dot_kernel<T><<<DOT_BLOCKS, NT, 0, st>>>(a, b, n, s.part);
    QSM_HIP(hipGetLastError());
    sum_kernel<<<1, NT, 0, st>>>(s.part, DOT_BLOCKS, s.sum);
    QSM_HIP(hipGetLastError());
    QSM_HIP(hipMemcpyAsync(&out, s.sum, sizeof(double), hipMemcpyDeviceToHost, st));
    QSM_HIP(hipStreamSynchronize(st));
    return QSM_OK;
}

// n red-black sweeps on mask M.  Forward sweeps take the even colour
// ((i + j + k) even) first; reverse sweeps start from the volume's far
// corner, whose colour is that of parity nx + 1 (gauss_seidel_mex.c:141-163).
template <typename T>
int smooth(Level<T> &L, const uint8_t *M, int n, bool reverse, hipStream_t st)
{
    const dim3 g = grid_half(L.d), b(BX, BY);
    const int first = reverse ? ((L.d.nx + 1) & 1) : 0;
    for (int s = 0; s < n; ++s)
        for (int h = 0; h < 2; ++h) {
            gs_kernel<T><<<g, b, 0, st>>>(L.v, L.f, M, L.d, L.c, first ^ h);
            QSM_HIP(hipGetLastError());
        }
    return QSM_OK;
}

// n red-black sweeps on the level's band, from its index lists; the colour
// order of smooth().
template <typename T>
int smooth_band(Level<T> &L, int n, bool reverse, hipStream_t st)
{
    const int first = reverse ? ((L.d.nx + 1) & 1) : 0;
    const int *list[2] = {L.bl, L.bl + (L.d.n() + 1) / 2};
    for (int s = 0; s < n; ++s)
        for (int h = 0; h < 2; ++h) {
            const int c = first ^ h;
            if (L.bn[c] == 0) continue;
            gs_list_kernel<T><<<(L.bn[c] + NT - 1) / NT, NT, 0, st>>>(L.v, L.f, list[c], L.bn[c],
                                                                     L.d, L.c);
            QSM_HIP(hipGetLastError());
        }
    return QSM_OK;
}

// Spectrum bounds of hh A on the level's box (interior nx-2 x ny-2 x nz-2,
// Dirichlet): lmin = hh sum_a 2 h_a (1 - cos(pi / (m_a + 1))), lmax = 2 -
// lmin; the Chebyshev degree that reduces the error by CHEB_EPS over them.
struct Cheb {
    double lmin, lmax;
    int K;
};

Cheb cheb_of(const Dim &d, const Lap &c)
{
    const double pi = 3.14159265358979323846;
    const double hh = 1.0 / (2.0 * (c.hx + c.hy + c.hz));
    const double w[3] = {c.hx, c.hy, c.hz};
    const int m[3] = {d.nx - 2, d.ny - 2, d.nz - 2};
    double lmin = 0.0;
    for (int a = 0; a < 3; ++a) lmin += hh * 2.0 * w[a] * (1.0 - std::cos(pi / (m[a] + 1)));
    const double lmax = 2.0 - lmin;
    const double sigma = (lmax + lmin) / (lmax - lmin);
    const int K = int(std::ceil(std::acosh(1.0 / CHEB_EPS) / std::acosh(sigma)));
    return Cheb{lmin, lmax, std::max(1, std::min(K, CHEB_MAX))};
}

// The coarsest level: K Chebyshev steps from the current v (restriction
// zeroed it, except on a one-level hierarchy's later cycles).
template <typename T>
int coarse_solve(Solver<T> &s, Level<T> &L, hipStream_t st)
{
    const Cheb ch = cheb_of(L.d, L.c);
    const double theta = 0.5 * (ch.lmax + ch.lmin), delta = 0.5 * (ch.lmax - ch.lmin);
    const double sigma = theta / delta;
    const dim3 g = grid3(L.d), b(BX, BY);
    T *d0 = s.cp, *d1 = s.cq;
    cheb_start_kernel<T><<<g, b, 0, st>>>(L.r, d0, L.v, L.f, L.G, L.d, L.c, T(1.0 / theta));
    QSM_HIP(hipGetLastError());
    double rho = 1.0 / sigma;
    for (int k = 0; k < ch.K; ++k) {
        const double rho1 = 1.0 / (2.0 * sigma - rho);
        cheb_step_kernel<T><<<g, b, 0, st>>>(L.v, L.r, d0, d1, L.G, L.d, L.c, T(rho1 * rho),
                                             T(2.0 * rho1 / delta));
        QSM_HIP(hipGetLastError());
        rho = rho1;
        std::swap(d0, d1);
    }
    return QSM_OK;
}

struct MgParams {
    double tol;
    int cycles, mu, npre, npost, nboundary;
};

// mgpcg_mex.c: mx_cycle on level l.
template <typename T>
int cycle(Solver<T> &s, int l, const MgParams &m, hipStream_t st)
{
    int rc;
    Level<T> &L = s.lv[l];
    const dim3 b(BX, BY);
    if (l < s.L - 1) {
        const int bs = (1 << l) * m.nboundary;
        if ((rc = smooth(L, L.G, m.npre, false, st))) return rc;
        if ((rc = smooth_band(L, bs, false, st))) return rc;
        residual_kernel<T><<<grid3(L.d), b, 0, st>>>(L.r, L.f, L.v, L.G, L.d, L.c);
        QSM_HIP(hipGetLastError());
        Level<T> &C = s.lv[l + 1];
        restrict_kernel<T><<<grid3(C.d), b, 0, st>>>(C.f, C.v, L.r, C.G, C.d, L.d);
        QSM_HIP(hipGetLastError());
        for (int i = 0; i < m.mu; ++i) {
            if ((rc = cycle(s, l + 1, m, st))) return rc;
            if (l + 1 == s.L - 1) break;
        }
        correct_kernel<T><<<grid3(C.d), b, 0, st>>>(L.v, C.v, L.G, L.d, C.d);
        QSM_HIP(hipGetLastError());
        if ((rc = smooth_band(L, bs, true, st))) return rc;
        if ((rc = smooth(L, L.G, m.npost, true, st))) return rc;
    } else {
        if ((rc = coarse_solve(s, L, st))) return rc;
    }
    return QSM_OK;
}

// Solves -Lap v = f on the mask (sz, unpadded, device), v out (sz).  src:
// f itself, or with unwrap != 0 the wrapped Laplacian d2u whose right-hand
// side is -m .* d2u (unwrapLaplacian.m:84).
//
// x0 (may be null): the initial guess (sz, device); PCG then starts from
// r = f - A x0, with the stopping rule still relative to ||f||
// (mgpcg_mex.c: mx_pcg).
template <typename T>
int mgpcg_t(const T *src, bool unwrap, const uint8_t *mask, const size_t sz[3],
            const double vsz[3], double tol, int maxit, const MgParams &m, int nlevels, T *v,
            void *work, int *iters, double *resnorm, hipStream_t st, const T *x0 = nullptr)
{
    int rc;
    Solver<T> s;
    solver_layout<T>(sz, vsz, nlevels, &s, static_cast<char *>(work));
    Level<T> &L0 = s.lv[0];
    const size_t n0 = L0.d.n();
    const dim3 b(BX, BY);
    // level 0: r (held in f) = padded right-hand side, G; masks of the hierarchy
    pad_kernel<T><<<grid3(L0.d), b, 0, st>>>(L0.f, src, L0.G, mask, L0.d, unwrap ? T(-1) : T(0));
    QSM_HIP(hipGetLastError());
    for (int l = 0; l < s.L; ++l) {
        Level<T> &L = s.lv[l];
        if (l > 0) {
            coarsen_kernel<<<grid3(L.d), b, 0, st>>>(L.G, s.lv[l - 1].G, L.d, s.lv[l - 1].d);
            QSM_HIP(hipGetLastError());
        }
        band_seed_kernel<<<grid3(L.d), b, 0, st>>>(s.tmp, L.G, L.d);
        band_kernel<<<grid3(L.d), b, 0, st>>>(L.B, s.tmp, L.G, L.d);
        QSM_HIP(hipGetLastError());
        if (l < s.L - 1) {
            // the band's index lists per colour
            const int rows = L.d.ny * L.d.nz, rb = (rows + NT / 64 - 1) / (NT / 64);
            band_count_kernel<<<rb, NT, 0, st>>>(s.rowcnt, L.B, L.d);
            band_scan_kernel<<<1, 1024, 0, st>>>(s.rowcnt, rows, s.btot + 2 * l);
            band_fill_kernel<<<rb, NT, 0, st>>>(L.bl, L.bl + (L.d.n() + 1) / 2, s.rowcnt, L.B,
                                                L.d);
            QSM_HIP(hipGetLastError());
        }
    }
    if (s.L > 1) {
        int tot[2 * MAX_LEVELS];
        QSM_HIP(hipMemcpyAsync(tot, s.btot, 2 * (s.L - 1) * sizeof(int), hipMemcpyDeviceToHost,
                               st));
        QSM_HIP(hipStreamSynchronize(st));
        for (int l = 0; l < s.L - 1; ++l) {
            s.lv[l].bn[0] = tot[2 * l];
            s.lv[l].bn[1] = tot[2 * l + 1];
        }
    }
    // PCG (mgpcg_mex.c: mx_pcg): reltol from ||f||, then r = f - A x0
    zero_kernel<T><<<g1(n0), NT, 0, st>>>(s.p, n0);
    QSM_HIP(hipGetLastError());
    T *r = L0.f, *q = L0.v;
    double res;
    if ((rc = dot(r, r, n0, s, res, st))) return rc;
    res = std::sqrt(res);
    const double reltol = tol * res;
    if (x0) {
        pad_kernel<T><<<grid3(L0.d), b, 0, st>>>(s.x, x0, nullptr, mask, L0.d, T(0));
        apply_kernel<T><<<grid3(L0.d), b, 0, st>>>(q, s.x, L0.G, L0.d, L0.c);
        sub_kernel<T><<<g1(n0), NT, 0, st>>>(r, q, n0);
        QSM_HIP(hipGetLastError());
        if ((rc = dot(r, r, n0, s, res, st))) return rc;
        res = std::sqrt(res);
    } else {
        zero_kernel<T><<<g1(n0), NT, 0, st>>>(s.x, n0);
        QSM_HIP(hipGetLastError());
    }
    // rho, prevrho, pq and ||r||^2 stay on the device; each iteration reads
    // back ||r||^2 and the breakdown flag once, at its end
    const Pcg g0{1.0, 1.0, 0.0, 0.0, 0, 0};  // rho = 1 before the first iteration
    QSM_HIP(hipMemcpyAsync(s.pcg, &g0, sizeof(Pcg), hipMemcpyHostToDevice, st));
    int it = 0;
    for (; it < maxit; ++it) {
        if (res <= reltol) break;
        // q = M \ r: m.cycles V-cycles from q = 0 (mgpcg_mex.c: mx_mg; with
        // mgopts.tol > 0 they stop once ||r - A q|| < mgopts.tol ||r||)
        zero_kernel<T><<<g1(n0), NT, 0, st>>>(q, n0);
        QSM_HIP(hipGetLastError());
        const double mgtol = m.tol > 0.0 ? m.tol * res : 0.0;
        for (int k = 0; k < m.cycles; ++k) {
            if (m.tol > 0.0) {
                residual_kernel<T><<<grid3(L0.d), b, 0, st>>>(L0.r, L0.f, L0.v, L0.G, L0.d, L0.c);
                QSM_HIP(hipGetLastError());
                double rn;
                if ((rc = dot(L0.r, L0.r, n0, s, rn, st))) return rc;
                if (std::sqrt(rn) < mgtol) break;
            }
            if ((rc = cycle(s, 0, m, st))) return rc;
        }
        // rho = r.q (prevrho = the last one); breakdown when !(rho > 0): the
        // residual is exactly zero on the mask (the rule above never stops
        // a right-hand side that is nonzero off the mask, since ||r|| counts
        // the whole array), so no direction is left
        dot_kernel<T><<<DOT_BLOCKS, NT, 0, st>>>(r, q, n0, s.part);
        pcg_sum_kernel<1><<<1, NT, 0, st>>>(s.part, DOT_BLOCKS, s.pcg);
        // p = q + (rho / prevrho) p; q = A p; pq = p.q (breakdown: !(pq > 0))
        xpby_pcg_kernel<T><<<g1(n0), NT, 0, st>>>(s.p, q, s.pcg, n0);
        apply_kernel<T><<<grid3(L0.d), b, 0, st>>>(q, s.p, L0.G, L0.d, L0.c);
        dot_kernel<T><<<DOT_BLOCKS, NT, 0, st>>>(s.p, q, n0, s.part);
        pcg_sum_kernel<2><<<1, NT, 0, st>>>(s.part, DOT_BLOCKS, s.pcg);
        // x += (rho / pq) p; r -= (rho / pq) q; ||r||^2
        update_pcg_kernel<T><<<g1(n0), NT, 0, st>>>(s.x, r, s.p, q, s.pcg, n0);
        dot_kernel<T><<<DOT_BLOCKS, NT, 0, st>>>(r, r, n0, s.part);
        pcg_sum_kernel<0><<<1, NT, 0, st>>>(s.part, DOT_BLOCKS, s.pcg);
        QSM_HIP(hipGetLastError());
        Pcg g;
        QSM_HIP(hipMemcpyAsync(&g, s.pcg, sizeof(Pcg), hipMemcpyDeviceToHost, st));
        QSM_HIP(hipStreamSynchronize(st));
        if (g.brk) break;  // x and r are as the reference leaves them
        res = std::sqrt(g.rr);
        if (resnorm) resnorm[it] = res;
    }
    if (iters) *iters = it;
    unpad_kernel<T><<<grid3(L0.d), b, 0, st>>>(v, s.x, L0.d);
    QSM_HIP(hipGetLastError());
    return QSM_OK;
}

int default_levels(const size_t sz[3])
{
    // mgpcg.m:51: max(1, floor(log2(min(size(f)))) - 3)
    const size_t mn = std::min(sz[0], std::min(sz[1], sz[2]));
    return std::max(1, int(std::floor(std::log2(double(mn)))) - 3);
}

int check_mg(int dtype, const size_t sz[3], const double vsz[3], const qsm_mg_opts *o, int &L)
{
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    QSM_ARG(sz && vsz && o, "null pointer");
    int rc = check_grid(sz, vsz, 1);
    if (rc) return rc;
    L = o->nlevels > 0 ? o->nlevels : default_levels(sz);
    QSM_ARG(L >= 1 && L <= MAX_LEVELS, "mgopts.nlevels %d out of range 1..%d", L, MAX_LEVELS);
    Dim d[MAX_LEVELS];
    level_dims(sz, L, d);
    QSM_ARG(d[L - 1].nx > 2 && d[L - 1].ny > 2 && d[L - 1].nz > 2,
            "mgopts.nlevels %d too many for this size", L);
    QSM_ARG(o->maxit >= 1 && o->mu >= 1 && o->npre >= 0 && o->npost >= 0 && o->nboundary >= 0 &&
            std::isfinite(o->tol), "invalid mgopts");
    return QSM_OK;
}

}  // namespace

extern "C" {

void qsm_mg_opts_default(qsm_mg_opts *o)
{
    // mgpcg.m:11-52
    o->tol = -1.0;
    o->maxit = 1;
    o->mu = 1;
    o->npre = 1;
    o->npost = 1;
    o->nboundary = 2;
    o->nlevels = 0;
}

size_t qsm_mgpcg_work_bytes(int dtype, const size_t sz[3], int nlevels)
{
    if (!sz) return 0;
    const int L = nlevels > 0 ? nlevels : default_levels(sz);
    if (L < 1 || L > MAX_LEVELS) return 0;
    const double one[3] = {1.0, 1.0, 1.0};
    return dtype == QSM_F32 ? solver_layout<float>(sz, one, L, nullptr, nullptr)
                            : solver_layout<double>(sz, one, L, nullptr, nullptr);
}

int qsm_mgpcg_dev(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
                  const double vsz[3], double tol, int maxit, const qsm_mg_opts *mgopts, void *v,
                  void *work, int *iters, double *resnorm, void *stream)
{
    int L, rc = check_mg(dtype, sz, vsz, mgopts, L);
    if (rc) return rc;
    QSM_ARG(f && mask && v && work, "null pointer");
    QSM_ARG(maxit >= 1 && std::isfinite(tol), "tol must be finite and maxit >= 1");
    const MgParams m{mgopts->tol, mgopts->maxit, mgopts->mu, mgopts->npre, mgopts->npost,
                     mgopts->nboundary};
    const hipStream_t st = as_stream(stream);
    return dtype == QSM_F64
        ? mgpcg_t<double>(static_cast<const double *>(f), false, mask, sz, vsz, tol, maxit, m, L,
                          static_cast<double *>(v), work, iters, resnorm, st)
        : mgpcg_t<float>(static_cast<const float *>(f), false, mask, sz, vsz, tol, maxit, m, L,
                         static_cast<float *>(v), work, iters, resnorm, st);
}

int qsm_mgpcg(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
              const double vsz[3], double tol, int maxit, const qsm_mg_opts *mgopts, void *v,
              int *iters, double *resnorm)
{
    int L, rc = check_mg(dtype, sz, vsz, mgopts, L);
    if (rc) return rc;
    QSM_ARG(f && mask && v, "null pointer");
    QSM_ARG(maxit >= 1 && std::isfinite(tol), "tol must be finite and maxit >= 1");
    if ((rc = ensure_device())) return rc;
    const size_t N = sz[0] * sz[1] * sz[2], es = dtype_size(dtype);
    DevBuf df, dm, dv, dw;
    if ((rc = df.alloc(N * es)) || (rc = dm.alloc(N)) || (rc = dv.alloc(N * es))
        || (rc = dw.alloc(qsm_mgpcg_work_bytes(dtype, sz, L))))
        return rc;
    QSM_HIP(hipMemcpy(df.p, f, N * es, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(dm.p, mask, N, hipMemcpyHostToDevice));
    if ((rc = qsm_mgpcg_dev(dtype, df.p, static_cast<const uint8_t *>(dm.p), sz, vsz, tol, maxit,
                            mgopts, dv.p, dw.p, iters, resnorm, nullptr)))
        return rc;
    QSM_HIP(hipStreamSynchronize(nullptr));
    QSM_HIP(hipMemcpy(v, dv.p, N * es, hipMemcpyDeviceToHost));
    return QSM_OK;
}

int qsm_mgpcg_x0(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
                 const double vsz[3], double tol, int maxit, const qsm_mg_opts *mgopts, void *v,
                 int *iters, double *resnorm)
{
    int L, rc = check_mg(dtype, sz, vsz, mgopts, L);
    if (rc) return rc;
    QSM_ARG(f && mask && v, "null pointer");
    QSM_ARG(maxit >= 0 && std::isfinite(tol), "tol must be finite and maxit >= 0");
    if ((rc = ensure_device())) return rc;
    const size_t N = sz[0] * sz[1] * sz[2], es = dtype_size(dtype);
    DevBuf df, dm, dv, dx, dw;
    if ((rc = df.alloc(N * es)) || (rc = dm.alloc(N)) || (rc = dv.alloc(N * es))
        || (rc = dx.alloc(N * es)) || (rc = dw.alloc(qsm_mgpcg_work_bytes(dtype, sz, L))))
        return rc;
    QSM_HIP(hipMemcpy(df.p, f, N * es, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(dm.p, mask, N, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(dx.p, v, N * es, hipMemcpyHostToDevice));
    const MgParams m{mgopts->tol, mgopts->maxit, mgopts->mu, mgopts->npre, mgopts->npost,
                     mgopts->nboundary};
    const hipStream_t st = nullptr;
    rc = dtype == QSM_F64
        ? mgpcg_t<double>(static_cast<const double *>(df.p), false,
                          static_cast<const uint8_t *>(dm.p), sz, vsz, tol, maxit, m, L,
                          static_cast<double *>(dv.p), dw.p, iters, resnorm, st,
                          static_cast<const double *>(dx.p))
        : mgpcg_t<float>(static_cast<const float *>(df.p), false,
                         static_cast<const uint8_t *>(dm.p), sz, vsz, tol, maxit, m, L,
                         static_cast<float *>(dv.p), dw.p, iters, resnorm, st,
                         static_cast<const float *>(dx.p));
    if (rc) return rc;
    QSM_HIP(hipStreamSynchronize(nullptr));
    QSM_HIP(hipMemcpy(v, dv.p, N * es, hipMemcpyDeviceToHost));
    return QSM_OK;
}

int qsm_unwrap_laplacian_dev(int dtype, const void *u, const uint8_t *mask, const size_t sz[3],
                             const double vsz[3], void *uphas, void *d2u, void *work,
                             int *iters, void *stream)
{
    // unwrapLaplacian.m:78-98 on an already cropped volume: d2u = lapw(u,
    // h), then mgpcg(-m .* d2u, m, h, 1e-8, ceil(sqrt(numel(m))), mgopts
    // with maxit = npre = npost = nboundary = 2)
    qsm_mg_opts o;
    qsm_mg_opts_default(&o);
    o.maxit = 2;
    o.npre = 2;
    o.npost = 2;
    o.nboundary = 2;
    int L, rc = check_mg(dtype, sz, vsz, &o, L);
    if (rc) return rc;
    QSM_ARG(u && mask && uphas && d2u && work, "null pointer");
    const size_t N = sz[0] * sz[1] * sz[2];
    const hipStream_t st = as_stream(stream);
    // lapw leaves the faces untouched: they are zero in the reference
    QSM_HIP(hipMemsetAsync(d2u, 0, N * dtype_size(dtype), st));
    if ((rc = qsm_lapw_dev(dtype, d2u, u, sz, vsz, stream))) return rc;
    const int maxit = int(std::ceil(std::sqrt(double(N))));
    const MgParams m{o.tol, o.maxit, o.mu, o.npre, o.npost, o.nboundary};
    return dtype == QSM_F64
        ? mgpcg_t<double>(static_cast<const double *>(d2u), true, mask, sz, vsz, 1e-8, maxit, m,
                          L, static_cast<double *>(uphas), work, iters, nullptr, st)
        : mgpcg_t<float>(static_cast<const float *>(d2u), true, mask, sz, vsz, 1e-8, maxit, m, L,
                         static_cast<float *>(uphas), work, iters, nullptr, st);
}

int qsm_unwrap_laplacian(int dtype, const void *phas, const uint8_t *mask, const size_t sz[3],
                         size_t nechoes, const double vsz[3], void *uphas, void *d2phas)
{
    // unwrapLaplacian.m:60-75: crop to the mask's box grown by one voxel
    // (clamped to the volume), per echo unwrap_, uncrop into zeros
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    QSM_ARG(phas && mask && sz && vsz && uphas && nechoes >= 1, "null pointer or no echo");
    int rc = check_grid(sz, vsz, 1);
    if (rc) return rc;
    const size_t es = dtype_size(dtype), N = sz[0] * sz[1] * sz[2];
    std::vector<char> any[3];
    axis_any(mask, sz, any);
    size_t c0[3], cs[3];
    for (int a = 0; a < 3; ++a) {
        size_t lo = 0, hi = sz[a];
        while (lo < sz[a] && !any[a][lo]) ++lo;
        while (hi > lo && !any[a][hi - 1]) --hi;
        if (lo == hi) {  // empty mask
            std::memset(uphas, 0, N * nechoes * es);
            if (d2phas) std::memset(d2phas, 0, N * nechoes * es);
            return QSM_OK;
        }
        c0[a] = lo > 0 ? lo - 1 : 0;
        cs[a] = std::min(sz[a] - 1, hi) - c0[a] + 1;  // through hi - 1 + 1, clamped
    }
    qsm_mg_opts o;
    qsm_mg_opts_default(&o);
    int L;
    if ((rc = check_mg(dtype, cs, vsz, &o, L))) return rc;
    if ((rc = ensure_device())) return rc;
    const size_t n = cs[0] * cs[1] * cs[2];
    // the crop and the uncrop run on the device: dP and dD hold a whole
    // echo (phase in, then the unwrapped phase; the Laplacian)
    std::vector<int> ix;
    for (int a = 0; a < 3; ++a)
        for (size_t i = 0; i < cs[a]; ++i) ix.push_back(int(c0[a] + i));
    DevBuf dP, dD, dM, dix, du, dm, dout, dd2, dw;
    if ((rc = dP.alloc(N * es)) || (rc = dM.alloc(N)) || (d2phas && (rc = dD.alloc(N * es))) ||
        (rc = dix.alloc(ix.size() * sizeof(int))) || (rc = du.alloc(n * es)) ||
        (rc = dm.alloc(n)) || (rc = dout.alloc(n * es)) || (rc = dd2.alloc(n * es)) ||
        (rc = dw.alloc(qsm_mgpcg_work_bytes(dtype, cs, L))))
        return rc;
    QSM_HIP(hipMemcpy(dM.p, mask, N, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(dix.p, ix.data(), ix.size() * sizeof(int), hipMemcpyHostToDevice));
    const int *dixp = static_cast<const int *>(dix.p);
    const CropMap cm{dixp, dixp + cs[0], dixp + cs[0] + cs[1], int(cs[0]), int(cs[1]),
                     int(cs[2]), sz[0], sz[1], N, n};
    crop_gather_kernel<uint8_t, true><<<crop_blocks(cm, 1), 256>>>(
        static_cast<uint8_t *>(dm.p), static_cast<const uint8_t *>(dM.p), cm, 1);
    QSM_HIP(hipGetLastError());
    // one echo's crop in and out (es bytes per sample)
    auto gather = [&](void *dst, const void *src) {
        if (dtype == QSM_F32)
            crop_gather_kernel<float><<<crop_blocks(cm, 1), 256>>>(
                static_cast<float *>(dst), static_cast<const float *>(src), cm, 1);
        else
            crop_gather_kernel<double><<<crop_blocks(cm, 1), 256>>>(
                static_cast<double *>(dst), static_cast<const double *>(src), cm, 1);
    };
    auto scatter = [&](void *dst, const void *src) {
        if (dtype == QSM_F32)
            crop_scatter_kernel<float><<<crop_blocks(cm, 1), 256>>>(
                static_cast<float *>(dst), static_cast<const float *>(src), cm, 1);
        else
            crop_scatter_kernel<double><<<crop_blocks(cm, 1), 256>>>(
                static_cast<double *>(dst), static_cast<const double *>(src), cm, 1);
    };
    for (size_t e = 0; e < nechoes; ++e) {
        QSM_HIP(hipMemcpy(dP.p, static_cast<const char *>(phas) + e * N * es, N * es,
                          hipMemcpyHostToDevice));
        gather(du.p, dP.p);
        QSM_HIP(hipGetLastError());
        if ((rc = qsm_unwrap_laplacian_dev(dtype, du.p, static_cast<const uint8_t *>(dm.p), cs,
                                           vsz, dout.p, dd2.p, dw.p, nullptr, nullptr)))
            return rc;
        QSM_HIP(hipMemsetAsync(dP.p, 0, N * es, nullptr));
        scatter(dP.p, dout.p);
        if (d2phas) {
            QSM_HIP(hipMemsetAsync(dD.p, 0, N * es, nullptr));
            scatter(dD.p, dd2.p);
        }
        QSM_HIP(hipGetLastError());
        QSM_HIP(hipMemcpy(static_cast<char *>(uphas) + e * N * es, dP.p, N * es,
                          hipMemcpyDeviceToHost));
        if (d2phas)
            QSM_HIP(hipMemcpy(static_cast<char *>(d2phas) + e * N * es, dD.p, N * es,
                              hipMemcpyDeviceToHost));
    }
    return QSM_OK;
}

}  // extern "C"
