// fd.hip -- finite-difference operators of QSM.m (src/inversion/mex/gradfp*,
// src/utils/fd/mex/grad{f,fm,b,bm}{,_adj}_mex.c, gradc{,m}_mex.c), the masked
// Laplacian (src/utils/fd/mex/lap1_mex.c) and the wrapped-phase Laplacian
// (src/unwrap/mex/lapw_mex.c) for gfx950.
//
// One thread per voxel, x fastest, so every wavefront reads and writes 64
// consecutive voxels; the +/-1 neighbours come from the same or adjacent
// cache lines.  These operators are HBM-bound elementwise stencils: 4 words
// per voxel for the gradients (1 read, 3 written) and their adjoints (3 read,
// 1 written), 2 for lapw.  The per-voxel formulas (including the order of the
// three axis terms of the adjoints) are the reference's, built with
// -ffp-contract=off, so results are bitwise the reference's.
#include <cmath>

#include "common.h"

using namespace qsm;

namespace {

constexpr int FX = 64, FY = 4;

struct G3 {
    int nx, ny, nz;
};

__device__ __forceinline__ bool vox(G3 g, int &i, int &j, int &k, int &l)
{
    i = blockIdx.x * FX + threadIdx.x;
    j = blockIdx.y * FY + threadIdx.y;
    k = blockIdx.z;
    l = i + g.nx * (j + g.ny * k);
    return i < g.nx && j < g.ny;
}

// gradfp_mex.c:163-241: periodic forward differences.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradfp_kernel(T *dx, T *dy, T *dz,
                                                        const T *u, G3 g,
                                                        T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const int lx = (i + 1 < g.nx) ? l + 1 : l - (g.nx - 1);
    const int ly = (j + 1 < g.ny) ? l + sy : l - sy * (g.ny - 1);
    const int lz = (k + 1 < g.nz) ? l + sk : l - sk * (g.nz - 1);
    const T c = u[l];
    dx[l] = hx * (u[lx] - c);
    dy[l] = hy * (u[ly] - c);
    dz[l] = hz * (u[lz] - c);
}

// gradfp_adj_mex.c:174-254: -(periodic backward divergence), with the
// reference's per-region term order: k == 0 -> (X+Y)+Z; k > 0, j == 0 ->
// (Z+X)+Y; otherwise (Z+Y)+X.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradfp_adj_kernel(T *du, const T *x,
                                                            const T *y, const T *z,
                                                            G3 g, T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const int lx = (i > 0) ? l - 1 : l + (g.nx - 1);
    const int ly = (j > 0) ? l - sy : l + sy * (g.ny - 1);
    const int lz = (k > 0) ? l - sk : l + sk * (g.nz - 1);
    const T X = hx * (x[l] - x[lx]);
    const T Y = hy * (y[l] - y[ly]);
    const T Z = hz * (z[l] - z[lz]);
    du[l] = (k == 0) ? (X + Y) + Z : (j == 0) ? (Z + X) + Y : (Z + Y) + X;
}

// gradf_mex.c:167-243: forward differences, backward on the last face.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradf_kernel(T *dx, T *dy, T *dz,
                                                       const T *u, G3 g,
                                                       T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const T c = u[l];
    dx[l] = (i + 1 < g.nx) ? hx * (u[l + 1] - c) : hx * (c - u[l - 1]);
    dy[l] = (j + 1 < g.ny) ? hy * (u[l + sy] - c) : hy * (c - u[l - sy]);
    dz[l] = (k + 1 < g.nz) ? hz * (u[l + sk] - c) : hz * (c - u[l - sk]);
}

// gradfm_mex.c:123-150: inside G, forward if the +1 neighbour is in G, else
// backward if the -1 neighbour is, else 0; outside G untouched.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradfm_kernel(T *dx, T *dy, T *dz,
                                                        const T *u,
                                                        const uint8_t *G, G3 g,
                                                        T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    if (!G[l]) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const T c = u[l];
    dx[l] = (i + 1 < g.nx && G[l + 1]) ? hx * (u[l + 1] - c)
          : (i > 0 && G[l - 1]) ? hx * (c - u[l - 1]) : T(0);
    dy[l] = (j + 1 < g.ny && G[l + sy]) ? hy * (u[l + sy] - c)
          : (j > 0 && G[l - sy]) ? hy * (c - u[l - sy]) : T(0);
    dz[l] = (k + 1 < g.nz && G[l + sk]) ? hz * (u[l + sk] - c)
          : (k > 0 && G[l - sk]) ? hz * (c - u[l - sk]) : T(0);
}

// gradf_adj_mex.c:168-245: (X+Y)+Z, backward differences scaled by -1/h,
// forward on the first face.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradf_adj_kernel(T *du, const T *x,
                                                           const T *y, const T *z,
                                                           G3 g, T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const T X = (i > 0) ? hx * (x[l] - x[l - 1]) : hx * (x[l + 1] - x[l]);
    const T Y = (j > 0) ? hy * (y[l] - y[l - sy]) : hy * (y[l + sy] - y[l]);
    const T Z = (k > 0) ? hz * (z[l] - z[l - sk]) : hz * (z[l + sk] - z[l]);
    du[l] = (X + Y) + Z;
}

// gradfm_adj_mex.c:131-157: inside G, backward if the -1 neighbour is in G,
// else forward if the +1 neighbour is, else 0; du = (dx+dy)+dz.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradfm_adj_kernel(T *du, const T *x,
                                                            const T *y, const T *z,
                                                            const uint8_t *G, G3 g,
                                                            T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    if (!G[l]) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const T DZ = (k > 0 && G[l - sk]) ? hz * (z[l] - z[l - sk])
               : (k + 1 < g.nz && G[l + sk]) ? hz * (z[l + sk] - z[l]) : T(0);
    const T DY = (j > 0 && G[l - sy]) ? hy * (y[l] - y[l - sy])
               : (j + 1 < g.ny && G[l + sy]) ? hy * (y[l + sy] - y[l]) : T(0);
    const T DX = (i > 0 && G[l - 1]) ? hx * (x[l] - x[l - 1])
               : (i + 1 < g.nx && G[l + 1]) ? hx * (x[l + 1] - x[l]) : T(0);
    du[l] = (DX + DY) + DZ;
}

// gradb_mex.c:142-226: backward differences, forward on the first face.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradb_kernel(T *dx, T *dy, T *dz,
                                                       const T *u, G3 g,
                                                       T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const T c = u[l];
    dx[l] = (i > 0) ? hx * (c - u[l - 1]) : hx * (u[l + 1] - c);
    dy[l] = (j > 0) ? hy * (c - u[l - sy]) : hy * (u[l + sy] - c);
    dz[l] = (k > 0) ? hz * (c - u[l - sk]) : hz * (u[l + sk] - c);
}

// gradbm_mex.c:121-150: inside G, backward if the -1 neighbour is in G,
// else forward if the +1 neighbour is, else 0; outside G untouched.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradbm_kernel(T *dx, T *dy, T *dz,
                                                        const T *u,
                                                        const uint8_t *G, G3 g,
                                                        T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    if (!G[l]) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const T c = u[l];
    dx[l] = (i > 0 && G[l - 1]) ? hx * (c - u[l - 1])
          : (i + 1 < g.nx && G[l + 1]) ? hx * (u[l + 1] - c) : T(0);
    dy[l] = (j > 0 && G[l - sy]) ? hy * (c - u[l - sy])
          : (j + 1 < g.ny && G[l + sy]) ? hy * (u[l + sy] - c) : T(0);
    dz[l] = (k > 0 && G[l - sk]) ? hz * (c - u[l - sk])
          : (k + 1 < g.nz && G[l + sk]) ? hz * (u[l + sk] - c) : T(0);
}

// gradb_adj_mex.c:159-251: (X+Y)+Z, forward differences scaled by -1/h,
// backward on the last face; the reference's loops never write (nx-1,
// ny-1, nz-1), which keeps the caller's value.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradb_adj_kernel(T *du, const T *x,
                                                           const T *y, const T *z,
                                                           G3 g, T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    if (i + 1 == g.nx && j + 1 == g.ny && k + 1 == g.nz) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const T X = (i + 1 < g.nx) ? hx * (x[l + 1] - x[l]) : hx * (x[l] - x[l - 1]);
    const T Y = (j + 1 < g.ny) ? hy * (y[l + sy] - y[l]) : hy * (y[l] - y[l - sy]);
    const T Z = (k + 1 < g.nz) ? hz * (z[l + sk] - z[l]) : hz * (z[l] - z[l - sk]);
    du[l] = (X + Y) + Z;
}

// gradbm_adj_mex.c:131-157: inside G, forward if the +1 neighbour is in G,
// else backward if the -1 neighbour is, else 0; du = (dx+dy)+dz.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradbm_adj_kernel(T *du, const T *x,
                                                            const T *y, const T *z,
                                                            const uint8_t *G, G3 g,
                                                            T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    if (!G[l]) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const T DX = (i + 1 < g.nx && G[l + 1]) ? hx * (x[l + 1] - x[l])
               : (i > 0 && G[l - 1]) ? hx * (x[l] - x[l - 1]) : T(0);
    const T DY = (j + 1 < g.ny && G[l + sy]) ? hy * (y[l + sy] - y[l])
               : (j > 0 && G[l - sy]) ? hy * (y[l] - y[l - sy]) : T(0);
    const T DZ = (k + 1 < g.nz && G[l + sk]) ? hz * (z[l + sk] - z[l])
               : (k > 0 && G[l - sk]) ? hz * (z[l] - z[l - sk]) : T(0);
    du[l] = (DX + DY) + DZ;
}

// gradc_mex.c:309-554: per axis the central difference (0.5/h), forward on
// the first face, backward on the last.  As the reference's loops leave it:
// in plane k = 0 every l in [NX + nx, NX + NY) (NX = nx-1, NY = nx(ny-1))
// holds what its "i = nx-1" loop writes last (it steps j by 1: dx backward,
// dy central, dz forward), and (nx-1, ny-1, nz-1) is never written.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradc_kernel(T *dx, T *dy, T *dz, const T *u, G3 g,
                                                       T hx, T hy, T hz, T hx2, T hy2, T hz2)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    if (i + 1 == g.nx && j + 1 == g.ny && k + 1 == g.nz) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    if (k == 0 && l >= (g.nx - 1) + g.nx && l < (g.nx - 1) + g.nx * (g.ny - 1)) {
        dx[l] = hx * (u[l] - u[l - 1]);
        dy[l] = hy2 * (u[l + sy] - u[l - sy]);
        dz[l] = hz * (u[l + sk] - u[l]);
        return;
    }
    dx[l] = (i == 0) ? hx * (u[l + 1] - u[l])
          : (i + 1 == g.nx) ? hx * (u[l] - u[l - 1]) : hx2 * (u[l + 1] - u[l - 1]);
    dy[l] = (j == 0) ? hy * (u[l + sy] - u[l])
          : (j + 1 == g.ny) ? hy * (u[l] - u[l - sy]) : hy2 * (u[l + sy] - u[l - sy]);
    dz[l] = (k == 0) ? hz * (u[l + sk] - u[l])
          : (k + 1 == g.nz) ? hz * (u[l] - u[l - sk]) : hz2 * (u[l + sk] - u[l - sk]);
}

// gradcm_mex.c:122-186: inside G, central if both neighbours are in G, else
// forward if the +1 neighbour is, else backward if the -1 neighbour is,
// else 0; outside G untouched.
template <typename T>
__global__ __launch_bounds__(FX *FY) void gradcm_kernel(T *dx, T *dy, T *dz, const T *u,
                                                        const uint8_t *G, G3 g, T hx, T hy,
                                                        T hz, T hx2, T hy2, T hz2)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    if (!G[l]) return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const bool xm = i > 0 && G[l - 1], xp = i + 1 < g.nx && G[l + 1];
    const bool ym = j > 0 && G[l - sy], yp = j + 1 < g.ny && G[l + sy];
    const bool zm = k > 0 && G[l - sk], zp = k + 1 < g.nz && G[l + sk];
    dx[l] = (xm && xp) ? hx2 * (u[l + 1] - u[l - 1])
          : xp ? hx * (u[l + 1] - u[l]) : xm ? hx * (u[l] - u[l - 1]) : T(0);
    dy[l] = (ym && yp) ? hy2 * (u[l + sy] - u[l - sy])
          : yp ? hy * (u[l + sy] - u[l]) : ym ? hy * (u[l] - u[l - sy]) : T(0);
    dz[l] = (zm && zp) ? hz2 * (u[l + sk] - u[l - sk])
          : zp ? hz * (u[l + sk] - u[l]) : zm ? hz * (u[l] - u[l - sk]) : T(0);
}

__device__ __forceinline__ double fmod_t(double a, double b) { return fmod(a, b); }
__device__ __forceinline__ float fmod_t(float a, float b) { return fmodf(a, b); }

// lapw_mex.c:170-187: wrap to [-pi, pi] by fmod, applied when
// d < -pi || d > pi (lapw_mex.c:10-11).
template <typename T>
__device__ __forceinline__ T wrap_ref(T d)
{
    const T pi = T(3.14159265358979323846);
    const T tau = T(2 * 3.14159265358979323846);
    if (!(d < -pi || d > pi)) return d;
    const T x = d + pi;
    return x < T(0) ? fmod_t(x, tau) + pi : fmod_t(x, tau) - pi;
}

// The same wrap, bitwise, for |d + pi| < 2 tau.  On a wrapped lane
// x = d + pi is nonzero, and fmod(x, tau) is x for |x| < tau, else
// x - copysign(tau, x), a subtraction that is exact (Sterbenz: tau <= |x| <
// 2 tau); the zero fmod returns at |x| = tau differs from this one in sign
// only, which the following -/+ pi removes.  `slow` is raised where that
// range does not hold (|x| >= 2 tau, Inf, NaN).
template <typename T>
__device__ __forceinline__ T wrap_fast(T d, bool &slow)
{
    const T pi = T(3.14159265358979323846);
    const T tau = T(2 * 3.14159265358979323846);
    const T x = d + pi;
    const T ax = fabs(x);
    slow |= !(ax < T(2) * tau);
    const T r = ax < tau ? x : x - copysign(tau, x);
    const T w = r - copysign(pi, x);
    return fabs(d) > pi ? w : d;
}

// One voxel from its seven samples, with the reference's term order
// (lapw_mex.c:139-162).
template <typename T>
__device__ __noinline__ T lapw_voxel_ref(T c, T xm, T xp, T ym, T yp, T zm, T zp,
                                         T hx, T hy, T hz)
{
    T acc = hx * (wrap_ref(xp - c) - wrap_ref(c - xm));
    acc = acc + hy * (wrap_ref(yp - c) - wrap_ref(c - ym));
    return acc + hz * (wrap_ref(zp - c) - wrap_ref(c - zm));
}

// lapw_mex.c:133-164: interior voxels only.  Lanes whose differences leave
// the fast wrap's range (phase jumps beyond 3 pi, Inf, NaN) redo the voxel
// with the library fmod.
template <typename T>
__global__ __launch_bounds__(FX *FY) void lapw_kernel(T *du, const T *u, G3 g,
                                                      T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    if (i < 1 || i + 1 >= g.nx || j < 1 || j + 1 >= g.ny || k < 1 || k + 1 >= g.nz)
        return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const T c = u[l], xm = u[l - 1], xp = u[l + 1];
    const T ym = u[l - sy], yp = u[l + sy], zm = u[l - sk], zp = u[l + sk];
    bool slow = false;
    T acc = hx * (wrap_fast(xp - c, slow) - wrap_fast(c - xm, slow));
    acc = acc + hy * (wrap_fast(yp - c, slow) - wrap_fast(c - ym, slow));
    acc = acc + hz * (wrap_fast(zp - c, slow) - wrap_fast(c - zm, slow));
    if (slow)
        acc = lapw_voxel_ref(c, xm, xp, ym, yp, zm, zp, hx, hy, hz);
    du[l] = acc;
}

// lap1_mex.c:208-315: the second difference along one axis at a voxel of
// G, index ii of extent n >= 2 along the axis; U(d) / M(d) give u and "in
// G" at offset d in [-3, 3] along it (read only where ii + d is inside).
// Inside the axis the +/-1 neighbours in G pick central / forward /
// backward / none; on the first and last index forward / backward whatever
// G holds.  One-sided forms are third order when the next two cells are in
// G, else second order when the second is, else the first difference
// (backward's with the reference's sign, u[-1] - u[0]).
template <typename T, typename FU, typename FM>
__device__ __forceinline__ T lap1_axis(const FU &U, const FM &M, int ii, int n)
{
    int idx;
    if (ii >= 1 && ii + 2 <= n)
        idx = 2 * M(1) + M(-1);
    else
        idx = 2 * (ii == 0) + (ii + 1 == n);
    const T two = T(2), four = T(4), five = T(5);
    if (idx == 3) return (U(-1) - two * U(0)) + U(1);
    if (idx == 2) {
        if (ii + 3 < n && M(2) && M(3))
            return ((two * U(0) - five * U(1)) + four * U(2)) - U(3);
        if (ii + 2 < n && M(2)) return (U(0) - two * U(1)) + U(2);
        return U(1) - U(0);
    }
    if (idx == 1) {
        if (ii >= 3 && M(-3) && M(-2))
            return ((-U(-3) + four * U(-2)) - five * U(-1)) + two * U(0);
        if (ii >= 2 && M(-2)) return (U(-2) - two * U(-1)) + U(0);
        return U(-1) - U(0);
    }
    return T(0);
}

// lap1_mex.c:63-205: on G, d2u = (hx D_x + hy D_y) + hz D_z, each partial sum
// rounded in T as the reference's "+=" does; outside G untouched.
//
// The +/-3 reach is served without re-reading: a block of LX x LY threads
// owns a tile of columns and marches up LZ planes; each thread keeps its
// column's u and mask at planes k-3 .. k+3 in registers (the z terms), and
// the block stages plane k of the tile plus the axis-aligned halo the x and
// y terms read (3 cells beyond each edge, no corners) in LDS, two plane
// slots so one barrier per plane suffices.  Cells outside the grid hold 0
// and are never read (the rule tests ii against the extent first).
constexpr int LX = 64, LY = 8, LZ = 32, LH = 3;
constexpr int LW = LX + 2 * LH, LHT = LY + 2 * LH;
constexpr int LHALO = 2 * LH * LX + 2 * LH * LY;  // y rows + x columns

template <typename T>
__global__ __launch_bounds__(LX *LY) void lap1_kernel(T *du, const T *u, const uint8_t *G,
                                                      G3 g, T hx, T hy, T hz)
{
    __shared__ T su[2][LHT][LW];
    __shared__ uint8_t sg[2][LHT][LW];
    const int tx = threadIdx.x, ty = threadIdx.y, t = tx + LX * ty;
    const int i0 = blockIdx.x * LX, j0 = blockIdx.y * LY;
    const int kb = blockIdx.z * LZ, ke = min(kb + LZ, g.nz);
    const int i = i0 + tx, j = j0 + ty;
    const bool own = i < g.nx && j < g.ny;
    const size_t sk = size_t(g.nx) * g.ny;
    const size_t col = own ? size_t(i) + size_t(g.nx) * j : 0;
    // this thread's halo cell (at most one: LHALO <= LX * LY) and its column
    int hx_ = -1, hy_ = -1;
    bool hin = false;
    size_t hcol = 0;
    if (t < LHALO) {
        if (t < 2 * LH * LX) {
            const int r = t / LX;
            hx_ = LH + t % LX;
            hy_ = r < LH ? r : r + LY;
        } else {
            const int q = t - 2 * LH * LX, c = q % (2 * LH);
            hy_ = LH + q / (2 * LH);
            hx_ = c < LH ? c : LX + c;
        }
        const int gi = i0 + hx_ - LH, gj = j0 + hy_ - LH;
        hin = gi >= 0 && gi < g.nx && gj >= 0 && gj < g.ny;
        hcol = hin ? size_t(gi) + size_t(g.nx) * gj : 0;
    }
    T wu[7];
    uint8_t wg[7];
#pragma unroll
    for (int d = 0; d < 7; ++d) {
        const int kk = kb - 3 + d;
        const bool in = own && kk >= 0 && kk < g.nz;
        wu[d] = in ? u[col + sk * kk] : T(0);
        wg[d] = in ? uint8_t(G[col + sk * kk] != 0) : uint8_t(0);
    }
    for (int k = kb; k < ke; ++k) {
        const int slot = k & 1;
        // next plane of the own column, consumed at the end of this step
        const int kn = k + 4;
        const bool nin = own && kn < g.nz;
        const T un = nin ? u[col + sk * kn] : T(0);
        const uint8_t gn = nin ? uint8_t(G[col + sk * kn] != 0) : uint8_t(0);
        su[slot][ty + LH][tx + LH] = wu[3];
        sg[slot][ty + LH][tx + LH] = wg[3];
        if (t < LHALO) {
            su[slot][hy_][hx_] = hin ? u[hcol + sk * k] : T(0);
            sg[slot][hy_][hx_] = hin ? uint8_t(G[hcol + sk * k] != 0) : uint8_t(0);
        }
        __syncthreads();
        if (own && wg[3]) {
            const int cy = ty + LH, cx = tx + LH;
            auto UX = [&](int d) { return su[slot][cy][cx + d]; };
            auto MX = [&](int d) { return int(sg[slot][cy][cx + d]); };
            auto UY = [&](int d) { return su[slot][cy + d][cx]; };
            auto MY = [&](int d) { return int(sg[slot][cy + d][cx]); };
            auto UZ = [&](int d) { return wu[3 + d]; };
            auto MZ = [&](int d) { return int(wg[3 + d]); };
            T d2 = hx * lap1_axis<T>(UX, MX, i, g.nx);
            d2 = d2 + hy * lap1_axis<T>(UY, MY, j, g.ny);
            d2 = d2 + hz * lap1_axis<T>(UZ, MZ, k, g.nz);
            du[col + sk * k] = d2;
        }
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            wu[d] = wu[d + 1];
            wg[d] = wg[d + 1];
        }
        wu[6] = un;
        wg[6] = gn;
    }
}

dim3 lap1_grid(G3 g) { return dim3((g.nx + LX - 1) / LX, (g.ny + LY - 1) / LY, (g.nz + LZ - 1) / LZ); }

dim3 fd_grid(G3 g) { return dim3((g.nx + FX - 1) / FX, (g.ny + FY - 1) / FY, g.nz); }

enum Op { GRADFP, GRADFP_ADJ, GRADF, GRADFM, GRADF_ADJ, GRADFM_ADJ, LAPW,
          GRADB, GRADBM, GRADB_ADJ, GRADBM_ADJ, GRADC, GRADCM, LAP1 };

bool is_adj(Op op)
{
    return op == GRADFP_ADJ || op == GRADF_ADJ || op == GRADFM_ADJ || op == GRADB_ADJ ||
           op == GRADBM_ADJ;
}

bool is_masked(Op op)
{
    return op == GRADFM || op == GRADFM_ADJ || op == GRADBM || op == GRADBM_ADJ ||
           op == GRADCM || op == LAP1;
}

// Coefficients as the reference forms them: 1/h (gradients), -1/h (adjoints),
// 1/h^2 (lapw, lap1), in double and cast on the single path.
template <typename T>
void coefs(Op op, const double h[3], T c[3])
{
    for (int a = 0; a < 3; ++a) {
        if (op == LAPW || op == LAP1) c[a] = T(1.0 / (h[a] * h[a]));
        else if (is_adj(op)) c[a] = T(-1.0 / h[a]);
        else c[a] = T(1.0 / h[a]);
    }
}

// Launches op on device arrays: a3 = the three vector arrays (outputs of the
// gradients, inputs of the adjoints), s = the scalar array.
template <typename T>
hipError_t launch_op(Op op, T *s, T *a0, T *a1, T *a2, const uint8_t *G,
                     G3 g, const double h[3], hipStream_t st)
{
    T c[3];
    coefs<T>(op, h, c);
    const dim3 grid = fd_grid(g), blk(FX, FY);
    switch (op) {
    case GRADFP: gradfp_kernel<T><<<grid, blk, 0, st>>>(a0, a1, a2, s, g, c[0], c[1], c[2]); break;
    case GRADFP_ADJ: gradfp_adj_kernel<T><<<grid, blk, 0, st>>>(s, a0, a1, a2, g, c[0], c[1], c[2]); break;
    case GRADF: gradf_kernel<T><<<grid, blk, 0, st>>>(a0, a1, a2, s, g, c[0], c[1], c[2]); break;
    case GRADFM: gradfm_kernel<T><<<grid, blk, 0, st>>>(a0, a1, a2, s, G, g, c[0], c[1], c[2]); break;
    case GRADF_ADJ: gradf_adj_kernel<T><<<grid, blk, 0, st>>>(s, a0, a1, a2, g, c[0], c[1], c[2]); break;
    case GRADFM_ADJ: gradfm_adj_kernel<T><<<grid, blk, 0, st>>>(s, a0, a1, a2, G, g, c[0], c[1], c[2]); break;
    case LAPW: lapw_kernel<T><<<grid, blk, 0, st>>>(a0, s, g, c[0], c[1], c[2]); break;
    case LAP1:
        lap1_kernel<T><<<lap1_grid(g), dim3(LX, LY), 0, st>>>(a0, s, G, g, c[0], c[1], c[2]);
        break;
    case GRADB: gradb_kernel<T><<<grid, blk, 0, st>>>(a0, a1, a2, s, g, c[0], c[1], c[2]); break;
    case GRADBM: gradbm_kernel<T><<<grid, blk, 0, st>>>(a0, a1, a2, s, G, g, c[0], c[1], c[2]); break;
    case GRADB_ADJ: gradb_adj_kernel<T><<<grid, blk, 0, st>>>(s, a0, a1, a2, g, c[0], c[1], c[2]); break;
    case GRADBM_ADJ: gradbm_adj_kernel<T><<<grid, blk, 0, st>>>(s, a0, a1, a2, G, g, c[0], c[1], c[2]); break;
    case GRADC:
    case GRADCM: {
        // 0.5/h formed in double and cast, as gradc_mex.c / gradcm_mex.c do
        const T c2[3] = {T(0.5 / h[0]), T(0.5 / h[1]), T(0.5 / h[2])};
        if (op == GRADC)
            gradc_kernel<T><<<grid, blk, 0, st>>>(a0, a1, a2, s, g, c[0], c[1], c[2], c2[0], c2[1], c2[2]);
        else
            gradcm_kernel<T><<<grid, blk, 0, st>>>(a0, a1, a2, s, G, g, c[0], c[1], c[2], c2[0], c2[1], c2[2]);
        break;
    }
    }
    return hipGetLastError();
}

// the unmasked one-sided forms read the +/-1 neighbour on every face, and so
// do lap1's (lap1_mex.c reads out of bounds on an axis of extent 1)
size_t min_dim(Op op)
{
    return (op == GRADF || op == GRADF_ADJ || op == GRADB || op == GRADB_ADJ || op == GRADC ||
            op == LAP1) ? 2 : 1;
}

int run_dev(Op op, int dtype, void *s, void *a0, void *a1, void *a2,
            const uint8_t *G, const size_t sz[3], const double h[3], void *stream)
{
    int rc = check_grid(sz, h, min_dim(op));
    if (rc) return rc;
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    QSM_ARG(s && a0 && (op == LAPW || op == LAP1 || (a1 && a2)), "null array pointer");
    QSM_ARG(!is_masked(op) || G, "null mask pointer");
    const G3 g{int(sz[0]), int(sz[1]), int(sz[2])};
    if (dtype == QSM_F64)
        QSM_HIP(launch_op<double>(op, (double *)s, (double *)a0, (double *)a1, (double *)a2, G, g, h, as_stream(stream)));
    else
        QSM_HIP(launch_op<float>(op, (float *)s, (float *)a0, (float *)a1, (float *)a2, G, g, h, as_stream(stream)));
    return QSM_OK;
}

// Host-pointer form: stage to the device, run, copy the outputs back.
// in_s/in_v: whether the scalar / vector arrays are inputs; outputs that the
// operator leaves partly untouched (masked variants, lapw's faces) are staged
// in too so the untouched voxels keep the caller's values.
int run_host(Op op, int dtype, void *s, void *a0, void *a1, void *a2,
             const uint8_t *G, const size_t sz[3], const double h[3])
{
    int rc = check_grid(sz, h, min_dim(op));
    if (rc) return rc;
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    const bool lap = op == LAPW || op == LAP1;
    QSM_ARG(s && a0 && (lap || (a1 && a2)), "null array pointer");
    const bool masked = is_masked(op);
    QSM_ARG(!masked || G, "null mask pointer");
    // outputs with voxels the operator leaves untouched (besides the masked
    // forms): the corner gradb_adj / gradc never write
    const bool keep = masked || op == GRADB_ADJ || op == GRADC;
    if ((rc = ensure_device())) return rc;
    const size_t N = sz[0] * sz[1] * sz[2], bytes = N * dtype_size(dtype);
    const bool vec_out = op == GRADFP || op == GRADF || op == GRADFM || op == GRADB ||
                         op == GRADBM || op == GRADC || op == GRADCM;
    const int nv = lap ? 1 : 3;
    DevBuf ds, dv[3], dg;
    if ((rc = ds.alloc(bytes))) return rc;
    for (int v = 0; v < nv; ++v)
        if ((rc = dv[v].alloc(bytes))) return rc;
    void *hv[3] = {a0, a1, a2};
    // scalar array: input of gradients; output of adjoints (staged when masked)
    if (vec_out || keep || lap)
        QSM_HIP(hipMemcpy(ds.p, s, bytes, hipMemcpyHostToDevice));
    for (int v = 0; v < nv; ++v)
        if (!vec_out || keep || lap)
            QSM_HIP(hipMemcpy(dv[v].p, hv[v], bytes, hipMemcpyHostToDevice));
    if (masked) {
        if ((rc = dg.alloc(N))) return rc;
        QSM_HIP(hipMemcpy(dg.p, G, N, hipMemcpyHostToDevice));
    }
    if ((rc = run_dev(op, dtype, ds.p, dv[0].p, lap ? dv[0].p : dv[1].p,
                      lap ? dv[0].p : dv[2].p, (const uint8_t *)dg.p, sz, h, nullptr)))
        return rc;
    if (vec_out || lap) {
        for (int v = 0; v < nv; ++v)
            QSM_HIP(hipMemcpy(hv[v], dv[v].p, bytes, hipMemcpyDeviceToHost));
    } else {
        QSM_HIP(hipMemcpy(s, ds.p, bytes, hipMemcpyDeviceToHost));
    }
    return QSM_OK;
}

}  // namespace

extern "C" {

int qsm_gradfp(int dtype, void *dx, void *dy, void *dz, const void *u,
               const size_t sz[3], const double h[3])
{
    return run_host(GRADFP, dtype, const_cast<void *>(u), dx, dy, dz, nullptr, sz, h);
}

int qsm_gradfp_adj(int dtype, void *du, const void *x, const void *y,
                   const void *z, const size_t sz[3], const double h[3])
{
    return run_host(GRADFP_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                    const_cast<void *>(z), nullptr, sz, h);
}

int qsm_gradf(int dtype, void *dx, void *dy, void *dz, const void *u,
              const size_t sz[3], const double h[3])
{
    return run_host(GRADF, dtype, const_cast<void *>(u), dx, dy, dz, nullptr, sz, h);
}

int qsm_gradfm(int dtype, void *dx, void *dy, void *dz, const void *u,
               const uint8_t *G, const size_t sz[3], const double h[3])
{
    return run_host(GRADFM, dtype, const_cast<void *>(u), dx, dy, dz, G, sz, h);
}

int qsm_gradf_adj(int dtype, void *du, const void *x, const void *y,
                  const void *z, const size_t sz[3], const double h[3])
{
    return run_host(GRADF_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                    const_cast<void *>(z), nullptr, sz, h);
}

int qsm_gradfm_adj(int dtype, void *du, const void *x, const void *y,
                   const void *z, const uint8_t *G, const size_t sz[3],
                   const double h[3])
{
    return run_host(GRADFM_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                    const_cast<void *>(z), G, sz, h);
}

int qsm_gradb(int dtype, void *dx, void *dy, void *dz, const void *u,
              const size_t sz[3], const double h[3])
{
    return run_host(GRADB, dtype, const_cast<void *>(u), dx, dy, dz, nullptr, sz, h);
}

int qsm_gradbm(int dtype, void *dx, void *dy, void *dz, const void *u, const uint8_t *G,
               const size_t sz[3], const double h[3])
{
    return run_host(GRADBM, dtype, const_cast<void *>(u), dx, dy, dz, G, sz, h);
}

int qsm_gradb_adj(int dtype, void *du, const void *x, const void *y, const void *z,
                  const size_t sz[3], const double h[3])
{
    return run_host(GRADB_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                    const_cast<void *>(z), nullptr, sz, h);
}

int qsm_gradbm_adj(int dtype, void *du, const void *x, const void *y, const void *z,
                   const uint8_t *G, const size_t sz[3], const double h[3])
{
    return run_host(GRADBM_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                    const_cast<void *>(z), G, sz, h);
}

int qsm_gradc(int dtype, void *dx, void *dy, void *dz, const void *u,
              const size_t sz[3], const double h[3])
{
    return run_host(GRADC, dtype, const_cast<void *>(u), dx, dy, dz, nullptr, sz, h);
}

int qsm_gradcm(int dtype, void *dx, void *dy, void *dz, const void *u, const uint8_t *G,
               const size_t sz[3], const double h[3])
{
    return run_host(GRADCM, dtype, const_cast<void *>(u), dx, dy, dz, G, sz, h);
}

int qsm_lap1(int dtype, void *d2u, const void *u, const uint8_t *G, const size_t sz[3],
             const double h[3])
{
    return run_host(LAP1, dtype, const_cast<void *>(u), d2u, nullptr, nullptr, G, sz, h);
}

int qsm_lapw(int dtype, void *d2u, const void *u, const size_t sz[3],
             const double h[3])
{
    return run_host(LAPW, dtype, const_cast<void *>(u), d2u, nullptr, nullptr, nullptr, sz, h);
}

int qsm_gradfp_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                   const size_t sz[3], const double h[3], void *stream)
{
    return run_dev(GRADFP, dtype, const_cast<void *>(u), dx, dy, dz, nullptr, sz, h, stream);
}

int qsm_gradfp_adj_dev(int dtype, void *du, const void *x, const void *y,
                       const void *z, const size_t sz[3], const double h[3],
                       void *stream)
{
    return run_dev(GRADFP_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                   const_cast<void *>(z), nullptr, sz, h, stream);
}

int qsm_gradf_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                  const size_t sz[3], const double h[3], void *stream)
{
    return run_dev(GRADF, dtype, const_cast<void *>(u), dx, dy, dz, nullptr, sz, h, stream);
}

int qsm_gradfm_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                   const uint8_t *G, const size_t sz[3], const double h[3],
                   void *stream)
{
    return run_dev(GRADFM, dtype, const_cast<void *>(u), dx, dy, dz, G, sz, h, stream);
}

int qsm_gradf_adj_dev(int dtype, void *du, const void *x, const void *y,
                      const void *z, const size_t sz[3], const double h[3],
                      void *stream)
{
    return run_dev(GRADF_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                   const_cast<void *>(z), nullptr, sz, h, stream);
}

int qsm_gradfm_adj_dev(int dtype, void *du, const void *x, const void *y,
                       const void *z, const uint8_t *G, const size_t sz[3],
                       const double h[3], void *stream)
{
    return run_dev(GRADFM_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                   const_cast<void *>(z), G, sz, h, stream);
}

int qsm_gradb_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                  const size_t sz[3], const double h[3], void *stream)
{
    return run_dev(GRADB, dtype, const_cast<void *>(u), dx, dy, dz, nullptr, sz, h, stream);
}

int qsm_gradbm_dev(int dtype, void *dx, void *dy, void *dz, const void *u, const uint8_t *G,
                   const size_t sz[3], const double h[3], void *stream)
{
    return run_dev(GRADBM, dtype, const_cast<void *>(u), dx, dy, dz, G, sz, h, stream);
}

int qsm_gradb_adj_dev(int dtype, void *du, const void *x, const void *y, const void *z,
                      const size_t sz[3], const double h[3], void *stream)
{
    return run_dev(GRADB_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                   const_cast<void *>(z), nullptr, sz, h, stream);
}

int qsm_gradbm_adj_dev(int dtype, void *du, const void *x, const void *y, const void *z,
                       const uint8_t *G, const size_t sz[3], const double h[3], void *stream)
{
    return run_dev(GRADBM_ADJ, dtype, du, const_cast<void *>(x), const_cast<void *>(y),
                   const_cast<void *>(z), G, sz, h, stream);
}

int qsm_gradc_dev(int dtype, void *dx, void *dy, void *dz, const void *u,
                  const size_t sz[3], const double h[3], void *stream)
{
    return run_dev(GRADC, dtype, const_cast<void *>(u), dx, dy, dz, nullptr, sz, h, stream);
}

int qsm_gradcm_dev(int dtype, void *dx, void *dy, void *dz, const void *u, const uint8_t *G,
                   const size_t sz[3], const double h[3], void *stream)
{
    return run_dev(GRADCM, dtype, const_cast<void *>(u), dx, dy, dz, G, sz, h, stream);
}

int qsm_lapw_dev(int dtype, void *d2u, const void *u, const size_t sz[3],
                 const double h[3], void *stream)
{
    return run_dev(LAPW, dtype, const_cast<void *>(u), d2u, nullptr, nullptr, nullptr, sz, h, stream);
}

int qsm_lap1_dev(int dtype, void *d2u, const void *u, const uint8_t *G, const size_t sz[3],
                 const double h[3], void *stream)
{
    return run_dev(LAP1, dtype, const_cast<void *>(u), d2u, nullptr, nullptr, G, sz, h, stream);
}

}  // extern "C"
