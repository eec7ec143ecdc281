// admm.hip -- total-variation dipole inversion by ADMM (src/inversion/tv.m)
// on gfx950.
//
// tv.m:124-308 per echo: F = mu conj(D) fft3(f), iA = 1/(mu |D|^2 + rho L)
// with L the symbol of the negative Laplacian; u = f, v = grad(u), l = 0;
// each iteration
//   u-subproblem   u = real(ifft3(iA .* (F + fft3(gradAdj(rho v - l)))))
//   v-subproblem   over-relaxed gradient of u, soft-thresholded by 1/rho
//                  (anisotropic per component, isotropic on the vector)
//   multipliers    l += rho (dx1 - v)
//   convergence    primal ||grad u - v||_{2,1}, dual ||rho gradAdj(dv)||_2
//                  against absolute + relative tolerances; rho adapted
// and finally x = mask .* u.
//
// Device layout: every real field is a column-major volume of the solver
// grid; the spectra F and the work spectrum hold the half spectrum along x
// ((nx/2 + 1) ny nz complex).  Per iteration up to three launches touch
// HBM: the right-hand side (gradAdj of rho v - l, 6 reads + 1 write per
// voxel), the half-spectrum solve (five FFT passes, the k-space multipliers
// generated in-kernel from per-axis tables), and the v/l update with the
// partial sums of both residuals (7 reads + 7 writes; (v, l) kept in two
// sets so that the dual residual's -1 neighbour terms are recomputed from
// the previous set instead of stored).  The update also writes the next
// right-hand side at the current rho; the right-hand-side launch does work
// only when the rule has changed rho (and in the first iteration).  Sums
// are accumulated in double per block and reduced in a fixed order, so
// they are reproducible.
//
// The loop runs on the device: the stopping rule and the penalty update
// (tv.m:279-306) are taken by a one-block kernel after each iteration's
// reduction, from a control block in device memory that every kernel of
// the next iteration reads (rho, 1/rho, stop).  The host enqueues maxit
// iterations without waiting; once the rule stops the loop, the remaining
// launches return at once.  One synchronisation per solve.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "fft_dev.h"

using namespace qsm;
using namespace qsm::fft;

namespace {

constexpr int BX = 64, BY = 4, NT = BX * BY;
// partial sums per block: [0..2] from the v-update (nr, na, nb), [3..4]
// from the dual norms (||s||^2, ||rho gradAdj(l)||^2); the initial |v| sum
// uses [0]
constexpr int NPART = 5;
// first-level reduction blocks (each sums a contiguous range of the
// per-block partials; the decision kernel sums their RB results)
constexpr int RB = 128;
// iterations the device trace holds
constexpr int MAX_ITERS = 65536;

// tv.m's loop scalars, in device memory.
struct Ctl {
    double rho, irho;  // penalty and its inverse for the next iteration
    double nr;         // last primal residual (nr_ of the next check)
    int flag;          // penalty-update counter (tv.m:287-306)
    int stop;          // set when the rule stops: later launches return
    int iters;         // iterations done
    int rkeep;         // rho unchanged: r already holds the next right-hand side
};

// Constants of the stopping rule and the penalty update.
struct Rule {
    double atolp, atold, rtol, rtau, stau, gamma;
};

struct G3 {
    int nx, ny, nz;
};

__device__ __forceinline__ bool vox(G3 g, int &i, int &j, int &k, int &l)
{
    i = blockIdx.x * BX + threadIdx.x;
    j = blockIdx.y * BY + threadIdx.y;
    k = blockIdx.z;
    l = i + g.nx * (j + g.ny * k);
    return i < g.nx && j < g.ny;
}

dim3 vox_grid(G3 g) { return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, g.nz); }

int vox_blocks(G3 g)
{
    const dim3 d = vox_grid(g);
    return int(d.x * d.y * d.z);
}

__device__ __forceinline__ int flat_block()
{
    return int(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}

// Tile t of the voxel grid (BX x BY voxels of one plane; t = bx + gx (by +
// gy k), the block index vox_grid would give it): the thread's voxel,
// false outside the volume.
__device__ __forceinline__ bool tile_vox(G3 g, int t, int &i, int &j, int &k, int &l)
{
    const int gx = (g.nx + BX - 1) / BX, gy = (g.ny + BY - 1) / BY;
    const int r = t / gx;
    i = (t - r * gx) * BX + threadIdx.x;
    j = (r % gy) * BY + threadIdx.y;
    k = r / gy;
    l = i + g.nx * (j + g.ny * k);
    return i < g.nx && j < g.ny;
}

// Sum of v[q] over the block's NT threads into out[q] for q < NQ, in a
// fixed order (thread 0 ends with the totals).
template <int NQ>
__device__ __forceinline__ void block_sum(double (&v)[NQ], double *out)
{
    __shared__ double red[NQ][NT];
    const int t = threadIdx.x + BX * threadIdx.y;
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        red[q][t] = v[q];
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s)
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                red[q][t] += red[q][t + s];
        __syncthreads();
    }
    if (t == 0)
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            out[q] = red[q][0];
}

// Periodic -1 neighbours (gradfp_adj_mex.c:174-254) and the reference's
// per-region term order of the adjoint: k == 0 -> (X+Y)+Z; k > 0, j == 0
// -> (Z+X)+Y; otherwise (Z+Y)+X.
__device__ __forceinline__ void minus_nbrs(G3 g, int i, int j, int k, int l, int &lx, int &ly,
                                           int &lz)
{
    const int sy = g.nx, sk = g.nx * g.ny;
    lx = (i > 0) ? l - 1 : l + (g.nx - 1);
    ly = (j > 0) ? l - sy : l + sy * (g.ny - 1);
    lz = (k > 0) ? l - sk : l + sk * (g.nz - 1);
}

template <typename T>
__device__ __forceinline__ T adj_sum(T X, T Y, T Z, int j, int k)
{
    return (k == 0) ? (X + Y) + Z : (j == 0) ? (Z + X) + Y : (Z + Y) + X;
}

// Periodic +1 neighbours (gradfp_mex.c:163-241).
__device__ __forceinline__ void plus_nbrs(G3 g, int i, int j, int k, int l, int &lx, int &ly,
                                          int &lz)
{
    const int sy = g.nx, sk = g.nx * g.ny;
    lx = (i + 1 < g.nx) ? l + 1 : l - (g.nx - 1);
    ly = (j + 1 < g.ny) ? l + sy : l - sy * (g.ny - 1);
    lz = (k + 1 < g.nz) ? l + sk : l - sk * (g.nz - 1);
}

template <typename T>
struct Vec3 {
    T *x, *y, *z;
};

template <typename T>
struct CVec3 {
    const T *x, *y, *z;
};

// tv.m:145-149: v = grad(f) (gradfp_mex with 1/h), l = 0, and the partial
// sums of |v| for nr.
template <typename T>
__global__ __launch_bounds__(NT) void init_kernel(Vec3<T> v, Vec3<T> lm, const T *f, G3 g,
                                                  T hx, T hy, T hz, double *part)
{
    int i, j, k, l;
    double s[1] = {0.0};
    if (vox(g, i, j, k, l)) {
        int px, py, pz;
        plus_nbrs(g, i, j, k, l, px, py, pz);
        const T c = f[l];
        const T a = hx * (f[px] - c), b = hy * (f[py] - c), d = hz * (f[pz] - c);
        v.x[l] = a;
        v.y[l] = b;
        v.z[l] = d;
        lm.x[l] = T(0);
        lm.y[l] = T(0);
        lm.z[l] = T(0);
        s[0] = double(sqrt((a * a + b * b) + d * d));
    }
    block_sum<1>(s, part + NPART * flat_block());
}

// tv.m:196-200: r = gradAdj(rho v - l) (gradfp_adj_mex with -1/h), the
// differences rho*v - l formed per sample as the reference forms r1..r3.
// Returns at once when the previous update already wrote r at this rho.
template <typename T>
__global__ __launch_bounds__(NT) void rhs_kernel(T *r, CVec3<T> v, CVec3<T> lm, G3 g,
                                                 const Ctl *ctl, T hx, T hy, T hz,
                                                 int ntiles)
{
    if (ctl->stop || ctl->rkeep) return;
    const T rho = T(ctl->rho);
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int i, j, k, l;
        if (!tile_vox(g, t, i, j, k, l)) continue;
        int mx, my, mz;
        minus_nbrs(g, i, j, k, l, mx, my, mz);
        const T X = hx * ((rho * v.x[l] - lm.x[l]) - (rho * v.x[mx] - lm.x[mx]));
        const T Y = hy * ((rho * v.y[l] - lm.y[l]) - (rho * v.y[my] - lm.y[my]));
        const T Z = hz * ((rho * v.z[l] - lm.z[l]) - (rho * v.z[mz] - lm.z[mz]));
        r[l] = adj_sum(X, Y, Z, j, k);
    }
}

template <typename T>
__device__ __forceinline__ T sgn_of(T x)
{
    return x > T(0) ? T(1) : (x < T(0) ? T(-1) : x);  // sign(): 0 -> 0, NaN -> NaN
}

// tv.m:205-252 at one voxel (i, j, k) = l: the gradient d of u (gradfp
// with 1/h), the over-relaxed d1, the shrunk v (new), the multiplier l
// (new) and dv = v_new - v_old, from u and the previous v, l.
template <typename T>
struct Upd {
    T d[3], n[3], m[3], dv[3];
};

template <typename T, bool ISO>
__device__ __forceinline__ Upd<T> update_at(const T *u, CVec3<T> v, CVec3<T> lm, G3 g, int i,
                                            int j, int k, int l, T hx, T hy, T hz, T alpha,
                                            T oma, T rho, T irho)
{
    Upd<T> r;
    int px, py, pz;
    plus_nbrs(g, i, j, k, l, px, py, pz);
    const T c = u[l];
    r.d[0] = hx * (u[px] - c);
    r.d[1] = hy * (u[py] - c);
    r.d[2] = hz * (u[pz] - c);
    const T vo[3] = {v.x[l], v.y[l], v.z[l]};
    const T lo[3] = {lm.x[l], lm.y[l], lm.z[l]};
    // over-relaxation, then the shrinkage argument
    T d1[3], q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        d1[a] = alpha * r.d[a] + oma * vo[a];
        q[a] = d1[a] + irho * lo[a];
    }
    if constexpr (!ISO) {
        // max(abs(r)-irho, 0) .* sign(r)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            r.n[a] = fmax(fabs(q[a]) - irho, T(0)) * sgn_of(q[a]);
    } else {
        // s = (s - irho) .* (s > irho) ./ s, non-finite -> 0
        const T sn = sqrt((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]);
        T f = ((sn - irho) * T(sn > irho ? 1 : 0)) / sn;
        if (!isfinite(f)) f = T(0);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            r.n[a] = q[a] * f;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.m[a] = lo[a] + rho * (d1[a] - r.n[a]);
        r.dv[a] = r.n[a] - vo[a];
    }
    return r;
}

// One ADMM update after the u-subproblem: the v-subproblem and the
// multipliers (tv.m:205-256) and both residuals' partial sums (tv.m:
// 258-277).  Reads u and the previous (v, l), writes the new (vn, ln) to
// separate buffers: the dual terms need dv and the new l at the periodic -1
// neighbours, which each thread recomputes from the unchanged previous
// values (the same operations on the same inputs, so the same bits as the
// neighbour's own update), so dv is never stored.  Partial sums per tile:
// |grad u - v| (nr), |grad u| (na), |v| (nb), |rho gradAdj(dv)|^2 (dual
// residual) and |rho gradAdj(l)|^2 (its scale); gradAdj with the -1/h
// coefficients (ax, ay, az) as gradfp_adj_mex forms it.  Also writes r =
// gradAdj(rho v - l) of the new (v, l) at this rho, bitwise what
// rhs_kernel would form, for the next iteration when the rule keeps rho.
template <typename T, bool ISO>
__global__ __launch_bounds__(NT) void update_kernel(CVec3<T> v, CVec3<T> lm, Vec3<T> vn,
                                                    Vec3<T> ln, T *r, const T *u, G3 g, T hx,
                                                    T hy, T hz, T ax, T ay, T az, T alpha,
                                                    T oma, const Ctl *ctl, double *part,
                                                    int ntiles)
{
    if (ctl->stop) return;
    const T rho = T(ctl->rho), irho = T(ctl->irho);
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int i, j, k, l;
        double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (tile_vox(g, t, i, j, k, l)) {
            int mx, my, mz;
            minus_nbrs(g, i, j, k, l, mx, my, mz);
            const int im = i > 0 ? i - 1 : g.nx - 1, jm = j > 0 ? j - 1 : g.ny - 1,
                      km = k > 0 ? k - 1 : g.nz - 1;
            const Upd<T> c = update_at<T, ISO>(u, v, lm, g, i, j, k, l, hx, hy, hz, alpha,
                                               oma, rho, irho);
            const Upd<T> X = update_at<T, ISO>(u, v, lm, g, im, j, k, mx, hx, hy, hz, alpha,
                                               oma, rho, irho);
            const Upd<T> Y = update_at<T, ISO>(u, v, lm, g, i, jm, k, my, hx, hy, hz, alpha,
                                               oma, rho, irho);
            const Upd<T> Z = update_at<T, ISO>(u, v, lm, g, i, j, km, mz, hx, hy, hz, alpha,
                                               oma, rho, irho);
            vn.x[l] = c.n[0];
            vn.y[l] = c.n[1];
            vn.z[l] = c.n[2];
            ln.x[l] = c.m[0];
            ln.y[l] = c.m[1];
            ln.z[l] = c.m[2];
            const T e1 = c.d[0] - c.n[0], e2 = c.d[1] - c.n[1], e3 = c.d[2] - c.n[2];
            s[0] = double(sqrt((e1 * e1 + e2 * e2) + e3 * e3));
            s[1] = double(sqrt((c.d[0] * c.d[0] + c.d[1] * c.d[1]) + c.d[2] * c.d[2]));
            s[2] = double(sqrt((c.n[0] * c.n[0] + c.n[1] * c.n[1]) + c.n[2] * c.n[2]));
            const T a = rho * adj_sum(ax * (c.dv[0] - X.dv[0]), ay * (c.dv[1] - Y.dv[1]),
                                      az * (c.dv[2] - Z.dv[2]), j, k);
            const T b = rho * adj_sum(ax * (c.m[0] - X.m[0]), ay * (c.m[1] - Y.m[1]),
                                      az * (c.m[2] - Z.m[2]), j, k);
            s[3] = double(a) * double(a);
            s[4] = double(b) * double(b);
            // the next right-hand side at this rho, as rhs_kernel forms it
            r[l] = adj_sum(ax * ((rho * c.n[0] - c.m[0]) - (rho * X.n[0] - X.m[0])),
                           ay * ((rho * c.n[1] - c.m[1]) - (rho * Y.n[1] - Y.m[1])),
                           az * ((rho * c.n[2] - c.m[2]) - (rho * Z.n[2] - Z.m[2])), j, k);
        }
        block_sum<5>(s, part + NPART * t);
    }
}

// First reduction level: block b of RB sums the NPART partials of a
// contiguous range of voxel blocks (a fixed order).
__global__ __launch_bounds__(NT) void reduce1_kernel(const double *part, int nblk, double *out,
                                                     const Ctl *ctl)
{
    if (ctl && ctl->stop) return;
    const int chunk = (nblk + RB - 1) / RB;
    const int b0 = int(blockIdx.x) * chunk, b1 = min(nblk, b0 + chunk);
    double s[NPART];
#pragma unroll
    for (int q = 0; q < NPART; ++q)
        s[q] = 0.0;
    const int t = threadIdx.x + BX * threadIdx.y;
    for (int b = b0 + t; b < b1; b += NT)
#pragma unroll
        for (int q = 0; q < NPART; ++q)
            s[q] += part[NPART * b + q];
    block_sum<NPART>(s, out + NPART * blockIdx.x);
}

__device__ __forceinline__ void sum_level1(const double *p1, double (&tot)[NPART])
{
    double s[NPART];
#pragma unroll
    for (int q = 0; q < NPART; ++q)
        s[q] = 0.0;
    const int t = threadIdx.x + BX * threadIdx.y;
    for (int b = t; b < RB; b += NT)
#pragma unroll
        for (int q = 0; q < NPART; ++q)
            s[q] += p1[NPART * b + q];
    __shared__ double res[NPART];
    block_sum<NPART>(s, res);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NPART; ++q)
        tot[q] = res[q];
}

// tv.m:141-149 scalars: nr = sum |grad f| (the init kernel's sums), rho and
// the counters at their start values.
__global__ __launch_bounds__(NT) void start_kernel(const double *p1, Ctl *ctl, double rho)
{
    double tot[NPART];
    sum_level1(p1, tot);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        ctl->rho = rho;
        ctl->irho = 1.0 / rho;
        ctl->nr = tot[0];
        ctl->flag = 0;
        ctl->stop = 0;
        ctl->iters = 0;
        ctl->rkeep = 0;
    }
}

// tv.m:258-306 after iteration it: the residual norms, the stopping rule
// and the penalty update; writes the trace row (||r||, eps_pri, ||s||,
// eps_dual, rho after the update).
__global__ __launch_bounds__(NT) void decide_kernel(const double *p1, Ctl *ctl, Rule r,
                                                    double *trace, int it)
{
    if (ctl->stop) return;
    double tot[NPART];
    sum_level1(p1, tot);
    if (threadIdx.x != 0 || threadIdx.y != 0) return;
    const double nr_ = ctl->nr;
    const double nr = tot[0], na = tot[1], nb = tot[2];
    const double ns = sqrt(tot[3]);
    const double ep = r.atolp + r.rtol * fmax(na, nb);
    const double ed = r.atold + r.rtol * sqrt(tot[4]);
    double rho = ctl->rho;
    int flag = ctl->flag;
    const bool stop = nr < ep && ns < ed;
    if (!stop) {
        if (nr > r.stau * ns || (nr > r.rtau * nr_ && flag < 3)) {
            rho = r.gamma * rho;
            if (flag > 5)
                flag = 0;
        } else if (nr < ns / r.stau) {
            rho = rho / (r.gamma * r.gamma);
            flag = flag + 1;
        }
    }
    double *row = trace + 5 * (it - 1);
    row[0] = nr;
    row[1] = ep;
    row[2] = ns;
    row[3] = ed;
    row[4] = rho;
    ctl->nr = nr;
    ctl->rkeep = rho == ctl->rho ? 1 : 0;
    ctl->rho = rho;
    ctl->irho = 1.0 / rho;
    ctl->flag = flag;
    ctl->iters = it;
    ctl->stop = stop ? 1 : 0;
}

// tv.m:308: x = mask .* u.
template <typename T>
__global__ __launch_bounds__(NT) void mask_kernel(T *x, const T *u, const uint8_t *m, G3 g)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    x[l] = T(m[l] ? 1 : 0) * u[l];
}

template <typename T>
__global__ __launch_bounds__(NT) void zero_kernel(T *x, G3 g)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    x[l] = T(0);
}

// The forward transform of f is kept as it comes from the x pass (doubled).
template <typename T>
struct KeepOp {
    __device__ Cx<T> operator()(int, int, int, Cx<T> v) const { return v; }
};

// k-space operation of the u-subproblem at half-spectrum sample (x, y, z).
// tv.m:159-174, :203-204: iA(k) = 1/(mu D(k)^2 + rho L(k)) (0 where not
// finite) and u = real(ifft3(iA .* (mu conj(D) fft3(f) + fft3(r)))).
// rts.m:145-157, :203-204 (RTS): M(k) = mu [|D(k)| > delta] (D in the class
// of f), iA(k) = 1/(M(k) + rho L(k)) and x = real(ifft3(iA .* (M fft3(x1) +
// fft3(r)))).  With the source s (f or x1) and r real, that is the inverse
// transform of A1s fft3(s) + A2s fft3(r), A1 = iA mu D (tv) or iA M (rts),
// A2 = iA, As(k) = (A(k) + A(-k))/2; As = A except on the Nyquist plane of
// an even axis with a tilted B0 (D differs at k and -k there; L does not).
// L(k) = sum_a (2 - 2 cos(2 pi m_a / n_a)) / h_a^2 is the symbol of
// -laplaceKernel (laplaceKernel.m:14-24) padded and circularly shifted to
// the origin (tv.m:163-165), with L(0) = 0 exactly.  The k coordinates and
// the per-axis terms of L come from tables (tab: k and L along x for H
// entries, then y for M1, then z for M2).  The weights are formed in the
// class the reference forms iA in: double, except for tv.m with a single
// f, where K and iA are single; then rounded once to T with the transform
// scale.
template <typename T, bool RTS>
struct SolveOp {
    const Cx<T> *F;    // 2 fft3(source) on the half spectrum
    const double *tab;
    const Ctl *ctl;
    int M0, M1, M2, H;
    double b0, b1, b2, mu, delta, scale;

    __device__ void weights(double D, double L, double rho, double &a1, double &a2) const
    {
        double K, g;
        if constexpr (RTS) {
            K = fabs(T(D)) > T(delta) ? mu : 0.0;
            g = K;
        } else {
            K = mu * D * D;
            g = mu * D;
        }
        double iA = 1.0 / (K + rho * L);
        if (!isfinite(iA)) iA = 0.0;
        a1 = iA * g;
        a2 = iA;
    }

    // tv.m in single (tv.m:89, 159-168): D formed in double and cast to
    // single, K = mu |D|^2 and iA = 1 ./ (K + rho L) in single (rho L
    // formed in double, then cast), data factor mu conj(D) in single.
    __device__ void weights_single(float D, double L, double rho, float &a1, float &a2) const
    {
        const float K = float(mu) * (D * D);
        float iA = 1.0f / (K + float(rho * L));
        if (!isfinite(iA)) iA = 0.0f;
        a1 = iA * (float(mu) * D);
        a2 = iA;
    }

    __device__ Cx<T> operator()(int x, int y, int z, Cx<T> fr) const
    {
        const double *ty = tab + 2 * H, *tz = ty + 2 * M1;
        const double X = tab[x], Y = ty[y], Z = tz[z];
        const double L = (tab[H + x] + ty[M1 + y]) + tz[M2 + z];
        const double rho = ctl->rho;
        const bool nx = 2 * x == M0, ny = 2 * y == M1, nz = 2 * z == M2;
        double a1, a2;
        if constexpr (!RTS && sizeof(T) == 4) {
            // rts.m's M = mu (|D| > delta) is double, so only tv.m's
            // weights follow the single class
            float s1, s2;
            weights_single(float(dip(X, Y, Z, b0, b1, b2)), L, rho, s1, s2);
            if (nx || ny || nz) {
                float m1, m2;
                weights_single(float(dip(nx ? X : -X, ny ? Y : -Y, nz ? Z : -Z, b0, b1, b2)), L,
                               rho, m1, m2);
                s1 = 0.5f * (s1 + m1);
                s2 = 0.5f * (s2 + m2);
            }
            a1 = s1;
            a2 = s2;
        } else {
            weights(dip(X, Y, Z, b0, b1, b2), L, rho, a1, a2);
            if (nx || ny || nz) {
                double m1, m2;
                weights(dip(nx ? X : -X, ny ? Y : -Y, nz ? Z : -Z, b0, b1, b2), L, rho, m1,
                        m2);
                a1 = 0.5 * (a1 + m1);
                a2 = 0.5 * (a2 + m2);
            }
        }
        const T c1 = T(a1 * scale), c2 = T(a2 * scale);
        const Cx<T> ff = F[size_t(x) + size_t(H) * (size_t(y) + size_t(M1) * z)];
        return Cx<T>{c1 * ff.re + c2 * fr.re, c1 * ff.im + c2 * fr.im};
    }
};

// Per-axis table of the solve: k coordinate (dipoleKernel.m grid followed
// by ifftshift, the same formula as kcoord) then (2 - 2 cos(2 pi m/n))/h^2
// (evaluated at min(m, n - m), so that m and n - m agree exactly).
void axis_table(double *k, double *L, int count, int n, double h)
{
    const double dk = 1.0 / (double(n) * h), ih2 = 1.0 / (h * h);
    const int low = n / 2;
    for (int m = 0; m < count; ++m) {
        k[m] = double(m < n - low ? m : m - n) * dk;
        const int mm = std::min(m, n - m);
        const long double a = 2.0L * 3.141592653589793238462643383279502884L * mm / n;
        L[m] = mm == 0 ? 0.0 : double((2.0L - 2.0L * cosl(a)) * ih2);
    }
}

// Work space, carved from one device buffer (256-byte aligned pieces).
template <typename T>
struct Work {
    Cx<T> *F, *W;
    T *r, *u;
    Vec3<T> v, l, v2, l2;  // (v, l) and the second set the update writes
    double *part, *p1, *tab, *trace;
    Ctl *ctl;
    // rts step 1 (LSMR on the full spectrum): u, v, h, hbar, x and D
    Cx<T> *lu, *lv, *lh, *lhb, *lx;
    T *D;
};

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// Grid of a tile-walking kernel: each block walks TPB tiles, gridDim.x
// apart, so that a launch that returns at once (rule stopped, r current)
// dispatches TPB times fewer blocks.
constexpr int TPB = 4;
int tile_grid(int ntiles) { return (ntiles + TPB - 1) / TPB; }

template <typename T>
size_t work_layout(const size_t sz[3], Work<T> *w, char *base, bool lsmr = false)
{
    const size_t N = sz[0] * sz[1] * sz[2];
    const size_t M[3] = {sz[0], sz[1], sz[2]};
    const size_t spec = align256(real_conv_work_bytes(M, sizeof(T)));
    const size_t vol = align256(N * sizeof(T));
    const G3 g{int(sz[0]), int(sz[1]), int(sz[2])};
    const size_t part = align256(size_t(vox_blocks(g)) * NPART * sizeof(double));
    size_t off = 0;
    auto take = [&](size_t n) { char *p = base ? base + off : nullptr; off += n; return p; };
    char *pF = take(spec), *pW = take(spec), *pr = take(vol), *pu = take(vol);
    char *pv[3], *pl[3], *pv2[3], *pl2[3];
    for (auto &p : pv) p = take(vol);
    for (auto &p : pl) p = take(vol);
    for (auto &p : pv2) p = take(vol);
    for (auto &p : pl2) p = take(vol);
    char *pp = take(part), *p1 = take(align256(RB * NPART * sizeof(double)));
    char *ptab = take(align256(2 * (sz[0] / 2 + 1 + sz[1] + sz[2]) * sizeof(double)));
    char *ptr = take(align256(size_t(MAX_ITERS) * 5 * sizeof(double)));
    char *pc = take(256);
    char *pls[5] = {nullptr, nullptr, nullptr, nullptr, nullptr}, *pD = nullptr;
    if (lsmr) {
        for (auto &p : pls) p = take(align256(N * sizeof(Cx<T>)));
        pD = take(vol);
    }
    if (w) {
        w->lu = reinterpret_cast<Cx<T> *>(pls[0]);
        w->lv = reinterpret_cast<Cx<T> *>(pls[1]);
        w->lh = reinterpret_cast<Cx<T> *>(pls[2]);
        w->lhb = reinterpret_cast<Cx<T> *>(pls[3]);
        w->lx = reinterpret_cast<Cx<T> *>(pls[4]);
        w->D = reinterpret_cast<T *>(pD);
        w->F = reinterpret_cast<Cx<T> *>(pF);
        w->W = reinterpret_cast<Cx<T> *>(pW);
        w->r = reinterpret_cast<T *>(pr);
        w->u = reinterpret_cast<T *>(pu);
        auto vec = [](char *const *q) {
            return Vec3<T>{reinterpret_cast<T *>(q[0]), reinterpret_cast<T *>(q[1]),
                           reinterpret_cast<T *>(q[2])};
        };
        w->v = vec(pv);
        w->l = vec(pl);
        w->v2 = vec(pv2);
        w->l2 = vec(pl2);
        w->part = reinterpret_cast<double *>(pp);
        w->p1 = reinterpret_cast<double *>(p1);
        w->tab = reinterpret_cast<double *>(ptab);
        w->trace = reinterpret_cast<double *>(ptr);
        w->ctl = reinterpret_cast<Ctl *>(pc);
    }
    return off;
}

// The ADMM parameters shared by tv.m and rts.m (step 2).
struct AdmmParams {
    double mu, delta;  // delta: rts only
    int tv, maxit, verbose;
    double rtol, atol, rho, alpha, rtau, stau, gamma;
};

// B0 direction normalised as dipoleKernel.m:35-38 does.
int unit_b(const double bdir[3], double b[3])
{
    const double bn = std::sqrt(bdir[0] * bdir[0] + bdir[1] * bdir[1] + bdir[2] * bdir[2]);
    QSM_ARG(bn > 0.0, "B0 direction must be non-zero");
    for (int a = 0; a < 3; ++a) b[a] = bn != 1.0 ? bdir[a] / bn : bdir[a];
    return QSM_OK;
}

// The ADMM loop of tv.m:141-308 (RTS = false: source f) and rts.m:145-308
// (RTS = true: source x1): F from the source, u = source, v = grad(u), l = 0,
// then at most maxit iterations decided on the device, and x = mask .* u.
// src may alias w.u.
template <typename T, bool RTS>
int admm_run(const T *src, const uint8_t *mask, const size_t sz[3], const double vsz[3],
             const double b[3], const AdmmParams &o, T *x, Work<T> &w, double *trace,
             int *iters, hipStream_t st)
{
    int rc;
    const G3 g{int(sz[0]), int(sz[1]), int(sz[2])};
    const size_t N = sz[0] * sz[1] * sz[2];
    const dim3 grid = vox_grid(g), blk(BX, BY);
    const int nblk = vox_blocks(g);

    // gradient coefficients as gradfp_mex / gradfp_adj_mex form them
    const T gx = T(1.0 / vsz[0]), gy = T(1.0 / vsz[1]), gz = T(1.0 / vsz[2]);
    const T ax = T(-1.0 / vsz[0]), ay = T(-1.0 / vsz[1]), az = T(-1.0 / vsz[2]);

    // per-axis tables of the solve
    const int H = g.nx / 2 + 1;
    std::vector<double> tab(2 * (H + g.ny + g.nz));
    axis_table(tab.data(), tab.data() + H, H, g.nx, vsz[0]);
    axis_table(tab.data() + 2 * H, tab.data() + 2 * H + g.ny, g.ny, g.ny, vsz[1]);
    axis_table(tab.data() + 2 * H + 2 * g.ny, tab.data() + 2 * H + 2 * g.ny + g.nz, g.nz, g.nz,
               vsz[2]);
    QSM_HIP(hipMemcpyAsync(w.tab, tab.data(), tab.size() * sizeof(double),
                           hipMemcpyHostToDevice, st));

    const size_t M[3] = {sz[0], sz[1], sz[2]}, off[3] = {0, 0, 0};
    // F: 2 fft3(source) on the half spectrum (the data-term factor is
    // applied in the solve)
    if ((rc = real_conv<T, false>(src, static_cast<T *>(nullptr), sz, M, off, w.F, KeepOp<T>{},
                                  st)))
        return rc;

    SolveOp<T, RTS> op;
    op.F = w.F;
    op.tab = w.tab;
    op.ctl = w.ctl;
    op.M0 = g.nx;
    op.M1 = g.ny;
    op.M2 = g.nz;
    op.H = H;
    op.b0 = b[0];
    op.b1 = b[1];
    op.b2 = b[2];
    op.mu = o.mu;
    op.delta = o.delta;
    // 1/(nx ny nz) of ifft3, 1/2 for the x pass's doubling
    op.scale = 0.5 / (double(g.nx) * double(g.ny) * double(g.nz));

    // tv.m:127-130 stopping tolerances, :279-306 penalty update
    const Rule rule{std::sqrt(3.0 * double(N)) * o.atol, std::sqrt(double(N)) * o.atol, o.rtol,
                    o.rtau, o.stau, o.gamma};

    // v = grad(u), l = 0, nr = sum |v| (tv.m:141-149)
    init_kernel<T><<<grid, blk, 0, st>>>(w.v, w.l, src, g, gx, gy, gz, w.part);
    QSM_HIP(hipGetLastError());
    reduce1_kernel<<<RB, blk, 0, st>>>(w.part, nblk, w.p1, nullptr);
    QSM_HIP(hipGetLastError());
    start_kernel<<<1, blk, 0, st>>>(w.p1, w.ctl, o.rho);
    QSM_HIP(hipGetLastError());

    const int rgrid = tile_grid(nblk), ugrid = rgrid;

    // (v, l) alternate between the two sets: the update reads one, writes
    // the other.  Once the rule stops the loop every later launch returns
    // at once, so which set holds the last (v, l) is immaterial.
    Vec3<T> vs[2] = {w.v, w.v2}, ls[2] = {w.l, w.l2};
    const int *stop = &w.ctl->stop;
    for (int it = 1; it <= o.maxit; ++it) {
        const int p = (it - 1) & 1;
        const CVec3<T> cv{vs[p].x, vs[p].y, vs[p].z}, cl{ls[p].x, ls[p].y, ls[p].z};
        // u-subproblem (tv.m:196-203)
        rhs_kernel<T><<<rgrid, blk, 0, st>>>(w.r, cv, cl, g, w.ctl, ax, ay, az, nblk);
        QSM_HIP(hipGetLastError());
        if ((rc = real_conv<T, true>(w.r, w.u, sz, M, off, w.W, op, st, stop))) return rc;
        // v-subproblem, multipliers, primal and dual sums (tv.m:205-277)
        if (o.tv == 1)
            update_kernel<T, false><<<ugrid, blk, 0, st>>>(cv, cl, vs[p ^ 1], ls[p ^ 1], w.r,
                                                          w.u, g, gx, gy, gz, ax, ay, az,
                                                          T(o.alpha), T(1.0 - o.alpha), w.ctl,
                                                          w.part, nblk);
        else
            update_kernel<T, true><<<ugrid, blk, 0, st>>>(cv, cl, vs[p ^ 1], ls[p ^ 1], w.r,
                                                         w.u, g, gx, gy, gz, ax, ay, az,
                                                         T(o.alpha), T(1.0 - o.alpha), w.ctl,
                                                         w.part, nblk);
        QSM_HIP(hipGetLastError());
        // the rule (tv.m:279-306)
        reduce1_kernel<<<RB, blk, 0, st>>>(w.part, nblk, w.p1, w.ctl);
        QSM_HIP(hipGetLastError());
        decide_kernel<<<1, blk, 0, st>>>(w.p1, w.ctl, rule, w.trace, it);
        QSM_HIP(hipGetLastError());
    }
    // x = mask .* u (tv.m:308)
    mask_kernel<T><<<grid, blk, 0, st>>>(x, w.u, mask, g);
    QSM_HIP(hipGetLastError());

    Ctl c;
    QSM_HIP(hipMemcpyAsync(&c, w.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    QSM_HIP(hipStreamSynchronize(st));
    std::vector<double> tr(size_t(c.iters) * 5);
    if (c.iters > 0)
        QSM_HIP(hipMemcpy(tr.data(), w.trace, tr.size() * sizeof(double), hipMemcpyDeviceToHost));
    if (o.verbose) {
        std::printf("\niter\t\t||r||\t\teps_pri\t\t||s||\t\teps_dual\n");
        for (int i = 0; i < c.iters; ++i)
            std::printf("%3d/%d\t\t%.4e\t%.4e\t%.4e\t%.4e\n", i + 1, o.maxit, tr[5 * i],
                        tr[5 * i + 1], tr[5 * i + 2], tr[5 * i + 3]);
        std::printf("\n");
    }
    if (trace && c.iters > 0)
        std::copy(tr.begin(), tr.end(), trace);
    if (iters)
        *iters = c.iters;
    return QSM_OK;
}

template <typename T>
int tv_dev_t(const T *f, const uint8_t *mask, const size_t sz[3], const double vsz[3],
             const double bdir[3], const qsm_tv_opts &o, T *x, void *work, double *trace,
             int *iters, hipStream_t st)
{
    double b[3];
    int rc = unit_b(bdir, b);
    if (rc) return rc;
    Work<T> w;
    work_layout<T>(sz, &w, static_cast<char *>(work));
    const AdmmParams p{o.mu, 0.0, o.tv, o.maxit, o.verbose, o.rtol, o.atol,
                       o.rho, o.alpha, o.rtau, o.stau, o.gamma};
    return admm_run<T, false>(f, mask, sz, vsz, b, p, x, w, trace, iters, st);
}

// ---- rts.m step 1: LSMR (third_party/lsmr/lsmr.m) with A = diag(D) ------

constexpr int LT = 256;  // threads per block of the 1-D LSMR kernels

__device__ __forceinline__ double abs2(Cx<float> a)
{
    return double(a.re) * double(a.re) + double(a.im) * double(a.im);
}
__device__ __forceinline__ double abs2(Cx<double> a) { return a.re * a.re + a.im * a.im; }

// Sum of one double per thread of a LT-thread block into out[0] (slot 0 of
// the block's NPART partials), fixed order.
__device__ __forceinline__ void block_sum1(double v, double *out)
{
    __shared__ double red[LT];
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = LT / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) out[0] = red[0];
}

// dipoleKernel.m:56-66 on the full grid, cast to T: D(k) for every k of the
// solver grid (x fastest), and u = (f, 0).
template <typename T>
__global__ __launch_bounds__(NT) void lsmr_setup_kernel(T *D, Cx<T> *u, const T *f, G3 g,
                                                        double dk0, double dk1, double dk2,
                                                        double b0, double b1, double b2)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    D[l] = T(dip(kcoord(i, g.nx, dk0), kcoord(j, g.ny, dk1), kcoord(k, g.nz, dk2), b0, b1, b2));
    u[l] = Cx<T>{f[l], T(0)};
}

// lsmr.m: u = A(v) - alpha u (mode 1 with alpha; the first call, alpha = 0,
// only measures u = b), partial sums of |u|^2.
template <typename T>
__global__ __launch_bounds__(LT) void lsmr_u_kernel(Cx<T> *u, const Cx<T> *v, const T *D,
                                                    size_t n, T alpha, int with_av,
                                                    double *part)
{
    const size_t l = size_t(blockIdx.x) * LT + threadIdx.x;
    double s = 0.0;
    if (l < n) {
        Cx<T> a = u[l];
        if (with_av) {
            const T d = D[l];
            const Cx<T> b = v[l];
            a = Cx<T>{d * b.re - alpha * a.re, d * b.im - alpha * a.im};
            u[l] = a;
        }
        s = abs2(a);
    }
    block_sum1(s, part + NPART * blockIdx.x);
}

// lsmr.m: u = (1/beta) u; v = A'(u) - beta v (beta_v = 0 on the first
// call: v = A'(u)), partial sums of |v|^2.
template <typename T>
__global__ __launch_bounds__(LT) void lsmr_v_kernel(Cx<T> *u, Cx<T> *v, const T *D, size_t n,
                                                    T ibeta, T beta_v, double *part)
{
    const size_t l = size_t(blockIdx.x) * LT + threadIdx.x;
    double s = 0.0;
    if (l < n) {
        const Cx<T> a = u[l];
        const Cx<T> un{ibeta * a.re, ibeta * a.im};
        u[l] = un;
        const T d = D[l];
        Cx<T> vn{d * un.re, d * un.im};
        if (beta_v != T(0)) {
            const Cx<T> b = v[l];
            vn = Cx<T>{vn.re - beta_v * b.re, vn.im - beta_v * b.im};
        }
        v[l] = vn;
        s = abs2(vn);
    }
    block_sum1(s, part + NPART * blockIdx.x);
}

// lsmr.m: v = (1/alpha) v (ialpha = 1 when alpha = 0); first call (init):
// h = v, hbar = 0, x = 0; later: hbar = h - c1 hbar; x = x + c2 hbar;
// h = v - c3 h.  Partial sums of |x|^2.
template <typename T>
__global__ __launch_bounds__(LT) void lsmr_x_kernel(Cx<T> *v, Cx<T> *h, Cx<T> *hb, Cx<T> *x,
                                                    size_t n, T ialpha, int init, T c1, T c2,
                                                    T c3, double *part)
{
    const size_t l = size_t(blockIdx.x) * LT + threadIdx.x;
    double s = 0.0;
    if (l < n) {
        const Cx<T> a = v[l];
        const Cx<T> vn{ialpha * a.re, ialpha * a.im};
        v[l] = vn;
        if (init) {
            h[l] = vn;
            hb[l] = Cx<T>{T(0), T(0)};
            x[l] = Cx<T>{T(0), T(0)};
        } else {
            const Cx<T> hh = h[l], hbo = hb[l], xo = x[l];
            const Cx<T> hbn{hh.re - c1 * hbo.re, hh.im - c1 * hbo.im};
            const Cx<T> xn{xo.re + c2 * hbn.re, xo.im + c2 * hbn.im};
            hb[l] = hbn;
            x[l] = xn;
            h[l] = Cx<T>{vn.re - c3 * hh.re, vn.im - c3 * hh.im};
            s = abs2(xn);
        }
    }
    block_sum1(s, part + NPART * blockIdx.x);
}

// rts.m:139: x1 = mask .* real(ifft3(x)).
template <typename T>
__global__ __launch_bounds__(NT) void real_mask_kernel(T *x1, const Cx<T> *x, const uint8_t *m,
                                                       G3 g)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    x1[l] = T(m[l] ? 1 : 0) * x[l].re;
}

// Total of slot 0 over nblk blocks' partials: two-level on the device, the
// RB first-level results summed on the host in a fixed order.
int lsmr_total(const double *part, int nblk, double *p1, double &out, hipStream_t st)
{
    reduce1_kernel<<<RB, dim3(BX, BY), 0, st>>>(part, nblk, p1, nullptr);
    QSM_HIP(hipGetLastError());
    double h[RB * NPART];
    QSM_HIP(hipMemcpyAsync(h, p1, sizeof(h), hipMemcpyDeviceToHost, st));
    QSM_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (int b = 0; b < RB; ++b) s += h[NPART * b];
    out = s;
    return QSM_OK;
}

// lsmr(@(x,~) D.*x, fft3(f), 0, eps, eps, [], itnlim) (rts.m:131-139):
// Golub-Kahan bidiagonalisation with A = A' = diag(D) on the full spectrum
// (vector updates and norms on the device, the scalar recurrence of
// lsmr.m on the host in double), then x1 = mask .* real(ifft3(x)).
template <typename T>
int rts_step1(const T *f, const uint8_t *mask, const size_t sz[3], const double vsz[3],
              const double b[3], int itnlim, T *x1, Work<T> &w, hipStream_t st)
{
    int rc;
    const G3 g{int(sz[0]), int(sz[1]), int(sz[2])};
    const size_t n = sz[0] * sz[1] * sz[2];
    const dim3 grid = vox_grid(g), blk(BX, BY);
    const int nb = int((n + LT - 1) / LT);
    const double eps = std::numeric_limits<T>::epsilon();
    lsmr_setup_kernel<T><<<grid, blk, 0, st>>>(w.D, w.lu, f, g, 1.0 / (double(g.nx) * vsz[0]),
                                               1.0 / (double(g.ny) * vsz[1]),
                                               1.0 / (double(g.nz) * vsz[2]), b[0], b[1], b[2]);
    QSM_HIP(hipGetLastError());
    if ((rc = fft3_inplace<T>(w.lu, sz, 0, st))) return rc;  // u = b = fft3(f)

    // u = b; beta = ||u||; u = u/beta; v = A'u; alpha = ||v||; v = v/alpha
    double beta, alpha, x2;
    lsmr_u_kernel<T><<<nb, LT, 0, st>>>(w.lu, w.lv, w.D, n, T(0), 0, w.part);
    QSM_HIP(hipGetLastError());
    if ((rc = lsmr_total(w.part, nb, w.p1, beta, st))) return rc;
    beta = std::sqrt(beta);
    lsmr_v_kernel<T><<<nb, LT, 0, st>>>(w.lu, w.lv, w.D, n, T(beta > 0 ? 1.0 / beta : 1.0), T(0),
                                        w.part);
    QSM_HIP(hipGetLastError());
    if ((rc = lsmr_total(w.part, nb, w.p1, alpha, st))) return rc;
    alpha = std::sqrt(alpha);
    lsmr_x_kernel<T><<<nb, LT, 0, st>>>(w.lv, w.lh, w.lhb, w.lx, n,
                                        T(alpha > 0 ? 1.0 / alpha : 1.0), 1, T(0), T(0), T(0),
                                        w.part);
    QSM_HIP(hipGetLastError());

    // scalar state (lsmr.m, lambda = 0, atol = btol = eps, conlim = 1e8)
    const double atol = eps, btol = eps, ctol = 1.0 / 1e8;
    double zetabar = alpha * beta, alphabar = alpha, rho = 1, rhobar = 1, cbar = 1, sbar = 0;
    double betadd = beta, betad = 0, rhodold = 1, tautildeold = 0, thetatilde = 0, zeta = 0;
    double d = 0, normA2 = alpha * alpha, maxrbar = 0, minrbar = 1e100;
    const double normb = beta;
    if (alpha * beta != 0.0) {  // otherwise the exact solution is x = 0
        for (int itn = 1; itn <= itnlim; ++itn) {
            lsmr_u_kernel<T><<<nb, LT, 0, st>>>(w.lu, w.lv, w.D, n, T(alpha), 1, w.part);
            QSM_HIP(hipGetLastError());
            if ((rc = lsmr_total(w.part, nb, w.p1, beta, st))) return rc;
            beta = std::sqrt(beta);
            if (beta > 0) {
                lsmr_v_kernel<T><<<nb, LT, 0, st>>>(w.lu, w.lv, w.D, n, T(1.0 / beta), T(beta),
                                                    w.part);
                QSM_HIP(hipGetLastError());
                if ((rc = lsmr_total(w.part, nb, w.p1, alpha, st))) return rc;
                alpha = std::sqrt(alpha);
            }
            // plane rotations (lambda = 0: alphahat = alphabar)
            const double alphahat = alphabar, chat = 1.0, shat = 0.0;
            const double rhoold = rho;
            rho = std::hypot(alphahat, beta);
            const double c = alphahat / rho, s = beta / rho;
            const double thetanew = s * alpha;
            alphabar = c * alpha;
            const double rhobarold = rhobar, zetaold = zeta;
            const double thetabar = sbar * rho, rhotemp = cbar * rho;
            rhobar = std::hypot(cbar * rho, thetanew);
            cbar = cbar * rho / rhobar;
            sbar = thetanew / rhobar;
            zeta = cbar * zetabar;
            zetabar = -sbar * zetabar;
            // hbar, x, h (with the new v scaled by 1/alpha first)
            const double c1 = thetabar * rho / (rhoold * rhobarold);
            const double c2 = zeta / (rho * rhobar);
            const double c3 = thetanew / rho;
            lsmr_x_kernel<T><<<nb, LT, 0, st>>>(w.lv, w.lh, w.lhb, w.lx, n,
                                                T(beta > 0 && alpha > 0 ? 1.0 / alpha : 1.0), 0,
                                                T(c1), T(c2), T(c3), w.part);
            QSM_HIP(hipGetLastError());
            if ((rc = lsmr_total(w.part, nb, w.p1, x2, st))) return rc;
            // estimate of ||r||
            const double betaacute = chat * betadd, betacheck = -shat * betadd;
            const double betahat = c * betaacute;
            betadd = -s * betaacute;
            const double thetatildeold = thetatilde;
            const double rhotildeold = std::hypot(rhodold, thetabar);
            const double ctildeold = rhodold / rhotildeold, stildeold = thetabar / rhotildeold;
            thetatilde = stildeold * rhobar;
            rhodold = ctildeold * rhobar;
            betad = -stildeold * betad + ctildeold * betahat;
            tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold;
            const double taud = (zeta - thetatilde * tautildeold) / rhodold;
            d = d + betacheck * betacheck;
            const double normr = std::sqrt(d + (betad - taud) * (betad - taud) + betadd * betadd);
            // estimates of ||A||, cond(A), ||A'r||
            normA2 = normA2 + beta * beta;
            const double normA = std::sqrt(normA2);
            normA2 = normA2 + alpha * alpha;
            maxrbar = std::max(maxrbar, rhobarold);
            if (itn > 1) minrbar = std::min(minrbar, rhobarold);
            const double condA = std::max(maxrbar, rhotemp) / std::min(minrbar, rhotemp);
            const double normAr = std::fabs(zetabar), normx = std::sqrt(x2);
            // stopping tests
            const double test1 = normr / normb, test2 = normAr / (normA * normr);
            const double test3 = 1.0 / condA;
            const double t1 = test1 / (1 + normA * normx / normb);
            const double rtol = btol + atol * normA * normx / normb;
            int istop = 0;
            if (itn >= itnlim) istop = 7;
            if (1 + test3 <= 1) istop = 6;
            if (1 + test2 <= 1) istop = 5;
            if (1 + t1 <= 1) istop = 4;
            if (test3 <= ctol) istop = 3;
            if (test2 <= atol) istop = 2;
            if (test1 <= rtol) istop = 1;
            if (istop > 0) break;
        }
        if ((rc = fft3_inplace<T>(w.lx, sz, 1, st))) return rc;
        real_mask_kernel<T><<<grid, blk, 0, st>>>(x1, w.lx, mask, g);
    } else {
        zero_kernel<T><<<grid, blk, 0, st>>>(x1, g);
    }
    QSM_HIP(hipGetLastError());
    return QSM_OK;
}

template <typename T>
int rts_dev_t(const T *f, const uint8_t *mask, const size_t sz[3], const double vsz[3],
              const double bdir[3], const qsm_rts_opts &o, T *x, void *work, double *trace,
              int *iters, hipStream_t st)
{
    double b[3];
    int rc = unit_b(bdir, b);
    if (rc) return rc;
    Work<T> w;
    work_layout<T>(sz, &w, static_cast<char *>(work), true);
    // step 1 into the ADMM's u buffer, the source of step 2
    if ((rc = rts_step1<T>(f, mask, sz, vsz, b, o.tolls, w.u, w, st))) return rc;
    const AdmmParams p{o.mu, o.delta, o.tv, o.maxit, o.verbose, o.rtol, o.atol,
                       o.rho, o.alpha, o.rtau, o.stau, o.gamma};
    return admm_run<T, true>(w.u, mask, sz, vsz, b, p, x, w, trace, iters, st);
}

template <typename O>
int check_admm(int dtype, const size_t sz[3], const double vsz[3], const double bdir[3],
               const O *o)
{
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    QSM_ARG(sz && vsz && bdir && o, "null pointer");
    int rc = check_grid(sz, vsz, 1);
    if (rc) return rc;
    QSM_ARG(o->tv == 1 || o->tv == 2, "opts.tv must be 1 (anisotropic) or 2 (isotropic)");
    QSM_ARG(o->maxit >= 1 && o->maxit <= MAX_ITERS, "opts.maxit must be in [1, %d]", MAX_ITERS);
    QSM_ARG(o->rho > 0.0 && o->gamma > 0.0 && o->stau > 0.0, "opts.rho, gamma, stau must be > 0");
    QSM_ARG(std::isfinite(o->mu) && std::isfinite(o->alpha), "opts.mu and alpha must be finite");
    return QSM_OK;
}

int check_tv(int dtype, const size_t sz[3], const double vsz[3], const double bdir[3],
             const qsm_tv_opts *o)
{
    return check_admm(dtype, sz, vsz, bdir, o);
}

int check_rts(int dtype, const size_t sz[3], const double vsz[3], const double bdir[3],
              const qsm_rts_opts *o)
{
    int rc = check_admm(dtype, sz, vsz, bdir, o);
    if (rc) return rc;
    QSM_ARG(o->tolls >= 0, "opts.tolls must be >= 0");
    QSM_ARG(std::isfinite(o->delta), "opts.delta must be finite");
    return QSM_OK;
}

}  // namespace

extern "C" {

void qsm_tv_opts_default(qsm_tv_opts *o)
{
    // tv.m:38-86
    o->mu = 7.5e3;
    o->tv = 1;
    o->rtol = 1e-1;
    o->atol = 1e-4;
    o->maxit = 20;
    o->verbose = 1;
    o->rho = 9.0;
    o->alpha = 1.5;
    o->rtau = 0.7;
    o->stau = 1e2;
    o->gamma = 3.0;
}

size_t qsm_tv_work_bytes(int dtype, const size_t sz[3])
{
    if (!sz) return 0;
    return dtype == QSM_F32 ? work_layout<float>(sz, nullptr, nullptr)
                            : work_layout<double>(sz, nullptr, nullptr);
}

int qsm_tv_dev(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
               const double vsz[3], const double bdir[3], const qsm_tv_opts *opts, void *x,
               void *work, double *trace, int *iters, void *stream)
{
    int rc = check_tv(dtype, sz, vsz, bdir, opts);
    if (rc) return rc;
    QSM_ARG(f && mask && x && work, "null pointer");
    const hipStream_t st = as_stream(stream);
    return dtype == QSM_F64
        ? tv_dev_t<double>(static_cast<const double *>(f), mask, sz, vsz, bdir, *opts,
                           static_cast<double *>(x), work, trace, iters, st)
        : tv_dev_t<float>(static_cast<const float *>(f), mask, sz, vsz, bdir, *opts,
                          static_cast<float *>(x), work, trace, iters, st);
}

int qsm_tv(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
           const double vsz[3], const double bdir[3], const qsm_tv_opts *opts, void *x,
           double *trace, int *iters)
{
    int rc = check_tv(dtype, sz, vsz, bdir, opts);
    if (rc) return rc;
    QSM_ARG(f && mask && x, "null pointer");
    if ((rc = ensure_device())) return rc;
    const size_t N = sz[0] * sz[1] * sz[2], es = dtype_size(dtype);
    DevBuf df, dm, dx, dw;
    if ((rc = df.alloc(N * es)) || (rc = dm.alloc(N)) || (rc = dx.alloc(N * es))
        || (rc = dw.alloc(qsm_tv_work_bytes(dtype, sz))))
        return rc;
    QSM_HIP(hipMemcpy(df.p, f, N * es, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(dm.p, mask, N, hipMemcpyHostToDevice));
    if ((rc = qsm_tv_dev(dtype, df.p, static_cast<const uint8_t *>(dm.p), sz, vsz, bdir, opts,
                         dx.p, dw.p, trace, iters, nullptr)))
        return rc;
    QSM_HIP(hipMemcpy(x, dx.p, N * es, hipMemcpyDeviceToHost));
    return QSM_OK;
}

void qsm_rts_opts_default(qsm_rts_opts *o)
{
    // rts.m:50-102
    o->tolls = 4;
    o->mu = 1e5;
    o->delta = 0.15;
    o->tv = 1;
    o->rtol = 1e-1;
    o->atol = 1e-2;
    o->maxit = 20;
    o->verbose = 1;
    o->rho = 10.0;
    o->alpha = 1.5;
    o->rtau = 0.7;
    o->stau = 1e2;
    o->gamma = 3.0;
}

size_t qsm_rts_work_bytes(int dtype, const size_t sz[3])
{
    if (!sz) return 0;
    return dtype == QSM_F32 ? work_layout<float>(sz, nullptr, nullptr, true)
                            : work_layout<double>(sz, nullptr, nullptr, true);
}

int qsm_rts_dev(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
                const double vsz[3], const double bdir[3], const qsm_rts_opts *opts, void *x,
                void *work, double *trace, int *iters, void *stream)
{
    int rc = check_rts(dtype, sz, vsz, bdir, opts);
    if (rc) return rc;
    QSM_ARG(f && mask && x && work, "null pointer");
    const hipStream_t st = as_stream(stream);
    return dtype == QSM_F64
        ? rts_dev_t<double>(static_cast<const double *>(f), mask, sz, vsz, bdir, *opts,
                            static_cast<double *>(x), work, trace, iters, st)
        : rts_dev_t<float>(static_cast<const float *>(f), mask, sz, vsz, bdir, *opts,
                           static_cast<float *>(x), work, trace, iters, st);
}

int qsm_rts(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
            const double vsz[3], const double bdir[3], const qsm_rts_opts *opts, void *x,
            double *trace, int *iters)
{
    int rc = check_rts(dtype, sz, vsz, bdir, opts);
    if (rc) return rc;
    QSM_ARG(f && mask && x, "null pointer");
    if ((rc = ensure_device())) return rc;
    const size_t N = sz[0] * sz[1] * sz[2], es = dtype_size(dtype);
    DevBuf df, dm, dx, dw;
    if ((rc = df.alloc(N * es)) || (rc = dm.alloc(N)) || (rc = dx.alloc(N * es))
        || (rc = dw.alloc(qsm_rts_work_bytes(dtype, sz))))
        return rc;
    QSM_HIP(hipMemcpy(df.p, f, N * es, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(dm.p, mask, N, hipMemcpyHostToDevice));
    if ((rc = qsm_rts_dev(dtype, df.p, static_cast<const uint8_t *>(dm.p), sz, vsz, bdir, opts,
                          dx.p, dw.p, trace, iters, nullptr)))
        return rc;
    QSM_HIP(hipMemcpy(x, dx.p, N * es, hipMemcpyDeviceToHost));
    return QSM_OK;
}

}  // extern "C"
