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
// ((nx/2 + 1) ny nz complex).  Per iteration four launches touch HBM: the
// right-hand side (gradAdj of rho v - l, 6 reads + 1 write per voxel), the
// half-spectrum solve (five FFT passes, the k-space multipliers generated
// in-kernel), the v/l update with its residual partial sums (7 reads + 9
// writes), and the dual residual norms (6 reads).  Sums are accumulated in
// double per block and reduced in a fixed order, so they are reproducible.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "fft_dev.h"

using namespace qsm;
using namespace qsm::fft;

namespace {

constexpr int BX = 64, BY = 4, NT = BX * BY;

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
    block_sum<1>(s, part + flat_block());
}

// tv.m:196-200: r = gradAdj(rho v - l) (gradfp_adj_mex with -1/h), the
// differences rho*v - l formed per sample as the reference forms r1..r3.
template <typename T>
__global__ __launch_bounds__(NT) void rhs_kernel(T *r, CVec3<T> v, CVec3<T> lm, G3 g, T rho,
                                                 T hx, T hy, T hz)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    int mx, my, mz;
    minus_nbrs(g, i, j, k, l, mx, my, mz);
    const T X = hx * ((rho * v.x[l] - lm.x[l]) - (rho * v.x[mx] - lm.x[mx]));
    const T Y = hy * ((rho * v.y[l] - lm.y[l]) - (rho * v.y[my] - lm.y[my]));
    const T Z = hz * ((rho * v.z[l] - lm.z[l]) - (rho * v.z[mz] - lm.z[mz]));
    r[l] = adj_sum(X, Y, Z, j, k);
}

template <typename T>
__device__ __forceinline__ T sgn_of(T x)
{
    return x > T(0) ? T(1) : (x < T(0) ? T(-1) : x);  // sign(): 0 -> 0, NaN -> NaN
}

// tv.m:205-260: the v-subproblem, the multiplier update and the primal
// residual sums.  Writes v (new), l (new), dv = v - v_old; partial sums of
// |grad u - v| (nr), |grad u| (na), |v| (nb) per block.
template <typename T, bool ISO>
__global__ __launch_bounds__(NT) void shrink_kernel(Vec3<T> v, Vec3<T> lm, Vec3<T> dv,
                                                    const T *u, G3 g, T hx, T hy, T hz,
                                                    T alpha, T oma, T rho, T irho, double *part)
{
    int i, j, k, l;
    double s[3] = {0.0, 0.0, 0.0};
    if (vox(g, i, j, k, l)) {
        int px, py, pz;
        plus_nbrs(g, i, j, k, l, px, py, pz);
        const T c = u[l];
        const T dx = hx * (u[px] - c), dy = hy * (u[py] - c), dz = hz * (u[pz] - c);
        const T v1 = v.x[l], v2 = v.y[l], v3 = v.z[l];
        const T l1 = lm.x[l], l2 = lm.y[l], l3 = lm.z[l];
        // over-relaxation
        const T dx1 = alpha * dx + oma * v1;
        const T dy1 = alpha * dy + oma * v2;
        const T dz1 = alpha * dz + oma * v3;
        const T r1 = dx1 + irho * l1, r2 = dy1 + irho * l2, r3 = dz1 + irho * l3;
        T n1, n2, n3;
        if constexpr (!ISO) {
            // max(abs(r)-irho, 0) .* sign(r)
            n1 = fmax(fabs(r1) - irho, T(0)) * sgn_of(r1);
            n2 = fmax(fabs(r2) - irho, T(0)) * sgn_of(r2);
            n3 = fmax(fabs(r3) - irho, T(0)) * sgn_of(r3);
        } else {
            // s = (s - irho) .* (s > irho) ./ s, non-finite -> 0
            const T sn = sqrt((r1 * r1 + r2 * r2) + r3 * r3);
            T f = ((sn - irho) * T(sn > irho ? 1 : 0)) / sn;
            if (!isfinite(f)) f = T(0);
            n1 = r1 * f;
            n2 = r2 * f;
            n3 = r3 * f;
        }
        v.x[l] = n1;
        v.y[l] = n2;
        v.z[l] = n3;
        lm.x[l] = l1 + rho * (dx1 - n1);
        lm.y[l] = l2 + rho * (dy1 - n2);
        lm.z[l] = l3 + rho * (dz1 - n3);
        dv.x[l] = n1 - v1;
        dv.y[l] = n2 - v2;
        dv.z[l] = n3 - v3;
        const T e1 = dx - n1, e2 = dy - n2, e3 = dz - n3;
        s[0] = double(sqrt((e1 * e1 + e2 * e2) + e3 * e3));
        s[1] = double(sqrt((dx * dx + dy * dy) + dz * dz));
        s[2] = double(sqrt((n1 * n1 + n2 * n2) + n3 * n3));
    }
    block_sum<3>(s, part + 3 * flat_block());
}

// tv.m:262-277: squared norms of rho gradAdj(dv) (dual residual) and of
// rho gradAdj(l) (the dual tolerance's scale).
template <typename T>
__global__ __launch_bounds__(NT) void dual_kernel(CVec3<T> dv, CVec3<T> lm, G3 g, T rho, T hx,
                                                  T hy, T hz, double *part)
{
    int i, j, k, l;
    double s[2] = {0.0, 0.0};
    if (vox(g, i, j, k, l)) {
        int mx, my, mz;
        minus_nbrs(g, i, j, k, l, mx, my, mz);
        const T a = rho * adj_sum(hx * (dv.x[l] - dv.x[mx]), hy * (dv.y[l] - dv.y[my]),
                                  hz * (dv.z[l] - dv.z[mz]), j, k);
        const T b = rho * adj_sum(hx * (lm.x[l] - lm.x[mx]), hy * (lm.y[l] - lm.y[my]),
                                  hz * (lm.z[l] - lm.z[mz]), j, k);
        s[0] = double(a) * double(a);
        s[1] = double(b) * double(b);
    }
    block_sum<2>(s, part + 2 * flat_block());
}

// Totals of NQ interleaved partial sums, in a fixed order.
template <int NQ>
__global__ __launch_bounds__(NT) void finish_kernel(const double *part, int nblk, double *out)
{
    double s[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q)
        s[q] = 0.0;
    const int t = threadIdx.x + BX * threadIdx.y;
    for (int b = t; b < nblk; b += NT)
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            s[q] += part[NQ * b + q];
    block_sum<NQ>(s, out);
}

// tv.m:308: x = mask .* u.
template <typename T>
__global__ __launch_bounds__(NT) void mask_kernel(T *x, const T *u, const uint8_t *m, G3 g)
{
    int i, j, k, l;
    if (!vox(g, i, j, k, l)) return;
    x[l] = T(m[l] ? 1 : 0) * u[l];
}

// The forward transform of f is kept as it comes from the x pass (doubled).
template <typename T>
struct KeepOp {
    __device__ Cx<T> operator()(int, int, int, Cx<T> v) const { return v; }
};

// k-space operation of the u-subproblem at half-spectrum sample (x, y, z),
// tv.m:159-174 and :203-204: with iA(k) = 1/(mu D(k)^2 + rho L(k)) (0 where
// not finite), the reference forms real(ifft3(iA .* (mu conj(D) fft3(f) +
// fft3(r)))).  For real f and r that is the inverse transform of
// A1s fft3(f) + A2s fft3(r), A1 = mu iA D, A2 = iA, As(k) = (A(k) + A(-k))/2;
// As = A except on the Nyquist plane of an even axis with a tilted B0 (D
// differs at k and -k there; L does not).  L(k) = sum_a (2 - 2 cos(2 pi m_a /
// n_a)) / h_a^2 is the symbol of -laplaceKernel (laplaceKernel.m:14-24)
// padded and circularly shifted to the origin (tv.m:163-165).  In double,
// rounded once to T.
template <typename T>
struct SolveOp {
    const Cx<T> *F;  // 2 fft3(f) on the half spectrum
    int M0, M1, M2, H;
    double dk0, dk1, dk2, b0, b1, b2;
    double mu, rho, lx, ly, lz, scale;

    __device__ static double lap1(int m, int n, double ih2)
    {
        const int mm = min(m, n - m);  // cos(2 pi m/n) = cos(2 pi (n-m)/n), evaluated once
        return (2.0 - 2.0 * cospi(2.0 * double(mm) / double(n))) * ih2;
    }

    __device__ Cx<T> operator()(int x, int y, int z, Cx<T> fr) const
    {
        const double X = kcoord(x, M0, dk0), Y = kcoord(y, M1, dk1), Z = kcoord(z, M2, dk2);
        const double L = (lap1(x, M0, lx) + lap1(y, M1, ly)) + lap1(z, M2, lz);
        double D = dip(X, Y, Z, b0, b1, b2);
        double iA = 1.0 / (mu * D * D + rho * L);
        if (!isfinite(iA)) iA = 0.0;
        double a1 = mu * iA * D, a2 = iA;
        const bool nx = 2 * x == M0, ny = 2 * y == M1, nz = 2 * z == M2;
        if (nx || ny || nz) {
            const double Dm = dip(nx ? X : -X, ny ? Y : -Y, nz ? Z : -Z, b0, b1, b2);
            double iAm = 1.0 / (mu * Dm * Dm + rho * L);
            if (!isfinite(iAm)) iAm = 0.0;
            a1 = 0.5 * (a1 + mu * iAm * Dm);
            a2 = 0.5 * (a2 + iAm);
        }
        const T c1 = T(a1 * scale), c2 = T(a2 * scale);
        const Cx<T> ff = F[size_t(x) + size_t(H) * (size_t(y) + size_t(M1) * z)];
        return Cx<T>{c1 * ff.re + c2 * fr.re, c1 * ff.im + c2 * fr.im};
    }
};

// Work space, carved from one device buffer (256-byte aligned pieces).
template <typename T>
struct Work {
    Cx<T> *F, *W;
    T *r, *u;
    Vec3<T> v, l, dv;
    double *part, *tot;
};

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

template <typename T>
size_t work_layout(const size_t sz[3], Work<T> *w, char *base)
{
    const size_t N = sz[0] * sz[1] * sz[2];
    const size_t M[3] = {sz[0], sz[1], sz[2]};
    const size_t spec = align256(real_conv_work_bytes(M, sizeof(T)));
    const size_t vol = align256(N * sizeof(T));
    const G3 g{int(sz[0]), int(sz[1]), int(sz[2])};
    const size_t part = align256(size_t(vox_blocks(g)) * 3 * sizeof(double));
    size_t off = 0;
    auto take = [&](size_t n) { char *p = base ? base + off : nullptr; off += n; return p; };
    char *pF = take(spec), *pW = take(spec), *pr = take(vol), *pu = take(vol);
    char *pv[3], *pl[3], *pd[3];
    for (auto &p : pv) p = take(vol);
    for (auto &p : pl) p = take(vol);
    for (auto &p : pd) p = take(vol);
    char *pp = take(part), *pt = take(256);
    if (w) {
        w->F = reinterpret_cast<Cx<T> *>(pF);
        w->W = reinterpret_cast<Cx<T> *>(pW);
        w->r = reinterpret_cast<T *>(pr);
        w->u = reinterpret_cast<T *>(pu);
        w->v = {reinterpret_cast<T *>(pv[0]), reinterpret_cast<T *>(pv[1]), reinterpret_cast<T *>(pv[2])};
        w->l = {reinterpret_cast<T *>(pl[0]), reinterpret_cast<T *>(pl[1]), reinterpret_cast<T *>(pl[2])};
        w->dv = {reinterpret_cast<T *>(pd[0]), reinterpret_cast<T *>(pd[1]), reinterpret_cast<T *>(pd[2])};
        w->part = reinterpret_cast<double *>(pp);
        w->tot = reinterpret_cast<double *>(pt);
    }
    return off;
}

// Reads NQ totals back (the loop's decisions are taken on the host, as in
// tv.m).
template <int NQ>
int read_totals(const double *part, int nblk, double *dtot, double *h, hipStream_t st)
{
    finish_kernel<NQ><<<1, dim3(BX, BY), 0, st>>>(part, nblk, dtot);
    QSM_HIP(hipGetLastError());
    QSM_HIP(hipMemcpyAsync(h, dtot, NQ * sizeof(double), hipMemcpyDeviceToHost, st));
    QSM_HIP(hipStreamSynchronize(st));
    return QSM_OK;
}

template <typename T>
int tv_dev_t(const T *f, const uint8_t *mask, const size_t sz[3], const double vsz[3],
             const double bdir[3], const qsm_tv_opts &o, T *x, void *work, double *trace,
             int *iters, hipStream_t st)
{
    int rc;
    const G3 g{int(sz[0]), int(sz[1]), int(sz[2])};
    const size_t N = sz[0] * sz[1] * sz[2];
    Work<T> w;
    work_layout<T>(sz, &w, static_cast<char *>(work));
    const dim3 grid = vox_grid(g), blk(BX, BY);
    const int nblk = vox_blocks(g);

    // normalised b (dipoleKernel.m:35-38)
    const double bn = std::sqrt(bdir[0] * bdir[0] + bdir[1] * bdir[1] + bdir[2] * bdir[2]);
    QSM_ARG(bn > 0.0, "B0 direction must be non-zero");
    double b[3] = {bdir[0], bdir[1], bdir[2]};
    if (bn != 1.0)
        for (double &e : b) e /= bn;

    // gradient coefficients as gradfp_mex / gradfp_adj_mex form them
    const T gx = T(1.0 / vsz[0]), gy = T(1.0 / vsz[1]), gz = T(1.0 / vsz[2]);
    const T ax = T(-1.0 / vsz[0]), ay = T(-1.0 / vsz[1]), az = T(-1.0 / vsz[2]);

    const size_t M[3] = {sz[0], sz[1], sz[2]}, off[3] = {0, 0, 0};
    // F: 2 fft3(f) on the half spectrum (the mu conj(D) factor is applied in
    // the solve, tv.m:159)
    if ((rc = real_conv<T, false>(f, static_cast<T *>(nullptr), sz, M, off, w.F, KeepOp<T>{}, st)))
        return rc;

    SolveOp<T> op;
    op.F = w.F;
    op.M0 = g.nx;
    op.M1 = g.ny;
    op.M2 = g.nz;
    op.H = g.nx / 2 + 1;
    op.dk0 = 1.0 / (double(g.nx) * vsz[0]);
    op.dk1 = 1.0 / (double(g.ny) * vsz[1]);
    op.dk2 = 1.0 / (double(g.nz) * vsz[2]);
    op.b0 = b[0];
    op.b1 = b[1];
    op.b2 = b[2];
    op.mu = o.mu;
    op.lx = 1.0 / (vsz[0] * vsz[0]);
    op.ly = 1.0 / (vsz[1] * vsz[1]);
    op.lz = 1.0 / (vsz[2] * vsz[2]);
    // 1/(nx ny nz) of ifft3, 1/2 for the x pass's doubling
    op.scale = 0.5 / (double(g.nx) * double(g.ny) * double(g.nz));

    // tv.m:127-130 stopping tolerances
    const double atolp = std::sqrt(3.0 * double(N)) * o.atol;
    const double atold = std::sqrt(double(N)) * o.atol;
    double rho = o.rho, irho = 1.0 / rho;

    // u = f, l = 0, v = grad(u), nr = sum |v| (tv.m:141-149)
    init_kernel<T><<<grid, blk, 0, st>>>(w.v, w.l, f, g, gx, gy, gz, w.part);
    QSM_HIP(hipGetLastError());
    double nr;
    if ((rc = read_totals<1>(w.part, nblk, w.tot, &nr, st))) return rc;

    if (o.verbose)
        std::printf("\niter\t\t||r||\t\teps_pri\t\t||s||\t\teps_dual\n");
    int flag = 0, ii = 0;
    bool stop = false;
    const CVec3<T> cv{w.v.x, w.v.y, w.v.z}, cl{w.l.x, w.l.y, w.l.z}, cd{w.dv.x, w.dv.y, w.dv.z};
    for (ii = 1; ii <= o.maxit && !stop; ++ii) {
        // u-subproblem (tv.m:196-203)
        rhs_kernel<T><<<grid, blk, 0, st>>>(w.r, cv, cl, g, T(rho), ax, ay, az);
        QSM_HIP(hipGetLastError());
        op.rho = rho;
        if ((rc = real_conv<T, true>(w.r, w.u, sz, M, off, w.W, op, st))) return rc;
        // v-subproblem, multipliers, primal sums (tv.m:205-256)
        if (o.tv == 1)
            shrink_kernel<T, false><<<grid, blk, 0, st>>>(w.v, w.l, w.dv, w.u, g, gx, gy, gz,
                                                         T(o.alpha), T(1.0 - o.alpha), T(rho),
                                                         T(irho), w.part);
        else
            shrink_kernel<T, true><<<grid, blk, 0, st>>>(w.v, w.l, w.dv, w.u, g, gx, gy, gz,
                                                        T(o.alpha), T(1.0 - o.alpha), T(rho),
                                                        T(irho), w.part);
        QSM_HIP(hipGetLastError());
        double p3[3];
        if ((rc = read_totals<3>(w.part, nblk, w.tot, p3, st))) return rc;
        dual_kernel<T><<<grid, blk, 0, st>>>(cd, cl, g, T(rho), ax, ay, az, w.part);
        QSM_HIP(hipGetLastError());
        double p2[2];
        if ((rc = read_totals<2>(w.part, nblk, w.tot, p2, st))) return rc;

        // convergence check (tv.m:258-285)
        const double nr_ = nr;
        nr = p3[0];
        const double ns = std::sqrt(p2[0]);
        const double na = p3[1], nb = p3[2];
        const double ep = atolp + o.rtol * std::max(na, nb);
        const double ed = atold + o.rtol * std::sqrt(p2[1]);
        if (o.verbose)
            std::printf("%3d/%d\t\t%.4e\t%.4e\t%.4e\t%.4e\n", ii, o.maxit, nr, ep, ns, ed);
        if (nr < ep && ns < ed)
            stop = true;
        // parameter update (tv.m:287-306)
        if (!stop) {
            if (nr > o.stau * ns || (nr > o.rtau * nr_ && flag < 3)) {
                rho = o.gamma * rho;
                irho = 1.0 / rho;
                if (flag > 5)
                    flag = 0;
            } else if (nr < ns / o.stau) {
                rho = rho / (o.gamma * o.gamma);
                irho = 1.0 / rho;
                flag = flag + 1;
            }
        }
        if (trace) {
            double *t = trace + 5 * (ii - 1);
            t[0] = nr;
            t[1] = ep;
            t[2] = ns;
            t[3] = ed;
            t[4] = rho;
        }
        if (stop)
            break;
    }
    if (o.verbose)
        std::printf("\n");
    if (iters)
        *iters = std::min(ii, o.maxit);
    mask_kernel<T><<<grid, blk, 0, st>>>(x, w.u, mask, g);
    QSM_HIP(hipGetLastError());
    return QSM_OK;
}

int check_tv(int dtype, const size_t sz[3], const double vsz[3], const double bdir[3],
             const qsm_tv_opts *o)
{
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    QSM_ARG(sz && vsz && bdir && o, "null pointer");
    int rc = check_grid(sz, vsz, 1);
    if (rc) return rc;
    QSM_ARG(o->tv == 1 || o->tv == 2, "opts.tv must be 1 (anisotropic) or 2 (isotropic)");
    QSM_ARG(o->maxit >= 1, "opts.maxit must be >= 1");
    QSM_ARG(o->rho > 0.0 && o->gamma > 0.0 && o->stau > 0.0, "opts.rho, gamma, stau must be > 0");
    QSM_ARG(std::isfinite(o->mu) && std::isfinite(o->alpha), "opts.mu and alpha must be finite");
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

}  // extern "C"
