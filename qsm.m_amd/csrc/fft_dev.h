// fft_dev.h -- the FFT engine shared by the dipole convolution (fft.hip) and
// the ADMM inversion (admm.hip), gfx950.
//
// One kernel = one axis pass over a batch of pencils: a block loads P
// pencils of length M into LDS (coalesced: rows along x, or P consecutive x
// positions for the y/z axes), runs a mixed-radix Stockham transform in LDS
// (radices 8, 13, 11, 7, 5, 4, 3, 2; twiddles from a table computed in double
// from exact rationals, cast for single precision), and stores.
//
// Real convolutions (a real volume in, a real volume out, a k-space
// operation between) run on the half spectrum along x in five passes: x
// forward with two real rows as the real and imaginary parts of one complex
// transform (H = M0/2 + 1 outputs per row kept; zero padding fused into the
// load, only rows that hold data); y forward (only planes that hold data,
// zero rows not read); z forward, the k-space operation and z inverse in one
// pass (only the data planes read and stored); y inverse (only data rows
// stored); x inverse (two rows per complex transform rebuilt from their half
// spectra by conjugate symmetry; crop and real output fused into the store).
// The padding is never materialised.
#pragma once

#include <cmath>

#include "common.h"

namespace qsm {
namespace fft {

constexpr int MAX_STAGES = 24;
constexpr int FFT_THREADS = 512;

template <typename T>
struct Cx {
    T re, im;
};

// Complex arithmetic of the transforms.  No FFT result is bitwise-pinned
// (the transforms are checked against numpy to a tolerance), so these
// helpers and the butterflies let multiply-adds fuse.
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
#pragma clang fp contract(fast)
    return Cx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b)
{
#pragma clang fp contract(fast)
    return Cx<T>{a.re + b.re, a.im + b.im};
}

template <typename T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b)
{
#pragma clang fp contract(fast)
    return Cx<T>{a.re - b.re, a.im - b.im};
}

// Stage plan of one transform length (host-built, read by the kernels).
struct Plan {
    int M;
    int nst;
    int R[MAX_STAGES];    // radix of stage s
    int Ns[MAX_STAGES];   // sub-transform length before stage s
    int tw[MAX_STAGES];   // offset of stage s's twiddles in the table
    unsigned mNs[MAX_STAGES];  // exact reciprocals (udiv) of Ns[s] and of M
    unsigned mM;
};

// n / d for n < 2^16 and 1 <= d <= 4096 as a multiply-high by
// m = ceil(2^32 / d) (exact in that range; m = 0 stands for d = 1).
__host__ __device__ inline unsigned recip(unsigned d)
{
    return d == 1 ? 0u : unsigned(((1ull << 32) + d - 1) / d);
}

__device__ __forceinline__ unsigned udiv(unsigned n, unsigned m)
{
    return m ? __umulhi(n, m) : n;
}

// Pencils of a block in LDS: element i of pencil p at [i*PP + p], P a
// power of two (shifts instead of divisions) and PP = P + 1 (the odd row
// stride keeps a wavefront's accesses along i on distinct LDS banks).
struct Pen {
    int logP, PP;
};

// Small DFT of v[0..R-1] in place: y_k = sum_r v_r exp(sgn*2*pi*i*r*k/R),
// with the textbook symmetric-pair butterflies (constants to 22 digits;
// cast to T).
template <typename T>
__device__ __forceinline__ Cx<T> mul_i(Cx<T> a, T sgn)  // a * (sgn*i)
{
    return Cx<T>{-sgn * a.im, sgn * a.re};
}

template <typename T>
__device__ __forceinline__ Cx<T> scale(Cx<T> a, T c)
{
#pragma clang fp contract(fast)
    return Cx<T>{a.re * c, a.im * c};
}

template <typename T>
__device__ __forceinline__ void dft4(Cx<T> &a0, Cx<T> &a1, Cx<T> &a2, Cx<T> &a3, T sgn)
{
#pragma clang fp contract(fast)
    const Cx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const Cx<T> t2 = cadd(a1, a3), t3 = mul_i(csub(a1, a3), sgn);
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

// cos and sin of 2*pi*m/R, m = 0..R-1, for the odd prime radices that have
// no hand-written butterfly (computed to 22 digits in decimal arithmetic).
template <int R>
struct Roots;
template <>
struct Roots<11> {
    static constexpr double c[11] = {
        1.0, 0.8412535328311811688618, 0.4154150130018864255293, -0.1423148382732851404438,
        -0.6548607339452850640569, -0.9594929736144973898904, -0.9594929736144973898904,
        -0.6548607339452850640569, -0.1423148382732851404438, 0.4154150130018864255293,
        0.8412535328311811688618};
    static constexpr double s[11] = {
        0.0, 0.5406408174555975821076, 0.9096319953545183714117, 0.9898214418809327323761,
        0.7557495743542582837740, 0.2817325568414296977114, -0.2817325568414296977114,
        -0.7557495743542582837740, -0.9898214418809327323761, -0.9096319953545183714117,
        -0.5406408174555975821076};
};
template <>
struct Roots<13> {
    static constexpr double c[13] = {
        1.0, 0.8854560256532098959004, 0.5680647467311558025118, 0.1205366802553230533491,
        -0.3546048870425356259696, -0.7485107481711010986346, -0.9709418174260520271570,
        -0.9709418174260520271570, -0.7485107481711010986346, -0.3546048870425356259696,
        0.1205366802553230533491, 0.5680647467311558025118, 0.8854560256532098959004};
    static constexpr double s[13] = {
        0.0, 0.4647231720437685456560, 0.8229838658936563945796, 0.9927088740980539928008,
        0.9350162426854148234398, 0.6631226582407952023768, 0.2393156642875577671488,
        -0.2393156642875577671488, -0.6631226582407952023768, -0.9350162426854148234398,
        -0.9927088740980539928008, -0.8229838658936563945796, -0.4647231720437685456560};
};

// Odd prime R: with t_r = v_r + v_{R-r} and d_r = v_r - v_{R-r} (r = 1..h,
// h = (R-1)/2), y_k = v_0 + sum_r cos(2 pi rk/R) t_r + sgn i sum_r
// sin(2 pi rk/R) d_r and y_{R-k} the same with the second sum negated.
template <typename T, int R>
__device__ __forceinline__ void odd_prime_dft(Cx<T> *v, T sgn)
{
#pragma clang fp contract(fast)
    constexpr int h = (R - 1) / 2;
    Cx<T> t[h], d[h];
    Cx<T> y0 = v[0];
#pragma unroll
    for (int r = 1; r <= h; ++r) {
        t[r - 1] = cadd(v[r], v[R - r]);
        d[r - 1] = csub(v[r], v[R - r]);
        y0 = cadd(y0, t[r - 1]);
    }
    Cx<T> y[R];
#pragma unroll
    for (int k = 1; k <= h; ++k) {
        Cx<T> a = v[0], b{T(0), T(0)};
#pragma unroll
        for (int r = 1; r <= h; ++r) {
            const int m = (r * k) % R;
            a = cadd(a, scale(t[r - 1], T(Roots<R>::c[m])));
            b = cadd(b, scale(d[r - 1], T(Roots<R>::s[m])));
        }
        b = mul_i(b, sgn);
        y[k] = cadd(a, b);
        y[R - k] = csub(a, b);
    }
    v[0] = y0;
#pragma unroll
    for (int k = 1; k < R; ++k)
        v[k] = y[k];
}

template <typename T, int R>
__device__ __forceinline__ void small_dft(Cx<T> *v, T sgn)
{
#pragma clang fp contract(fast)
    if constexpr (R == 2) {
        const Cx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const T S3 = T(0.8660254037844386467637);
        const Cx<T> t = cadd(v[1], v[2]);
        const Cx<T> d = mul_i(scale(csub(v[1], v[2]), S3), sgn);
        const Cx<T> m = csub(v[0], scale(t, T(0.5)));
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, d);
        v[2] = csub(m, d);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3], sgn);
    } else if constexpr (R == 5) {
        const T C1 = T(0.3090169943749474241022), C2 = T(-0.8090169943749474241023);
        const T S1 = T(0.9510565162951535721164), S2 = T(0.5877852522924731291687);
        const Cx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        const Cx<T> d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
        const Cx<T> a1 = cadd(v[0], cadd(scale(t1, C1), scale(t2, C2)));
        const Cx<T> a2 = cadd(v[0], cadd(scale(t1, C2), scale(t2, C1)));
        const Cx<T> b1 = mul_i(cadd(scale(d1, S1), scale(d2, S2)), sgn);
        const Cx<T> b2 = mul_i(csub(scale(d1, S2), scale(d2, S1)), sgn);
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(a1, b1);
        v[4] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[3] = csub(a2, b2);
    } else if constexpr (R == 7) {
        const T C1 = T(0.6234898018587335305250), C2 = T(-0.2225209339563144042889);
        const T C3 = T(-0.9009688679024191262361);
        const T S1 = T(0.7818314824680298087084), S2 = T(0.9749279121818236070181);
        const T S3 = T(0.4338837391175581204758);
        const Cx<T> t1 = cadd(v[1], v[6]), t2 = cadd(v[2], v[5]), t3 = cadd(v[3], v[4]);
        const Cx<T> d1 = csub(v[1], v[6]), d2 = csub(v[2], v[5]), d3 = csub(v[3], v[4]);
        // k = 1: cos(2 pi j/7) pairs (1,2,3); sin signs +,+,+
        const Cx<T> a1 = cadd(v[0], cadd(cadd(scale(t1, C1), scale(t2, C2)), scale(t3, C3)));
        const Cx<T> b1 = mul_i(cadd(cadd(scale(d1, S1), scale(d2, S2)), scale(d3, S3)), sgn);
        // k = 2: jk mod 7 = 2, 4, 6 -> cos C2, C3, C1; sin S2, -S3, -S1
        const Cx<T> a2 = cadd(v[0], cadd(cadd(scale(t1, C2), scale(t2, C3)), scale(t3, C1)));
        const Cx<T> b2 = mul_i(csub(csub(scale(d1, S2), scale(d2, S3)), scale(d3, S1)), sgn);
        // k = 3: jk mod 7 = 3, 6, 2 -> cos C3, C1, C2; sin S3, -S1, S2
        const Cx<T> a3 = cadd(v[0], cadd(cadd(scale(t1, C3), scale(t2, C1)), scale(t3, C2)));
        const Cx<T> b3 = mul_i(cadd(csub(scale(d1, S3), scale(d2, S1)), scale(d3, S2)), sgn);
        v[0] = cadd(v[0], cadd(cadd(t1, t2), t3));
        v[1] = cadd(a1, b1);
        v[6] = csub(a1, b1);
        v[2] = cadd(a2, b2);
        v[5] = csub(a2, b2);
        v[3] = cadd(a3, b3);
        v[4] = csub(a3, b3);
    } else if constexpr (R == 11 || R == 13) {
        odd_prime_dft<T, R>(v, sgn);
    } else {
        static_assert(R == 8, "radix");
        // 8 = 2 x 4: pair (r, r+4), twiddle the differences by W8^r, then
        // two DFT-4s give the even and the odd outputs.
        const T R2 = T(0.7071067811865475244008);
        Cx<T> t[4], u[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            t[r] = cadd(v[r], v[r + 4]);
            u[r] = csub(v[r], v[r + 4]);
        }
        // W8^1 = (1 + sgn i)/sqrt2, W8^2 = sgn i, W8^3 = (-1 + sgn i)/sqrt2
        u[1] = scale(cadd(u[1], mul_i(u[1], sgn)), R2);
        u[2] = mul_i(u[2], sgn);
        u[3] = scale(csub(mul_i(u[3], sgn), u[3]), R2);
        dft4(t[0], t[1], t[2], t[3], sgn);
        dft4(u[0], u[1], u[2], u[3], sgn);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            v[2 * m] = t[m];
            v[2 * m + 1] = u[m];
        }
    }
}

// One radix-R Stockham stage: A -> B.
template <typename T, int R>
__device__ __forceinline__ void stage(const Cx<T> *A, Cx<T> *B, Pen pn, int M, int Ns,
                                      unsigned mNs, const Cx<T> *__restrict__ tws, T sgn)
{
#pragma clang fp contract(fast)
    const int nb = M / R;
    const int PP = pn.PP, pm = (1 << pn.logP) - 1;
    for (int b = threadIdx.x; b < (nb << pn.logP); b += FFT_THREADS) {
        const int j = b >> pn.logP, p = b & pm;
        const int q = int(udiv(unsigned(j), mNs));
        const int k = j - q * Ns;
        Cx<T> v[R];
        v[0] = A[j * PP + p];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const Cx<T> w = tws[k * (R - 1) + (r - 1)];
            const Cx<T> wt{w.re, sgn < 0 ? w.im : -w.im};
            v[r] = cmul(A[(j + r * nb) * PP + p], wt);
        }
        small_dft<T, R>(v, sgn);
        const int base = q * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r)
            B[(base + r * Ns) * PP + p] = v[r];
    }
}

// Stockham transform of the block's pencils held in LDS; returns the
// buffer holding the result.  sgn = -1 forward, +1 inverse (twiddles
// conjugated).  Inlined so that the compiler sees the buffers are LDS (ds_
// instructions, not flat ones).
template <typename T>
__device__ __forceinline__ Cx<T> *stockham(Cx<T> *A, Cx<T> *B, Pen pn,
                                           const Plan *__restrict__ pl,
                                           const Cx<T> *__restrict__ tw, T sgn)
{
    const int M = pl->M, nst = pl->nst;
    for (int s = 0; s < nst; ++s) {
        const int R = pl->R[s], Ns = pl->Ns[s];
        const unsigned mNs = pl->mNs[s];
        const Cx<T> *tws = tw + pl->tw[s];
        switch (R) {
        case 2: stage<T, 2>(A, B, pn, M, Ns, mNs, tws, sgn); break;
        case 3: stage<T, 3>(A, B, pn, M, Ns, mNs, tws, sgn); break;
        case 4: stage<T, 4>(A, B, pn, M, Ns, mNs, tws, sgn); break;
        case 5: stage<T, 5>(A, B, pn, M, Ns, mNs, tws, sgn); break;
        case 7: stage<T, 7>(A, B, pn, M, Ns, mNs, tws, sgn); break;
        case 11: stage<T, 11>(A, B, pn, M, Ns, mNs, tws, sgn); break;
        case 13: stage<T, 13>(A, B, pn, M, Ns, mNs, tws, sgn); break;
        default: stage<T, 8>(A, B, pn, M, Ns, mNs, tws, sgn); break;
        }
        __syncthreads();
        Cx<T> *t = A; A = B; B = t;
    }
    return A;
}

// Where the real volume (nx, ny, nz) sits in the transform grid: offsets
// (o0, o1, o2) of zero padding before it; H = M0/2 + 1 half-spectrum
// columns; M1 the y length.
struct RealGeom {
    int nx, ny, nz, o0, o1, o2, H, M1;
};

}  // namespace fft
}  // namespace qsm

#include "fft_reg.h"

namespace qsm {
namespace fft {

// Pass geometry.  The volume is complex, dims M0 x M1 x M2, x fastest.
// axis 0: pencils are rows (i1, i2); the block takes P consecutive rows of
// one plane.  axis 1/2: pencils run along y/z; the block takes P
// consecutive x positions (coalesced) of one (i2) / (i1) coordinate.
struct Vol {
    int M0, M1, M2;
};

// Block copies between global memory and the LDS pencils, COPY_UNROLL
// accesses in flight per thread: all loads of a group are issued before
// the first LDS write (and all LDS reads before the first store), so a
// thread waits for memory once per group, not once per element.
constexpr int COPY_UNROLL = 8;

// For e in [0, total): A[dst] = v where v = src(e, dst) (src sets dst, or
// leaves it negative for nothing to write).
template <typename T, typename Src>
__device__ __forceinline__ void fill(Cx<T> *A, int total, Src src)
{
    for (int e0 = threadIdx.x; e0 < total; e0 += COPY_UNROLL * FFT_THREADS) {
        Cx<T> r[COPY_UNROLL];
        int d[COPY_UNROLL];
#pragma unroll
        for (int u = 0; u < COPY_UNROLL; ++u) {
            const int e = e0 + u * FFT_THREADS;
            d[u] = -1;
            r[u] = Cx<T>{T(0), T(0)};
            if (e < total) r[u] = src(e, d[u]);
        }
#pragma unroll
        for (int u = 0; u < COPY_UNROLL; ++u)
            if (d[u] >= 0) A[d[u]] = r[u];
    }
}

// For e in [0, total): put(e, R[at(e)]) (put stores or drops it).
template <typename T, typename At, typename Put>
__device__ __forceinline__ void drain(const Cx<T> *R, int total, At at, Put put)
{
    for (int e0 = threadIdx.x; e0 < total; e0 += COPY_UNROLL * FFT_THREADS) {
        Cx<T> r[COPY_UNROLL];
#pragma unroll
        for (int u = 0; u < COPY_UNROLL; ++u) {
            const int e = e0 + u * FFT_THREADS;
            if (e < total) r[u] = R[at(e)];
        }
#pragma unroll
        for (int u = 0; u < COPY_UNROLL; ++u) {
            const int e = e0 + u * FFT_THREADS;
            if (e < total) put(e, r[u]);
        }
    }
}

// k-space coordinate of index m on an axis of length n (dipoleKernel.m:
// 74-87 followed by ifftshift, :64); dk = 1 / (n h), formed on the host.
__device__ __forceinline__ double kcoord(int m, int n, double dk)
{
    const int low = n / 2;
    return double(m < n - low ? m : m - n) * dk;
}

// D(k) = 1/3 - (k.b)^2 / |k|^2, D(0) = 0 (dipoleKernel.m:59-66), in double.
__device__ __forceinline__ double dip(double X, double Y, double Z, double b0, double b1,
                                      double b2)
{
    const double kz = (b0 * X + b1 * Y) + b2 * Z;
    const double k2 = (X * X + Y * Y) + Z * Z;
    return (k2 == 0.0) ? 0.0 : (1.0 / 3.0 - (kz * kz) / k2);
}

// ---- generic in-place pass (plain fft3 / ifft3) -------------------------
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void pass_kernel(
    Cx<T> *data, Vol v, int axis, Pen pn, const Plan *pl, const Cx<T> *tw, T sgn, T scale)
{
    extern __shared__ unsigned char smem[];
    Cx<T> *A = reinterpret_cast<Cx<T> *>(smem);
    const int P = 1 << pn.logP, PP = pn.PP;
    Cx<T> *B = A + size_t(PP) * pl->M;
    const int M = pl->M;
    size_t base;
    size_t stride;  // element stride along the axis
    size_t pstride; // stride between the block's pencils
    int np;         // pencils in this block
    if (axis == 0) {
        const int rows = v.M1 * v.M2;
        const int r0 = blockIdx.x * P;
        np = min(P, rows - r0);
        base = size_t(r0) * v.M0;
        stride = 1;
        pstride = v.M0;
    } else {
        const int nxb = (v.M0 + P - 1) / P;
        const int lb = xcd_block();  // neighbouring x groups share L2 lines
        const int x0 = (lb % nxb) * P;
        const int other = lb / nxb;
        np = min(P, v.M0 - x0);
        stride = axis == 1 ? size_t(v.M0) : size_t(v.M0) * v.M1;
        base = axis == 1 ? size_t(x0) + size_t(other) * v.M0 * v.M1 : size_t(x0) + size_t(other) * v.M0;
        pstride = 1;
    }
    const unsigned mM = pl->mM;
    const int logP = pn.logP;
    auto pi = [=](int e, int &p, int &i) {
        if (axis == 0) { p = int(udiv(unsigned(e), mM)); i = e - p * M; }
        else { i = e >> logP; p = e & (P - 1); }
    };
    fill(A, P * M, [&](int e, int &d) {
        int p, i;
        pi(e, p, i);
        d = i * PP + p;
        return p < np ? data[base + size_t(p) * pstride + size_t(i) * stride] : Cx<T>{T(0), T(0)};
    });
    __syncthreads();
    Cx<T> *R = stockham(A, B, pn, pl, tw, sgn);
    drain(R, P * M, [&](int e) { int p, i; pi(e, p, i); return i * PP + p; },
          [&](int e, Cx<T> x) {
              int p, i;
              pi(e, p, i);
              if (p < np) {
                  x.re *= scale;
                  x.im *= scale;
                  data[base + size_t(p) * pstride + size_t(i) * stride] = x;
              }
          });
}

// ---- real-convolution passes ---------------------------------------------

// x pass, forward: reads the real input for rows (j, k) inside the data
// region only.  Rows 2p and 2p+1 of the block travel as the real and
// imaginary parts of pencil p; with Z = FFT(a + i b) and W(k) =
// conj(Z(M - k)) the row spectra are A = (Z + W)/2 and B = (Z - W)/(2i).
// Stored: 2A and 2B for k = 0..H-1 (the caller folds the 1/2 into its
// scaling) into the half-spectrum volume (H, M1, M2).
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void rx_fwd(
    Cx<T> *data, const T *in, RealGeom g, Pen pn, const Plan *pl, const Cx<T> *tw,
    const int *skip)
{
    if (skip && *skip)
        return;
    extern __shared__ unsigned char smem[];
    Cx<T> *A = reinterpret_cast<Cx<T> *>(smem);
    const int P = 1 << pn.logP, PP = pn.PP;
    Cx<T> *B = A + size_t(PP) * pl->M;
    const int M = pl->M;
    const int rows = g.ny * g.nz;
    const int r0 = blockIdx.x * 2 * P;
    const unsigned mM = pl->mM;
    fill(A, P * M, [&](int e, int &d) {
        const int p = int(udiv(unsigned(e), mM)), i = e - p * M;
        const int ra = r0 + 2 * p;
        T a = T(0), b = T(0);
        if (i >= g.o0 && i < g.o0 + g.nx) {
            if (ra < rows) a = in[size_t(i - g.o0) + size_t(g.nx) * ra];
            if (ra + 1 < rows) b = in[size_t(i - g.o0) + size_t(g.nx) * (ra + 1)];
        }
        d = i * PP + p;
        return Cx<T>{a, b};
    });
    __syncthreads();
    Cx<T> *R = stockham(A, B, pn, pl, tw, T(-1));
    const int H = g.H;
    const unsigned mH = recip(unsigned(H));
    for (int e = threadIdx.x; e < P * H; e += FFT_THREADS) {
        const int p = int(udiv(unsigned(e), mH)), k = e - p * H;
        const int ra = r0 + 2 * p;
        if (ra >= rows) continue;
        const Cx<T> z = R[k * PP + p];
        const Cx<T> w = R[(k == 0 ? 0 : M - k) * PP + p];  // conj taken below
        // 2A = z + conj(w), 2B = -i (z - conj(w))
        const Cx<T> sa{z.re + w.re, z.im - w.im};
        const Cx<T> sb{z.im + w.im, w.re - z.re};
        const int ja = ra % g.ny + g.o1, ka = ra / g.ny + g.o2;
        data[size_t(k) + size_t(H) * (size_t(ja) + size_t(g.M1) * ka)] = sa;
        if (ra + 1 < rows) {
            const int jb = (ra + 1) % g.ny + g.o1, kb = (ra + 1) / g.ny + g.o2;
            data[size_t(k) + size_t(H) * (size_t(jb) + size_t(g.M1) * kb)] = sb;
        }
    }
}

// y pass over planes k in [k0, k0 + nk) (forward or inverse).  Rows
// outside [ylo, yhi) are zero on input (not read); only rows in
// [slo, shi) are stored (the later passes read no others).
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void cy_pass(
    Cx<T> *data, Vol v, int k0, Pen pn, const Plan *pl, const Cx<T> *tw, T sgn,
    int ylo, int yhi, int slo, int shi, const int *skip)
{
    if (skip && *skip)
        return;
    extern __shared__ unsigned char smem[];
    Cx<T> *A = reinterpret_cast<Cx<T> *>(smem);
    const int P = 1 << pn.logP, PP = pn.PP;
    Cx<T> *B = A + size_t(PP) * pl->M;
    const int M = pl->M;
    const int nxb = (v.M0 + P - 1) / P;
    const int lb = xcd_block();  // neighbouring x groups share L2 lines
    const int x0 = (lb % nxb) * P;
    const int k = k0 + lb / nxb;
    const int np = min(P, v.M0 - x0);
    const size_t base = size_t(x0) + size_t(k) * v.M0 * v.M1;
    const int logP = pn.logP;
    fill(A, P * M, [&](int e, int &d) {
        const int i = e >> logP, p = e & (P - 1);
        d = i * PP + p;
        return (p < np && i >= ylo && i < yhi) ? data[base + p + size_t(i) * v.M0]
                                               : Cx<T>{T(0), T(0)};
    });
    __syncthreads();
    Cx<T> *R = stockham(A, B, pn, pl, tw, sgn);
    drain(R, P * M, [&](int e) { return (e >> logP) * PP + (e & (P - 1)); },
          [&](int e, Cx<T> x) {
              const int i = e >> logP, p = e & (P - 1);
              if (p < np && i >= slo && i < shi)
                  data[base + p + size_t(i) * v.M0] = x;
          });
}

// z pass over the half spectrum (x extent v.M0 = H): forward transform,
// op(x, y, z, value) on every k-space sample, and, when ROUND, the inverse
// transform.  Only planes [zlo, zhi) hold data on input and only they are
// needed on output (others neither read nor stored); without ROUND every
// plane of the spectrum is stored.
template <typename T, bool ROUND, typename Op>
__global__ __launch_bounds__(FFT_THREADS) void z_pass(
    Cx<T> *data, Vol v, Pen pn, const Plan *pl, const Cx<T> *tw, int zlo, int zhi, Op op,
    const int *skip)
{
    if (skip && *skip)
        return;
    extern __shared__ unsigned char smem[];
    Cx<T> *A = reinterpret_cast<Cx<T> *>(smem);
    const int P = 1 << pn.logP, PP = pn.PP;
    Cx<T> *B = A + size_t(PP) * pl->M;
    const int M = pl->M;
    const int nxb = (v.M0 + P - 1) / P;
    const int lb = xcd_block();  // neighbouring x groups share L2 lines
    const int x0 = (lb % nxb) * P;
    const int j = lb / nxb;
    const int np = min(P, v.M0 - x0);
    const size_t plane = size_t(v.M0) * v.M1;
    const size_t base = size_t(x0) + size_t(j) * v.M0;
    const int logP = pn.logP;
    fill(A, P * M, [&](int e, int &d) {
        const int i = e >> logP, p = e & (P - 1);
        d = i * PP + p;
        return (p < np && i >= zlo && i < zhi) ? data[base + p + size_t(i) * plane]
                                               : Cx<T>{T(0), T(0)};
    });
    __syncthreads();
    Cx<T> *R = stockham(A, B, pn, pl, tw, T(-1));
    Cx<T> *S = (R == A) ? B : A;
    for (int e = threadIdx.x; e < P * M; e += FFT_THREADS) {
        const int i = e >> logP, p = e & (P - 1);
        if (p < np)
            R[i * PP + p] = op(x0 + p, j, i, R[i * PP + p]);
    }
    __syncthreads();
    Cx<T> *Q = ROUND ? stockham(R, S, pn, pl, tw, T(1)) : R;
    const int slo = ROUND ? zlo : 0, shi = ROUND ? zhi : M;
    drain(Q, P * M, [&](int e) { return (e >> logP) * PP + (e & (P - 1)); },
          [&](int e, Cx<T> x) {
              const int i = e >> logP, p = e & (P - 1);
              if (p < np && i >= slo && i < shi)
                  data[base + p + size_t(i) * plane] = x;
          });
}

// z pass from a stored spectrum (all M2 planes, as real_conv<T, false>
// leaves it): op on every sample of src, the inverse z transform, and the
// planes [zlo, zhi) stored into data for the inverse y and x passes.  The
// op sees the values the fused z pass would hold after its forward
// transform, so results are bitwise those of real_conv<T, true>.
template <typename T, typename Op>
__global__ __launch_bounds__(FFT_THREADS) void z_inv_pass(
    const Cx<T> *src, Cx<T> *data, Vol v, Pen pn, const Plan *pl, const Cx<T> *tw, int zlo,
    int zhi, Op op)
{
    extern __shared__ unsigned char smem[];
    Cx<T> *A = reinterpret_cast<Cx<T> *>(smem);
    const int P = 1 << pn.logP, PP = pn.PP;
    Cx<T> *B = A + size_t(PP) * pl->M;
    const int M = pl->M;
    const int nxb = (v.M0 + P - 1) / P;
    const int lb = xcd_block();  // neighbouring x groups share L2 lines
    const int x0 = (lb % nxb) * P;
    const int j = lb / nxb;
    const int np = min(P, v.M0 - x0);
    const size_t plane = size_t(v.M0) * v.M1;
    const size_t base = size_t(x0) + size_t(j) * v.M0;
    const int logP = pn.logP;
    fill(A, P * M, [&](int e, int &d) {
        const int i = e >> logP, p = e & (P - 1);
        d = i * PP + p;
        return p < np ? op(x0 + p, j, i, src[base + p + size_t(i) * plane])
                      : Cx<T>{T(0), T(0)};
    });
    __syncthreads();
    Cx<T> *Q = stockham(A, B, pn, pl, tw, T(1));
    drain(Q, P * M, [&](int e) { return (e >> logP) * PP + (e & (P - 1)); },
          [&](int e, Cx<T> x) {
              const int i = e >> logP, p = e & (P - 1);
              if (p < np && i >= zlo && i < zhi)
                  data[base + p + size_t(i) * plane] = x;
          });
}

// x pass, inverse: rows inside the output region, two per complex
// transform.  From the half spectra A, B (k = 0..H-1) of rows 2p and 2p+1,
// pencil p is Z(k) = A(k) + i B(k), with A(k) = conj(A(M - k)) (likewise B)
// above H; the inverse transform of Z is a + i b.  Stores the cropped rows
// into out (nx, ny, nz).
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void rx_inv(
    const Cx<T> *data, T *out, RealGeom g, Pen pn, const Plan *pl, const Cx<T> *tw,
    const int *skip)
{
    if (skip && *skip)
        return;
    extern __shared__ unsigned char smem[];
    Cx<T> *A = reinterpret_cast<Cx<T> *>(smem);
    const int P = 1 << pn.logP, PP = pn.PP;
    Cx<T> *B = A + size_t(PP) * pl->M;
    const int M = pl->M;
    const int rows = g.ny * g.nz;
    const int r0 = blockIdx.x * 2 * P;
    const unsigned mM = pl->mM;
    const int H = g.H;
    fill(A, P * M, [&](int e, int &d) {
        const int p = int(udiv(unsigned(e), mM)), i = e - p * M;
        const int ra = r0 + 2 * p;
        const bool lo = i < H;
        const int k = lo ? i : M - i;
        Cx<T> a{T(0), T(0)}, b{T(0), T(0)};
        if (ra < rows) {
            const int j = ra % g.ny + g.o1, kk = ra / g.ny + g.o2;
            a = data[size_t(k) + size_t(H) * (size_t(j) + size_t(g.M1) * kk)];
        }
        if (ra + 1 < rows) {
            const int j = (ra + 1) % g.ny + g.o1, kk = (ra + 1) / g.ny + g.o2;
            b = data[size_t(k) + size_t(H) * (size_t(j) + size_t(g.M1) * kk)];
        }
        d = i * PP + p;
        // lo: a + i b; above H: conj(a) + i conj(b)
        return lo ? Cx<T>{a.re - b.im, a.im + b.re} : Cx<T>{a.re + b.im, b.re - a.im};
    });
    __syncthreads();
    Cx<T> *R = stockham(A, B, pn, pl, tw, T(1));
    drain(R, P * M,
          [&](int e) { const int p = int(udiv(unsigned(e), mM)); return (e - p * M) * PP + p; },
          [&](int e, Cx<T> x) {
              const int p = int(udiv(unsigned(e), mM)), i = e - p * M;
              const int ra = r0 + 2 * p;
              if (i >= g.o0 && i < g.o0 + g.nx) {
                  if (ra < rows) out[size_t(i - g.o0) + size_t(g.nx) * ra] = x.re;
                  if (ra + 1 < rows) out[size_t(i - g.o0) + size_t(g.nx) * (ra + 1)] = x.im;
              }
          });
}

// ---- host side ---------------------------------------------------------------

// LDS per block: two buffers of P pencils.  80 KB lets two blocks share a
// CU's 160 KB.
constexpr size_t LDS_BUDGET = 80 * 1024;
// one pencil of the longest supported length (2048 double, 4096 single)
constexpr size_t LDS_MAX = 64 * 1024;
static_assert(LDS_MAX <= LDS_BUDGET, "the pass kernels are allowed LDS_BUDGET bytes");

inline size_t lds_bytes(Pen pn, int M, size_t csize) { return 2 * size_t(pn.PP) * M * csize; }

// Pencils per block: the largest power of two P <= 16 (16 consecutive x
// positions are whole cache lines) whose buffers fit the budget.
inline Pen pencils(int M, size_t csize)
{
    for (int logP = 4; logP > 0; --logP) {
        const Pen pn{logP, (1 << logP) + 1};
        if (lds_bytes(pn, M, csize) <= LDS_BUDGET)
            return pn;
    }
    return Pen{0, 1};
}

// Lets kernel K use more than the default 64 KB of dynamic LDS (once per
// kernel and process).
template <auto K>
int allow_lds()
{
    static const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(K), hipFuncAttributeMaxDynamicSharedMemorySize,
        int(LDS_BUDGET));
    QSM_HIP(e);
    return QSM_OK;
}

struct AxisPlan {
    Plan pl;                    // host copy
    const Plan *dpl = nullptr;  // device copy (head of the buffer)
    const void *dtw = nullptr;  // device twiddle table in the transform's
                                // precision (after the plan); the single
                                // table is the double one rounded
    const void *dtwM = nullptr; // exp(-2 pi i j / M), j < M, for the
                                // register passes (after dtw)
};

template <typename T>
const Cx<T> *tw_of(const AxisPlan &ap)
{
    return static_cast<const Cx<T> *>(ap.dtw);
}

template <typename T>
const Cx<T> *twM_of(const AxisPlan &ap)
{
    return static_cast<const Cx<T> *>(ap.dtwM);
}

// Plan of one transform length on the current device, built and uploaded
// on first use and cached until the process ends (fft.hip).  Lengths must
// factor into 2, 3, 5, 7, 11 and 13 and fit one pencil in LDS_MAX.
template <typename T>
int make_axis(AxisPlan &ap, size_t M, const char *what);

// The y pass over planes [k0, k0 + nk) of the half spectrum in the
// register form (fft_reg.h), when M1 has a split: forward reads rows
// [lo, hi) and stores every row; inverse reads every row and stores rows
// [lo, hi).  False: no split, nothing launched.
template <typename T>
bool y_reg(Cx<T> *work, const Vol &v, int k0, int nk, const AxisPlan &ap, bool forward, int lo,
           int hi, const int *skip, hipStream_t st, hipError_t &err)
{
    reg::PenGeom pg;
    pg.H = v.M0;
    pg.npen = v.M0 * nk;
    pg.base0 = size_t(k0) * v.M0 * v.M1;
    pg.rstep = size_t(v.M0) * v.M1;
    pg.istride = size_t(v.M0);
    pg.lo_in = forward ? lo : 0;
    pg.hi_in = forward ? hi : v.M1;
    pg.lo_out = forward ? 0 : lo;
    pg.hi_out = forward ? v.M1 : hi;
    const reg::NoOp<T> none;
    return forward
        ? reg::launch<T, reg::FWD>(v.M1, work, work, pg, twM_of<T>(ap), none, skip, st, err)
        : reg::launch<T, reg::LOAD_K | reg::INV>(v.M1, work, work, pg, twM_of<T>(ap), none,
                                                 skip, st, err);
}

// The z pass in the register form, when M2 has a split: pencils (x, y) of
// the half spectrum, planes [zlo, zhi) read (n-order loads) and stored
// (n-order stores); MODE as reg::pen4.
template <typename T, int MODE, typename Op>
bool z_reg(Cx<T> *dst, const Cx<T> *src, const Vol &v, const AxisPlan &ap, int zlo, int zhi,
           const Op &op, const int *skip, hipStream_t st, hipError_t &err)
{
    reg::PenGeom pg;
    pg.H = v.M0;
    pg.npen = v.M0 * v.M1;
    pg.base0 = 0;
    pg.rstep = size_t(v.M0);
    pg.istride = size_t(v.M0) * v.M1;
    pg.lo_in = zlo;
    pg.hi_in = zhi;
    pg.lo_out = zlo;
    pg.hi_out = zhi;
    return reg::launch<T, MODE>(v.M2, dst, src, pg, twM_of<T>(ap), op, skip, st, err);
}

// A real convolution on the stream: out = crop(real(IFFT3(op . FFT3(pad(in)))))
// over the half spectrum, with the five passes described at the top.  M:
// transform lengths, off: leading zero padding, sz: the real volume.  The
// forward x pass doubles the spectrum (op sees 2 FFT3(pad(in)) and its
// result is transformed back without the 1/(M0 M1 M2)); work holds
// (M0/2 + 1) M1 M2 complex values and needs no clearing.  With
// ROUND = false the passes stop after op: work then holds op(2 FFT3(pad(in)))
// on every plane, and out is not touched.  skip (device flag, may be NULL):
// the passes do nothing when it is set at their turn on the stream.
// Inverse y pass on the output planes, then the inverse x pass on the
// output rows (two per transform, cropped into out).
template <typename T>
int inverse_yx(Cx<T> *work, T *out, const RealGeom &g, const Vol &v, const AxisPlan ap[3],
               hipStream_t st, const int *skip)
{
    const size_t cs = sizeof(Cx<T>);
    const int H = v.M0, M1 = v.M1, M0 = int(ap[0].pl.M);
    const int ny = g.ny, nz = g.nz;
    hipError_t e = hipSuccess;
    if (y_reg<T>(work, v, g.o2, nz, ap[1], false, g.o1, g.o1 + ny, skip, st, e)) {
        QSM_HIP(e);
    } else {
        const Pen pny = pencils(M1, cs);
        const int Py = 1 << pny.logP;
        cy_pass<T><<<((H + Py - 1) / Py) * nz, FFT_THREADS, lds_bytes(pny, M1, cs), st>>>(
            work, v, g.o2, pny, ap[1].dpl, tw_of<T>(ap[1]), T(1), 0, M1, g.o1, g.o1 + ny, skip);
        QSM_HIP(hipGetLastError());
    }
    if (reg::launch_x<T>(M0, false, work, nullptr, out, g, twM_of<T>(ap[0]), skip, st, e)) {
        QSM_HIP(e);
        return QSM_OK;
    }
    const Pen pn = pencils(M0, cs);
    const int P2 = 2 << pn.logP;
    rx_inv<T><<<(ny * nz + P2 - 1) / P2, FFT_THREADS, lds_bytes(pn, M0, cs), st>>>(
        work, out, g, pn, ap[0].dpl, tw_of<T>(ap[0]), skip);
    QSM_HIP(hipGetLastError());
    return QSM_OK;
}

template <typename T, bool ROUND, typename Op>
int real_conv(const T *in, T *out, const size_t sz[3], const size_t M[3], const size_t off[3],
              Cx<T> *work, Op op, hipStream_t st, const int *skip = nullptr)
{
    int rc;
    AxisPlan ap[3];
    for (int a = 0; a < 3; ++a)
        if ((rc = make_axis<T>(ap[a], M[a], "transform length"))) return rc;
    if ((rc = allow_lds<&rx_fwd<T>>()) || (rc = allow_lds<&cy_pass<T>>())
        || (rc = allow_lds<&z_pass<T, ROUND, Op>>()) || (rc = allow_lds<&rx_inv<T>>()))
        return rc;
    const int M0 = int(M[0]), M1 = int(M[1]), M2 = int(M[2]);
    const int H = M0 / 2 + 1;
    const Vol v{H, M1, M2};  // the volume the y and z passes see
    const RealGeom g{int(sz[0]), int(sz[1]), int(sz[2]), int(off[0]), int(off[1]), int(off[2]),
                     H, M1};
    const int ny = g.ny, nz = g.nz;
    const size_t cs = sizeof(Cx<T>);
    hipError_t e = hipSuccess;
    // x forward on data rows, two per transform
    if (reg::launch_x<T>(M0, true, work, in, nullptr, g, twM_of<T>(ap[0]), skip, st, e)) {
        QSM_HIP(e);
    } else {
        const Pen pn = pencils(M0, cs);
        const int P2 = 2 << pn.logP;
        rx_fwd<T><<<(ny * nz + P2 - 1) / P2, FFT_THREADS, lds_bytes(pn, M0, cs), st>>>(
            work, in, g, pn, ap[0].dpl, tw_of<T>(ap[0]), skip);
        QSM_HIP(hipGetLastError());
    }
    // y forward on data planes
    if (y_reg<T>(work, v, g.o2, nz, ap[1], true, g.o1, g.o1 + ny, skip, st, e)) {
        QSM_HIP(e);
    } else {
        const Pen pny = pencils(M1, cs);
        const int Py = 1 << pny.logP;
        cy_pass<T><<<((H + Py - 1) / Py) * nz, FFT_THREADS, lds_bytes(pny, M1, cs), st>>>(
            work, v, g.o2, pny, ap[1].dpl, tw_of<T>(ap[1]), T(-1), g.o1, g.o1 + ny, 0, M1, skip);
        QSM_HIP(hipGetLastError());
    }
    // z forward, op (, z inverse)
    constexpr int ZM = ROUND ? (reg::FWD | reg::OP | reg::INV) : (reg::FWD | reg::OP);
    if (z_reg<T, ZM>(work, work, v, ap[2], g.o2, g.o2 + nz, op, skip, st, e)) {
        QSM_HIP(e);
    } else {
        const Pen pn = pencils(M2, cs);
        const int P = 1 << pn.logP;
        z_pass<T, ROUND, Op><<<((H + P - 1) / P) * M1, FFT_THREADS, lds_bytes(pn, M2, cs), st>>>(
            work, v, pn, ap[2].dpl, tw_of<T>(ap[2]), g.o2, g.o2 + nz, op, skip);
        QSM_HIP(hipGetLastError());
    }
    if (!ROUND)
        return QSM_OK;
    return inverse_yx<T>(work, out, g, v, ap, st, skip);
}

// The convolution's second half from a stored spectrum S (real_conv<T,
// false>'s output for the same geometry): op, inverse z into work, then the
// inverse y and x passes into out.
template <typename T, typename Op>
int real_conv_from(const Cx<T> *S, T *out, const size_t sz[3], const size_t M[3],
                   const size_t off[3], Cx<T> *work, Op op, hipStream_t st)
{
    int rc;
    AxisPlan ap[3];
    for (int a = 0; a < 3; ++a)
        if ((rc = make_axis<T>(ap[a], M[a], "transform length"))) return rc;
    if ((rc = allow_lds<&cy_pass<T>>()) || (rc = allow_lds<&z_inv_pass<T, Op>>())
        || (rc = allow_lds<&rx_inv<T>>()))
        return rc;
    const int M0 = int(M[0]), M1 = int(M[1]), M2 = int(M[2]);
    const int H = M0 / 2 + 1;
    const Vol v{H, M1, M2};
    const RealGeom g{int(sz[0]), int(sz[1]), int(sz[2]), int(off[0]), int(off[1]), int(off[2]),
                     H, M1};
    const size_t cs = sizeof(Cx<T>);
    hipError_t e = hipSuccess;
    if (z_reg<T, reg::LOAD_K | reg::OP | reg::INV>(work, S, v, ap[2], g.o2, g.o2 + g.nz, op,
                                                     nullptr, st, e)) {
        QSM_HIP(e);
    } else {
        const Pen pn = pencils(M2, cs);
        const int P = 1 << pn.logP;
        z_inv_pass<T, Op><<<((H + P - 1) / P) * M1, FFT_THREADS, lds_bytes(pn, M2, cs), st>>>(
            S, work, v, pn, ap[2].dpl, tw_of<T>(ap[2]), g.o2, g.o2 + g.nz, op);
        QSM_HIP(hipGetLastError());
    }
    return inverse_yx<T>(work, out, g, v, ap, st, nullptr);
}

// In-place complex 3D transform of data (M[0] x M[1] x M[2], x fastest):
// fft3.m:5 (inverse = 0) or ifft3.m:5 (inverse = 1, 1/M per axis), axis 3
// then 2 then 1 as fftn.  Asynchronous on the stream.
template <typename T>
int fft3_inplace(Cx<T> *data, const size_t Msz[3], int inverse, hipStream_t st)
{
    AxisPlan ap[3];
    int rc;
    for (int a = 0; a < 3; ++a)
        if ((rc = make_axis<T>(ap[a], Msz[a], "FFT length"))) return rc;
    if ((rc = allow_lds<&pass_kernel<T>>())) return rc;
    const Vol v{int(Msz[0]), int(Msz[1]), int(Msz[2])};
    const T sgn = inverse ? T(1) : T(-1);
    for (int a = 2; a >= 0; --a) {
        const Plan &pl = ap[a].pl;
        if (pl.M == 1) continue;
        const Pen pn = pencils(pl.M, sizeof(Cx<T>));
        const int P = 1 << pn.logP;
        const size_t lds = lds_bytes(pn, pl.M, sizeof(Cx<T>));
        int nblk;
        if (a == 0) nblk = (v.M1 * v.M2 + P - 1) / P;
        else nblk = ((v.M0 + P - 1) / P) * (a == 1 ? v.M2 : v.M1);
        const T scale = inverse ? T(1.0 / double(pl.M)) : T(1);
        pass_kernel<T><<<nblk, FFT_THREADS, lds, st>>>(data, v, a, pn, ap[a].dpl, tw_of<T>(ap[a]),
                                                       sgn, scale);
        QSM_HIP(hipGetLastError());
    }
    return QSM_OK;
}

// padfastfft.m:34-51 ('both'): padded length and leading pad of one axis.
inline size_t next7(size_t n)
{
    // smallest 2,3,5,7-smooth number >= n (nextfastfft.m semantics for
    // n <= 2^20; above that the value is halved, looked up and doubled)
    size_t red = 0;
    double v = double(n);
    while (v > 1048576.0) { v /= 2.0; ++red; }
    size_t best = 0;
    for (size_t p2 = 1; p2 <= (1u << 21); p2 *= 2)
        for (size_t p3 = p2; p3 <= (1u << 21); p3 *= 3)
            for (size_t p5 = p3; p5 <= (1u << 21); p5 *= 5)
                for (size_t p7 = p5; p7 <= (1u << 21); p7 *= 7)
                    if (double(p7) >= v && (best == 0 || p7 < best))
                        best = p7;
    return best << red;
}

inline void pad_plan(const size_t sz[3], const size_t pad[3], size_t M[3], size_t off[3])
{
    for (int a = 0; a < 3; ++a) {
        if (pad[a] == 0) {
            M[a] = sz[a];
            off[a] = 0;
            continue;
        }
        M[a] = next7(sz[a] + 2 * pad[a]);
        const size_t tot = M[a] - sz[a];
        const size_t oddpre = (tot & 1) & (sz[a] & 1);
        off[a] = tot / 2 + oddpre;
    }
}

// Bytes of work space real_conv needs.
inline size_t real_conv_work_bytes(const size_t M[3], size_t real_size)
{
    return (M[0] / 2 + 1) * M[1] * M[2] * 2 * real_size;
}

}  // namespace fft
}  // namespace qsm
