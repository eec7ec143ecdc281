// fft.hip -- 3D FFT and the FFT dipole convolution for gfx950.
//
// Reference idioms replaced: fft3.m / ifft3.m (separable c2c transforms;
// ifft includes 1/M), dipoleKernel.m:56-66 (k-space dipole kernel) and
// padfastfft.m / unpadfastfft.m ('both' zero padding to 7-smooth sizes), in
// the composition real(ifft3(D .* fft3(pad(chi)))) cropped back.
//
// One kernel = one axis pass over a batch of pencils: a block loads P
// pencils of length M into LDS (coalesced: rows along x, or P consecutive x
// positions for the y/z axes), runs a mixed-radix Stockham transform in LDS
// (radices 8, 7, 5, 4, 3, 2; twiddles from a table computed in double from
// exact rationals, cast for single precision), and stores.  The dipole
// convolution uses five passes over a half spectrum along x (the input is
// real): x (two real rows per complex transform, zero padding fused into
// the load, only rows that hold data, H = M0/2 + 1 outputs per row kept),
// y (only planes that hold data, zero rows not read), z forward * D(k) * z
// inverse in one pass (D generated in-kernel; only the data planes read and
// stored), y inverse (only data rows stored), x inverse (two rows per
// complex transform rebuilt from their half spectra; crop and real output
// fused into the store).  The padding is never materialised.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "common.h"

using namespace qsm;

// No result here is bitwise-pinned (the FFT is checked against numpy to a
// tolerance), so multiply-adds may fuse.
#pragma clang fp contract(fast)

namespace {

constexpr int MAX_STAGES = 24;
#ifndef QSM_FFT_THREADS
#define QSM_FFT_THREADS 512
#endif
constexpr int FFT_THREADS = QSM_FFT_THREADS;

template <typename T>
struct Cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

template <typename T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) { return Cx<T>{a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b) { return Cx<T>{a.re - b.re, a.im - b.im}; }

// Stage plan of one transform length (host-built, passed by value).
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
// R in {2, 3, 4, 5, 7, 8}, with the textbook symmetric-pair butterflies
// (constants to 22 digits; cast to T).
template <typename T>
__device__ __forceinline__ Cx<T> mul_i(Cx<T> a, T sgn)  // a * (sgn*i)
{
    return Cx<T>{-sgn * a.im, sgn * a.re};
}

template <typename T>
__device__ __forceinline__ Cx<T> scale(Cx<T> a, T c)
{
    return Cx<T>{a.re * c, a.im * c};
}

template <typename T>
__device__ __forceinline__ void dft4(Cx<T> &a0, Cx<T> &a1, Cx<T> &a2, Cx<T> &a3, T sgn)
{
    const Cx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const Cx<T> t2 = cadd(a1, a3), t3 = mul_i(csub(a1, a3), sgn);
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <typename T, int R>
__device__ __forceinline__ void small_dft(Cx<T> *v, T sgn)
{
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
__device__ __forceinline__ Cx<T> *stockham(Cx<T> *A, Cx<T> *B, Pen pn, const Plan *__restrict__ pl,
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
        default: stage<T, 8>(A, B, pn, M, Ns, mNs, tws, sgn); break;
        }
        __syncthreads();
        Cx<T> *t = A; A = B; B = t;
    }
    return A;
}

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

// ---- generic in-place pass (plain fft3 / ifft3) -------------------------
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void fft_pass_kernel(
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

// ---- dipole convolution passes -------------------------------------------

// x pass, forward: reads the real input (nx, ny, nz) zero-padded by
// (o0, o1, o2) voxels before, for rows (j, k) inside the data region only.
// Rows 2p and 2p+1 of the block travel as the real and imaginary parts of
// pencil p; with Z = FFT(a + i b) and W(k) = conj(Z(M - k)) the row spectra
// are A = (Z + W)/2 and B = (Z - W)/(2i).  Stored: 2A and 2B for
// k = 0..H-1 (the 1/2 is folded into the z pass's scaling) into the half
// spectrum volume (H, M1, M2).
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void dip_x_fwd(
    Cx<T> *data, const T *chi, int H, int M1, int nx, int ny, int nz, int o0, int o1, int o2,
    Pen pn, const Plan *pl, const Cx<T> *tw)
{
    extern __shared__ unsigned char smem[];
    Cx<T> *A = reinterpret_cast<Cx<T> *>(smem);
    const int P = 1 << pn.logP, PP = pn.PP;
    Cx<T> *B = A + size_t(PP) * pl->M;
    const int M = pl->M;
    const int rows = ny * nz;
    const int r0 = blockIdx.x * 2 * P;
    const unsigned mM = pl->mM;
    fill(A, P * M, [&](int e, int &d) {
        const int p = int(udiv(unsigned(e), mM)), i = e - p * M;
        const int ra = r0 + 2 * p;
        T a = T(0), b = T(0);
        if (i >= o0 && i < o0 + nx) {
            if (ra < rows) a = chi[size_t(i - o0) + size_t(nx) * ra];
            if (ra + 1 < rows) b = chi[size_t(i - o0) + size_t(nx) * (ra + 1)];
        }
        d = i * PP + p;
        return Cx<T>{a, b};
    });
    __syncthreads();
    Cx<T> *R = stockham(A, B, pn, pl, tw, T(-1));
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
        const int ja = ra % ny + o1, ka = ra / ny + o2;
        data[size_t(k) + size_t(H) * (size_t(ja) + size_t(M1) * ka)] = sa;
        if (ra + 1 < rows) {
            const int jb = (ra + 1) % ny + o1, kb = (ra + 1) / ny + o2;
            data[size_t(k) + size_t(H) * (size_t(jb) + size_t(M1) * kb)] = sb;
        }
    }
}

// y pass over planes k in [k0, k0 + nk) (forward or inverse).  Rows
// outside [ylo, yhi) are zero on input (not read); only rows in
// [slo, shi) are stored (the later passes read no others).
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void dip_y(
    Cx<T> *data, Vol v, int k0, Pen pn, const Plan *pl, const Cx<T> *tw, T sgn,
    int ylo, int yhi, int slo, int shi)
{
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

// k-space coordinate of index m on an axis of length n, voxel size h
// (dipoleKernel.m:74-87 followed by ifftshift, :64).
// dk = 1 / (n h), formed on the host.
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

// z pass over the half spectrum (x extent v.M0 = H of a true length mx):
// forward transform, multiply by Ds(k) * scale, inverse.  The reference keeps
// real(ifftn(D .* fftn(x))); for a real x that is the inverse transform of
// fftn(x) .* Ds with Ds(k) = (D(k) + D(-k))/2, the form a half spectrum can
// carry.  Ds = D except where an even axis is at its Nyquist index (there
// -k has the same index, with the other components negated).  Only planes
// [zlo, zhi) hold data on input and only they are needed on output: the
// others are neither read (zero) nor stored.
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void dip_z(
    Cx<T> *data, Vol v, int mx, Pen pn, const Plan *pl, const Cx<T> *tw,
    double dk0, double dk1, double dk2, double b0, double b1, double b2, double scale,
    int zlo, int zhi)
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
        return (p < np && i >= zlo && i < zhi) ? data[base + p + size_t(i) * plane]
                                               : Cx<T>{T(0), T(0)};
    });
    __syncthreads();
    Cx<T> *R = stockham(A, B, pn, pl, tw, T(-1));
    Cx<T> *S = (R == A) ? B : A;
    const double Y = kcoord(j, v.M1, dk1);
    const bool ny_y = 2 * j == v.M1;
    for (int e = threadIdx.x; e < P * M; e += FFT_THREADS) {
        const int i = e >> logP, p = e & (P - 1);
        const double X = kcoord(x0 + p, mx, dk0);
        const double Z = kcoord(i, M, dk2);
        double D = dip(X, Y, Z, b0, b1, b2);
        const bool ny_x = 2 * (x0 + p) == mx, ny_z = 2 * i == M;
        if (ny_x || ny_y || ny_z)
            D = 0.5 * (D + dip(ny_x ? X : -X, ny_y ? Y : -Y, ny_z ? Z : -Z, b0, b1, b2));
        const T d = T(D * scale);
        Cx<T> x = R[i * PP + p];
        x.re *= d;
        x.im *= d;
        R[i * PP + p] = x;
    }
    __syncthreads();
    Cx<T> *Q = stockham(R, S, pn, pl, tw, T(1));
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
__global__ __launch_bounds__(FFT_THREADS) void dip_x_inv(
    const Cx<T> *data, T *out, int H, int M1, int nx, int ny, int nz, int o0, int o1, int o2,
    Pen pn, const Plan *pl, const Cx<T> *tw)
{
    extern __shared__ unsigned char smem[];
    Cx<T> *A = reinterpret_cast<Cx<T> *>(smem);
    const int P = 1 << pn.logP, PP = pn.PP;
    Cx<T> *B = A + size_t(PP) * pl->M;
    const int M = pl->M;
    const int rows = ny * nz;
    const int r0 = blockIdx.x * 2 * P;
    const unsigned mM = pl->mM;
    fill(A, P * M, [&](int e, int &d) {
        const int p = int(udiv(unsigned(e), mM)), i = e - p * M;
        const int ra = r0 + 2 * p;
        const bool lo = i < H;
        const int k = lo ? i : M - i;
        Cx<T> a{T(0), T(0)}, b{T(0), T(0)};
        if (ra < rows) {
            const int j = ra % ny + o1, kk = ra / ny + o2;
            a = data[size_t(k) + size_t(H) * (size_t(j) + size_t(M1) * kk)];
        }
        if (ra + 1 < rows) {
            const int j = (ra + 1) % ny + o1, kk = (ra + 1) / ny + o2;
            b = data[size_t(k) + size_t(H) * (size_t(j) + size_t(M1) * kk)];
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
              if (i >= o0 && i < o0 + nx) {
                  if (ra < rows) out[size_t(i - o0) + size_t(nx) * ra] = x.re;
                  if (ra + 1 < rows) out[size_t(i - o0) + size_t(nx) * (ra + 1)] = x.im;
              }
          });
}

// ---- host-side planning ----------------------------------------------------

bool factor(int M, Plan &pl)
{
    pl.M = M;
    pl.mM = recip(unsigned(M));
    pl.nst = 0;
    int rest = M, Ns = 1, off = 0;
    const int radices[] = {8, 7, 5, 4, 3, 2};
    while (rest > 1) {
        int R = 0;
        for (int r : radices)
            if (rest % r == 0) { R = r; break; }
        if (!R || pl.nst >= MAX_STAGES)
            return false;
        pl.R[pl.nst] = R;
        pl.Ns[pl.nst] = Ns;
        pl.mNs[pl.nst] = recip(unsigned(Ns));
        pl.tw[pl.nst] = off;
        off += Ns * (R - 1);
        Ns *= R;
        rest /= R;
        ++pl.nst;
    }
    return true;
}

// exp(-2*pi*i*a/b) in double from the exact rational a/b: the angle is
// reduced in long double before the sine and cosine.
Cx<double> twiddle(long a, long b)
{
    a %= b;
    const long double t = (long double)a / (long double)b;  // in [0, 1)
    const long double ang = 2.0L * 3.141592653589793238462643383279502884L * t;
    return Cx<double>{double(cosl(ang)), double(-sinl(ang))};
}

std::vector<Cx<double>> twiddles(const Plan &pl)
{
    std::vector<Cx<double>> t;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.R[s], Ns = pl.Ns[s];
        for (int k = 0; k < Ns; ++k)
            for (int r = 1; r < R; ++r)
                t.push_back(twiddle(long(r) * k, long(Ns) * R));
    }
    if (t.empty())
        t.push_back(Cx<double>{1.0, 0.0});
    return t;
}

// LDS per block: two buffers of P pencils.  80 KB lets two blocks share a
// CU's 160 KB.
constexpr size_t LDS_BUDGET = 80 * 1024;
// one pencil of the longest supported length (2048 double, 4096 single)
constexpr size_t LDS_MAX = 64 * 1024;
static_assert(LDS_MAX <= LDS_BUDGET, "the pass kernels are allowed LDS_BUDGET bytes");

size_t lds_bytes(Pen pn, int M, size_t csize) { return 2 * size_t(pn.PP) * M * csize; }

// Pencils per block: the largest power of two P <= 16 (16 consecutive x
// positions are whole cache lines) whose buffers fit the budget.
Pen pencils(int M, size_t csize)
{
    for (int logP = 4; logP > 0; --logP) {
        const Pen pn{logP, (1 << logP) + 1};
        if (lds_bytes(pn, M, csize) <= LDS_BUDGET)
            return pn;
    }
    return Pen{0, 1};
}

// Lets the pass kernels use more than the default 64 KB of dynamic LDS.
template <typename T>
int allow_lds()
{
    static const hipError_t e = [] {
        const void *k[] = {reinterpret_cast<const void *>(&fft_pass_kernel<T>),
                           reinterpret_cast<const void *>(&dip_x_fwd<T>),
                           reinterpret_cast<const void *>(&dip_y<T>),
                           reinterpret_cast<const void *>(&dip_z<T>),
                           reinterpret_cast<const void *>(&dip_x_inv<T>)};
        for (const void *f : k) {
            const hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                     int(LDS_BUDGET));
            if (r != hipSuccess)
                return r;
        }
        return hipSuccess;
    }();
    QSM_HIP(e);
    return QSM_OK;
}

struct AxisPlan {
    Plan pl;                    // host copy
    const Plan *dpl = nullptr;  // device copy (head of the buffer)
    const void *dtw = nullptr;  // device twiddle table in the transform's
                                // precision (after the plan); the single
                                // table is the double one rounded
};

template <typename T>
const Cx<T> *tw_of(const AxisPlan &ap)
{
    return static_cast<const Cx<T> *>(ap.dtw);
}

// Plans are cached per (device, length, precision): built and uploaded on first use,
// kept until the process ends (a few KB each), so the transforms launch
// without host synchronisation after the first call.  The device memory is
// deliberately never freed (freeing at static destruction would race the
// HIP runtime's own teardown).
struct PlanCache {
    std::mutex mu;
    std::map<std::pair<int, size_t>, AxisPlan> plans;
};

PlanCache &plan_cache()
{
    static PlanCache *c = new PlanCache;  // never destroyed, like its plans
    return *c;
}

// Factors M, builds its twiddles and uploads plan + table (once per device).
template <typename T>
int make_axis(AxisPlan &ap, size_t M, const char *what)
{
    QSM_ARG(M >= 1 && lds_bytes(Pen{0, 1}, int(M), sizeof(Cx<T>)) <= LDS_MAX,
            "%s %zu too long (at most %zu)", what, M, LDS_MAX / (2 * sizeof(Cx<T>)));
    int dev = 0;
    QSM_HIP(hipGetDevice(&dev));
    PlanCache &pc = plan_cache();
    std::lock_guard<std::mutex> lock(pc.mu);
    const auto key = std::make_pair(dev, M * 16 + sizeof(T));
    const auto it = pc.plans.find(key);
    if (it != pc.plans.end()) {
        ap = it->second;
        return QSM_OK;
    }
    AxisPlan np;
    QSM_ARG(factor(int(M), np.pl), "%s %zu is not 2,3,5,7-smooth", what, M);
    const std::vector<Cx<double>> td = twiddles(np.pl);
    std::vector<Cx<T>> t(td.size());
    for (size_t i = 0; i < td.size(); ++i)
        t[i] = Cx<T>{T(td[i].re), T(td[i].im)};
    // plan at the head, the table after it on a 256-byte boundary
    constexpr size_t head = (sizeof(Plan) + 255) / 256 * 256;
    void *p = nullptr;
    QSM_HIP(hipMalloc(&p, head + t.size() * sizeof(t[0])));
    if (hipMemcpy(p, &np.pl, sizeof(Plan), hipMemcpyHostToDevice) != hipSuccess
        || hipMemcpy(static_cast<char *>(p) + head, t.data(), t.size() * sizeof(t[0]),
                     hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        set_error("FFT plan upload failed");
        return QSM_EHIP;
    }
    np.dpl = static_cast<const Plan *>(p);
    np.dtw = static_cast<const char *>(p) + head;
    pc.plans.emplace(key, np);
    ap = np;
    return QSM_OK;
}

template <typename T>
int fft3_dev_t(void *data, const size_t Msz[3], int inverse, hipStream_t st)
{
    AxisPlan ap[3];
    int rc;
    for (int a = 0; a < 3; ++a)
        if ((rc = make_axis<T>(ap[a], Msz[a], "FFT length"))) return rc;
    if ((rc = allow_lds<T>())) return rc;
    const Vol v{int(Msz[0]), int(Msz[1]), int(Msz[2])};
    const T sgn = inverse ? T(1) : T(-1);
    // fft3.m:5 and ifft3.m:5 apply axis 3, then 2, then 1
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
        fft_pass_kernel<T><<<nblk, FFT_THREADS, lds, st>>>(
            static_cast<Cx<T> *>(data), v, a, pn, ap[a].dpl, tw_of<T>(ap[a]), sgn, scale);
        QSM_HIP(hipGetLastError());
    }
    return QSM_OK;
}

template <typename T>
int dipole_conv_dev_t(const void *chi, const size_t sz[3], const double h[3],
                      const double bdir[3], const size_t M[3], const size_t off[3],
                      void *field, void *work, hipStream_t st)
{
    int rc;
    AxisPlan ap[3];
    for (int a = 0; a < 3; ++a)
        if ((rc = make_axis<T>(ap[a], M[a], "padded length"))) return rc;
    if ((rc = allow_lds<T>())) return rc;
    const int M0 = int(M[0]), M1 = int(M[1]), M2 = int(M[2]);
    const int H = M0 / 2 + 1;        // half spectrum along x
    const Vol v{H, M1, M2};          // the volume the y and z passes see
    const int nx = int(sz[0]), ny = int(sz[1]), nz = int(sz[2]);
    const int o0 = int(off[0]), o1 = int(off[1]), o2 = int(off[2]);
    Cx<T> *d = static_cast<Cx<T> *>(work);
    const double total = double(M0) * M1 * M2;
    // The padding is never stored: the x pass writes the rows that hold
    // data, and the y and z passes treat the other rows and planes as zero
    // on load.  The work space therefore needs no clearing.
    // normalised b (dipoleKernel.m:35-38)
    const double bn = std::sqrt(bdir[0] * bdir[0] + bdir[1] * bdir[1] + bdir[2] * bdir[2]);
    QSM_ARG(bn > 0.0, "B0 direction must be non-zero");
    double b[3] = {bdir[0], bdir[1], bdir[2]};
    if (bn != 1.0)
        for (double &x : b) x /= bn;
    const size_t cs = sizeof(Cx<T>);
    {   // x forward on data rows, two per transform
        const Plan &pl = ap[0].pl;
        const Pen pn = pencils(pl.M, cs);
        const int P2 = 2 << pn.logP;
        dip_x_fwd<T><<<(ny * nz + P2 - 1) / P2, FFT_THREADS, lds_bytes(pn, pl.M, cs), st>>>(
            d, static_cast<const T *>(chi), H, M1, nx, ny, nz, o0, o1, o2, pn, ap[0].dpl,
            tw_of<T>(ap[0]));
        QSM_HIP(hipGetLastError());
    }
    {   // y forward on data planes
        const Plan &pl = ap[1].pl;
        const Pen pn = pencils(pl.M, cs);
        const int P = 1 << pn.logP;
        dip_y<T><<<((H + P - 1) / P) * nz, FFT_THREADS, lds_bytes(pn, pl.M, cs), st>>>(
            d, v, o2, pn, ap[1].dpl, tw_of<T>(ap[1]), T(-1), o1, o1 + ny, 0, M1);
        QSM_HIP(hipGetLastError());
    }
    {   // z forward * Ds * z inverse, scaled by 1/(2 M0 M1 M2) (the 2 undoes
        // the x pass's doubling)
        const Plan &pl = ap[2].pl;
        const Pen pn = pencils(pl.M, cs);
        const int P = 1 << pn.logP;
        dip_z<T><<<((H + P - 1) / P) * M1, FFT_THREADS, lds_bytes(pn, pl.M, cs), st>>>(
            d, v, M0, pn, ap[2].dpl, tw_of<T>(ap[2]), 1.0 / (double(M0) * h[0]),
            1.0 / (double(M1) * h[1]), 1.0 / (double(M2) * h[2]),
            b[0], b[1], b[2], 0.5 / total, o2, o2 + nz);
        QSM_HIP(hipGetLastError());
    }
    {   // y inverse on output planes
        const Plan &pl = ap[1].pl;
        const Pen pn = pencils(pl.M, cs);
        const int P = 1 << pn.logP;
        dip_y<T><<<((H + P - 1) / P) * nz, FFT_THREADS, lds_bytes(pn, pl.M, cs), st>>>(
            d, v, o2, pn, ap[1].dpl, tw_of<T>(ap[1]), T(1), 0, M1, o1, o1 + ny);
        QSM_HIP(hipGetLastError());
    }
    {   // x inverse on output rows, two per transform, cropped
        const Plan &pl = ap[0].pl;
        const Pen pn = pencils(pl.M, cs);
        const int P2 = 2 << pn.logP;
        dip_x_inv<T><<<(ny * nz + P2 - 1) / P2, FFT_THREADS, lds_bytes(pn, pl.M, cs), st>>>(
            d, static_cast<T *>(field), H, M1, nx, ny, nz, o0, o1, o2, pn, ap[0].dpl,
            tw_of<T>(ap[0]));
        QSM_HIP(hipGetLastError());
    }
    return QSM_OK;
}

// padfastfft.m:34-51 ('both'): padded length and leading pad of one axis.
size_t next7(size_t n)
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

void pad_plan(const size_t sz[3], const size_t pad[3], size_t M[3], size_t off[3])
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

}  // namespace

extern "C" {

int qsm_fft3_dev(int dtype, void *data, const size_t M[3], int inverse, void *stream)
{
    QSM_ARG(data && M, "null pointer");
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    return dtype == QSM_F64 ? fft3_dev_t<double>(data, M, inverse, as_stream(stream))
                            : fft3_dev_t<float>(data, M, inverse, as_stream(stream));
}

size_t qsm_nextfastfft(size_t n) { return n == 0 ? 1 : next7(n); }

int qsm_dipole_pad_size(const size_t sz[3], const size_t pad[3], size_t M[3], size_t off[3])
{
    QSM_ARG(sz && pad && M && off, "null pointer");
    pad_plan(sz, pad, M, off);
    return QSM_OK;
}

size_t qsm_dipole_work_bytes(int dtype, const size_t sz[3], const size_t pad[3])
{
    size_t M[3], off[3];
    pad_plan(sz, pad, M, off);
    return (M[0] / 2 + 1) * M[1] * M[2] * 2 * dtype_size(dtype);  // half spectrum along x
}

int qsm_dipole_conv_dev(int dtype, const void *chi, const size_t sz[3],
                        const double h[3], const double bdir[3], const size_t pad[3],
                        void *field, void *work, void *stream)
{
    int rc = check_grid(sz, h, 1);
    if (rc) return rc;
    QSM_ARG(chi && field && work && bdir && pad, "null pointer");
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    size_t M[3], off[3];
    pad_plan(sz, pad, M, off);
    return dtype == QSM_F64
        ? dipole_conv_dev_t<double>(chi, sz, h, bdir, M, off, field, work, as_stream(stream))
        : dipole_conv_dev_t<float>(chi, sz, h, bdir, M, off, field, work, as_stream(stream));
}

int qsm_dipole_conv(int dtype, const void *chi, const size_t sz[3], const double h[3],
                    const double bdir[3], const size_t pad[3], void *field)
{
    int rc = check_grid(sz, h, 1);
    if (rc) return rc;
    QSM_ARG(chi && field && bdir && pad, "null pointer");
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    if ((rc = ensure_device())) return rc;
    const size_t N = sz[0] * sz[1] * sz[2], es = dtype_size(dtype);
    DevBuf dc, df, dw;
    if ((rc = dc.alloc(N * es)) || (rc = df.alloc(N * es))
        || (rc = dw.alloc(qsm_dipole_work_bytes(dtype, sz, pad))))
        return rc;
    QSM_HIP(hipMemcpy(dc.p, chi, N * es, hipMemcpyHostToDevice));
    if ((rc = qsm_dipole_conv_dev(dtype, dc.p, sz, h, bdir, pad, df.p, dw.p, nullptr)))
        return rc;
    QSM_HIP(hipMemcpy(field, df.p, N * es, hipMemcpyDeviceToHost));
    return QSM_OK;
}

}  // extern "C"
