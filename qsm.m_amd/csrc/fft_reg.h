// fft_reg.h -- register-resident pencil transforms for the y and z passes
// of the real convolution (fft_dev.h), gfx950.
//
// A pencil of length M = N1 * N2 is transformed in the "four-step" form
// with n = N2 n1 + n2 and k = k1 + N1 k2:
//
//   X[k1 + N1 k2] = sum_n2 W_N2^(n2 k2) W_M^(n2 k1) sum_n1 W_N1^(n1 k1) x[N2 n1 + n2]
//
// A team of max(N1, N2) threads holds one pencil.  Thread n2 loads x[N2 n1
// + n2] for every n1 (n order), transforms them in registers (an N1-point
// DFT, itself split into small radices with compile-time twiddles),
// multiplies by W_M^(n2 k1), and the values cross once through LDS so that
// thread k1 holds every n2 and finishes with an N2-point DFT (k order:
// X[k1 + N1 k2] for every k2).  The inverse runs the same steps backwards
// (conjugate twiddles), so a forward transform, a k-space operation and the
// inverse cost two LDS exchanges in all, against one per radix stage (and
// the fill and drain) of the generic Stockham kernels.
//
// P consecutive pencils share a block, lanes running over them, so every
// global access of a wave is a run of consecutive x positions.
#pragma once

#include <type_traits>

namespace qsm {
namespace fft {
namespace reg {

// ---- compile-time roots of unity -------------------------------------------

constexpr double kHalfPi = 1.57079632679489661923132169163975144;

constexpr double sin_series(double x)
{
    double t = x, s = x;
    for (int k = 1; k < 12; ++k) {
        t *= -(x * x) / double((2 * k) * (2 * k + 1));
        s += t;
    }
    return s;
}

constexpr double cos_series(double x)
{
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 12; ++k) {
        t *= -(x * x) / double((2 * k - 1) * (2 * k));
        s += t;
    }
    return s;
}

struct CS {
    double c, s;
};

// cos and sin of 2 pi j / n: the quadrant and the reflection about pi/4
// are taken in integer arithmetic, the series is summed on [0, pi/4].
constexpr CS unit_root(long j, long n)
{
    j %= n;
    if (j < 0) j += n;
    const long q = (4 * j) / n, r = 4 * j - q * n;  // 2 pi j / n = q pi/2 + (pi/2) r / n
    double c = 1.0, s = 0.0;
    if (2 * r <= n) {
        const double x = kHalfPi * double(r) / double(n);
        c = cos_series(x);
        s = sin_series(x);
    } else {
        const double x = kHalfPi * double(n - r) / double(n);
        c = sin_series(x);
        s = cos_series(x);
    }
    switch (q) {
    case 0: return CS{c, s};
    case 1: return CS{-s, c};
    case 2: return CS{-c, -s};
    default: return CS{s, -c};
    }
}

// ---- compile-time loops ----------------------------------------------------

template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

// a * exp(SGN 2 pi i J / N); multiplications by 0 and 1 never appear.
template <typename T, int N, int J, int SGN>
__device__ __forceinline__ Cx<T> rot(Cx<T> a)
{
#pragma clang fp contract(fast)
    constexpr int j = ((J % N) + N) % N;
    constexpr T R2 = T(0.7071067811865475244008);
    if constexpr (j == 0) {
        return a;
    } else if constexpr (2 * j == N) {
        return Cx<T>{-a.re, -a.im};
    } else if constexpr (4 * j == N) {
        return mul_i(a, T(SGN));
    } else if constexpr (4 * j == 3 * N) {
        return mul_i(a, T(-SGN));
    } else if constexpr (8 * j == N) {          // (1 + SGN i) / sqrt 2
        return scale(cadd(a, mul_i(a, T(SGN))), R2);
    } else if constexpr (8 * j == 3 * N) {      // (-1 + SGN i) / sqrt 2
        return scale(csub(mul_i(a, T(SGN)), a), R2);
    } else if constexpr (8 * j == 5 * N) {      // (-1 - SGN i) / sqrt 2
        const Cx<T> b = scale(cadd(a, mul_i(a, T(SGN))), R2);
        return Cx<T>{-b.re, -b.im};
    } else if constexpr (8 * j == 7 * N) {      // (1 - SGN i) / sqrt 2
        return scale(csub(a, mul_i(a, T(SGN))), R2);
    } else {
        constexpr CS w = unit_root(j, N);
        return cmul(a, Cx<T>{T(w.c), T(SGN * w.s)});
    }
}

template <int N>
constexpr bool is_small()
{
    return N == 1 || N == 2 || N == 3 || N == 4 || N == 5 || N == 7 || N == 8 || N == 11 ||
           N == 13;
}

template <int N>
constexpr int radix_of()
{
    return N % 4 == 0 ? 4 : N % 3 == 0 ? 3 : N % 5 == 0 ? 5 : N % 7 == 0 ? 7 : 2;
}

// In-register DFT of v[0..N-1], natural order in and out:
// v_k <- sum_n v_n exp(SGN 2 pi i n k / N).  N = A B (A a small radix):
// A-point DFTs over the B strided subsequences, twiddles W_N^(nb ka), then
// B-point DFTs; every index is a compile-time constant, so the reordering
// is register renaming.
template <typename T, int N, int SGN>
__device__ __forceinline__ void dft(Cx<T> *v)
{
    if constexpr (N == 1) {
        return;
    } else if constexpr (is_small<N>()) {
        small_dft<T, N>(v, T(SGN));
    } else {
        constexpr int A = radix_of<N>(), B = N / A;
        Cx<T> y[N];  // y[ka * B + nb]
        sfor<0, B>([&](auto nbc) {
            constexpr int nb = decltype(nbc)::value;
            Cx<T> u[A];
            sfor<0, A>([&](auto ic) {
                constexpr int na = decltype(ic)::value;
                u[na] = v[B * na + nb];
            });
            small_dft<T, A>(u, T(SGN));
            sfor<0, A>([&](auto kc) {
                constexpr int ka = decltype(kc)::value;
                y[ka * B + nb] = rot<T, N, nb * ka, SGN>(u[ka]);
            });
        });
        sfor<0, A>([&](auto kc) {
            constexpr int ka = decltype(kc)::value;
            dft<T, B, SGN>(y + ka * B);
            sfor<0, B>([&](auto bc) {
                constexpr int kb = decltype(bc)::value;
                v[ka + A * kb] = y[ka * B + kb];
            });
        });
    }
}

// ---- the pencil pass ---------------------------------------------------------

// Pencil q of a pass: x = q % H, r = q / H; element i at
// base0 + x + r * rstep + i * istride (complex elements).
struct PenGeom {
    int H, npen;
    size_t base0, rstep, istride;
    int lo_in, hi_in;    // n-order loads: elements outside are zero, not read
    int lo_out, hi_out;  // n-order stores: elements outside are not written
};

// What a pass does, in order: load (n order, or k order with LOAD_K),
// forward transform (FWD, n -> k order), op(x, r, k, value) on every sample
// (OP), inverse transform (INV, k -> n order), store (k order: every
// element; n order: [lo_out, hi_out)).
enum : int { LOAD_K = 1, FWD = 2, OP = 4, INV = 8 };

template <int N1, int N2>
constexpr int team() { return N1 > N2 ? N1 : N2; }

template <typename T, int N1, int N2, int P, int MODE, typename Op>
__global__ __launch_bounds__(P * (N1 > N2 ? N1 : N2)) void pen4(
    Cx<T> *dst, const Cx<T> *src, PenGeom g, const Cx<T> *twM, Op op, const int *skip)
{
#pragma clang fp contract(fast)
    constexpr int M = N1 * N2, TM = team<N1, N2>(), NT = P * TM;
    constexpr bool kin = (MODE & LOAD_K) != 0, fwd = (MODE & FWD) != 0;
    constexpr bool useop = (MODE & OP) != 0, inv = (MODE & INV) != 0;
    static_assert(kin != fwd, "a pass loads in n order exactly when it transforms forward");
    __shared__ Cx<T> X[P * M];
    __shared__ Cx<T> W[M];
    if (skip && *skip)
        return;
    const int t = int(threadIdx.x), p = t % P, m = t / P;
    const int q = xcd_block() * P + p;
    const bool live = q < g.npen;
    const int x = live ? q % g.H : 0, r = live ? q / g.H : 0;
    const Cx<T> *in = src + (g.base0 + size_t(x) + size_t(r) * g.rstep);
    Cx<T> *out = dst + (g.base0 + size_t(x) + size_t(r) * g.rstep);
    const size_t is = g.istride;

    Cx<T> a[N1];  // n order: x[N2 n1 + m] (m < N2)
    Cx<T> b[N2];  // k order: X[m + N1 k2] (m < N1)
    if constexpr (!kin) {
        if (m < N2) {
            sfor<0, N1>([&](auto c) {
                constexpr int n1 = decltype(c)::value;
                const int i = N2 * n1 + m;
                a[n1] = Cx<T>{T(0), T(0)};
                if (live && i >= g.lo_in && i < g.hi_in)
                    a[n1] = in[size_t(i) * is];
            });
        }
    } else {
        if (m < N1) {
            sfor<0, N2>([&](auto c) {
                constexpr int k2 = decltype(c)::value;
                b[k2] = Cx<T>{T(0), T(0)};
                if (live)
                    b[k2] = in[size_t(m + N1 * k2) * is];
            });
        }
    }
    for (int e = t; e < M; e += NT)
        W[e] = twM[e];
    __syncthreads();

    if constexpr (fwd) {
        if (m < N2) {
            dft<T, N1, -1>(a);
            sfor<0, N1>([&](auto c) {
                constexpr int k1 = decltype(c)::value;
                X[(k1 * N2 + m) * P + p] = k1 == 0 ? a[0] : cmul(a[k1], W[m * k1]);
            });
        }
        __syncthreads();
        if (m < N1) {
            sfor<0, N2>([&](auto c) {
                constexpr int n2 = decltype(c)::value;
                b[n2] = X[(m * N2 + n2) * P + p];
            });
            dft<T, N2, -1>(b);
        }
    }
    if constexpr (useop) {
        if (m < N1 && live) {
            sfor<0, N2>([&](auto c) {
                constexpr int k2 = decltype(c)::value;
                b[k2] = op(x, r, m + N1 * k2, b[k2]);
            });
        }
    }
    if constexpr (inv) {
        if constexpr (fwd)
            __syncthreads();  // every read of the forward exchange is done
        if (m < N1) {
            dft<T, N2, 1>(b);
            sfor<0, N2>([&](auto c) {
                constexpr int n2 = decltype(c)::value;
                const Cx<T> w = W[m * n2];
                X[(n2 * N1 + m) * P + p] = n2 == 0 ? b[0] : cmul(b[n2], Cx<T>{w.re, -w.im});
            });
        }
        __syncthreads();
        if (m < N2) {
            sfor<0, N1>([&](auto c) {
                constexpr int k1 = decltype(c)::value;
                a[k1] = X[(m * N1 + k1) * P + p];
            });
            dft<T, N1, 1>(a);
            if (live) {
                sfor<0, N1>([&](auto c) {
                    constexpr int n1 = decltype(c)::value;
                    const int i = N2 * n1 + m;
                    if (i >= g.lo_out && i < g.hi_out)
                        out[size_t(i) * is] = a[n1];
                });
            }
        }
    } else {
        if (m < N1 && live) {
            sfor<0, N2>([&](auto c) {
                constexpr int k2 = decltype(c)::value;
                out[size_t(m + N1 * k2) * is] = b[k2];
            });
        }
    }
}

// Pencils per block: 16 (FP32) or 8 (FP64) consecutive x positions, 128
// bytes of each row.
template <typename T>
constexpr int pencils_per_block() { return sizeof(T) == 4 ? 16 : 8; }

// Lengths with a register split (N1, N2); other lengths use the Stockham
// kernels.
inline bool split_of(int M, int &N1, int &N2)
{
    switch (M) {
    case 176: N1 = 16; N2 = 11; return true;
    case 192: N1 = 16; N2 = 12; return true;
    case 256: N1 = 16; N2 = 16; return true;
    case 320: N1 = 20; N2 = 16; return true;
    case 360: N1 = 20; N2 = 18; return true;
    case 384: N1 = 24; N2 = 16; return true;
    case 448: N1 = 28; N2 = 16; return true;
    case 512: N1 = 32; N2 = 16; return true;
    case 640: N1 = 32; N2 = 20; return true;
    default: return false;
    }
}

template <typename T, int N1, int N2, int MODE, typename Op>
hipError_t launch_one(Cx<T> *dst, const Cx<T> *src, const PenGeom &g, const Cx<T> *twM,
                      const Op &op, const int *skip, hipStream_t st)
{
    constexpr int P = pencils_per_block<T>(), NT = P * team<N1, N2>();
    const int nblk = (g.npen + P - 1) / P;
    if (nblk > 0)
        pen4<T, N1, N2, P, MODE, Op><<<nblk, NT, 0, st>>>(dst, src, g, twM, op, skip);
    return hipGetLastError();
}

// Launches the register pass of length M; false when M has no split.
template <typename T, int MODE, typename Op>
bool launch(int M, Cx<T> *dst, const Cx<T> *src, const PenGeom &g, const Cx<T> *twM,
            const Op &op, const int *skip, hipStream_t st, hipError_t &err)
{
    switch (M) {
    case 176: err = launch_one<T, 16, 11, MODE>(dst, src, g, twM, op, skip, st); return true;
    case 192: err = launch_one<T, 16, 12, MODE>(dst, src, g, twM, op, skip, st); return true;
    case 256: err = launch_one<T, 16, 16, MODE>(dst, src, g, twM, op, skip, st); return true;
    case 320: err = launch_one<T, 20, 16, MODE>(dst, src, g, twM, op, skip, st); return true;
    case 360: err = launch_one<T, 20, 18, MODE>(dst, src, g, twM, op, skip, st); return true;
    case 384: err = launch_one<T, 24, 16, MODE>(dst, src, g, twM, op, skip, st); return true;
    case 448: err = launch_one<T, 28, 16, MODE>(dst, src, g, twM, op, skip, st); return true;
    case 512: err = launch_one<T, 32, 16, MODE>(dst, src, g, twM, op, skip, st); return true;
    case 640: err = launch_one<T, 32, 20, MODE>(dst, src, g, twM, op, skip, st); return true;
    default: return false;
    }
}

// ---- x passes of the real convolution --------------------------------------
//
// Pencils run along x, so the lanes of a wave take consecutive transform
// indices of one pencil (thread index m fastest): every global access is a
// run of a row.  Each complex pencil carries two real rows, 2p and 2p + 1 of
// the block, as its real and imaginary parts.  The exchange buffer holds
// entry (a, b) of a pencil at a (N2 + 1) + b: writes run along b, reads
// along a with an odd stride, both on distinct LDS banks.

template <int N1, int N2>
constexpr int xbuf() { return N1 * (N2 + 1) > N1 * N2 + N1 ? N1 * (N2 + 1) : N1 * N2 + N1; }

// Pencils (row pairs) per block of an x pass.
template <typename T>
constexpr int xpencils() { return sizeof(T) == 4 ? 8 : 4; }

// Forward: rows (j, k) of the real input inside the data region (zero
// padding fused into the load), stored as the half spectra 2A(k), 2B(k),
// k < H, of the two rows (A = (Z(k) + conj Z(M-k))/2, B = (Z(k) - conj
// Z(M-k))/2i with Z the spectrum of the packed pencil), as the Stockham
// rx_fwd stores them.
template <typename T, int N1, int N2, int P>
__global__ __launch_bounds__(P * (N1 > N2 ? N1 : N2)) void rx_fwd_reg(
    Cx<T> *data, const T *in, RealGeom g, const Cx<T> *twM, const int *skip)
{
#pragma clang fp contract(fast)
    constexpr int M = N1 * N2, TM = team<N1, N2>(), NT = P * TM, XB = xbuf<N1, N2>();
    __shared__ Cx<T> X[P * XB];
    __shared__ Cx<T> W[M];
    if (skip && *skip)
        return;
    const int t = int(threadIdx.x), m = t % TM, p = t / TM;
    const int rows = g.ny * g.nz;
    const int ra = int(blockIdx.x) * 2 * P + 2 * p;  // rows ra, ra + 1
    Cx<T> *Xp = X + p * XB;
    for (int e = t; e < M; e += NT)
        W[e] = twM[e];
    Cx<T> a[N1];
    if (m < N2) {
        const size_t b0 = size_t(g.nx) * ra, b1 = b0 + g.nx;  // rows ra, ra + 1
        sfor<0, N1>([&](auto c) {
            constexpr int n1 = decltype(c)::value;
            const int i = N2 * n1 + m;
            T re = T(0), im = T(0);
            if (i >= g.o0 && i < g.o0 + g.nx) {
                if (ra < rows) re = in[b0 + (i - g.o0)];
                if (ra + 1 < rows) im = in[b1 + (i - g.o0)];
            }
            a[n1] = Cx<T>{re, im};
        });
    }
    __syncthreads();
    if (m < N2) {
        dft<T, N1, -1>(a);
        sfor<0, N1>([&](auto c) {
            constexpr int k1 = decltype(c)::value;
            Xp[k1 * (N2 + 1) + m] = k1 == 0 ? a[0] : cmul(a[k1], W[m * k1]);
        });
    }
    __syncthreads();
    Cx<T> b[N2];
    if (m < N1) {
        sfor<0, N2>([&](auto c) {
            constexpr int n2 = decltype(c)::value;
            b[n2] = Xp[m * (N2 + 1) + n2];
        });
        dft<T, N2, -1>(b);
    }
    __syncthreads();
    if (m < N1) {
        sfor<0, N2>([&](auto c) {
            constexpr int k2 = decltype(c)::value;
            Xp[m + N1 * k2] = b[k2];  // natural order
        });
    }
    __syncthreads();
    if (ra >= rows)
        return;
    const int H = g.H;
    const int ja = ra % g.ny + g.o1, ka = ra / g.ny + g.o2;
    const int jb = (ra + 1) % g.ny + g.o1, kb = (ra + 1) / g.ny + g.o2;
    Cx<T> *oa = data + size_t(H) * (size_t(ja) + size_t(g.M1) * ka);
    Cx<T> *ob = data + size_t(H) * (size_t(jb) + size_t(g.M1) * kb);
    for (int k = m; k < H; k += TM) {
        const Cx<T> z = Xp[k];
        const Cx<T> w = Xp[k == 0 ? 0 : M - k];  // conj taken below
        oa[k] = Cx<T>{z.re + w.re, z.im - w.im};
        if (ra + 1 < rows)
            ob[k] = Cx<T>{z.im + w.im, w.re - z.re};
    }
}

// Inverse: two rows' half spectra A, B (k < H) rebuilt into the pencil
// Z(i) = A(i) + i B(i) (i < H) and conj A(M-i) + i conj B(M-i) above; the
// inverse transform is a + i b; the rows are cropped into out.
template <typename T, int N1, int N2, int P>
__global__ __launch_bounds__(P * (N1 > N2 ? N1 : N2)) void rx_inv_reg(
    const Cx<T> *data, T *out, RealGeom g, const Cx<T> *twM, const int *skip)
{
#pragma clang fp contract(fast)
    constexpr int M = N1 * N2, TM = team<N1, N2>(), NT = P * TM, XB = xbuf<N1, N2>();
    __shared__ Cx<T> X[P * XB];
    __shared__ Cx<T> W[M];
    if (skip && *skip)
        return;
    const int t = int(threadIdx.x), m = t % TM, p = t / TM;
    const int rows = g.ny * g.nz;
    const int ra = int(blockIdx.x) * 2 * P + 2 * p;
    const int H = g.H;
    Cx<T> *Xp = X + p * XB;
    for (int e = t; e < M; e += NT)
        W[e] = twM[e];
    Cx<T> a[N1];
    if (m < N2) {
        const Cx<T> *sa = nullptr, *sb = nullptr;
        if (ra < rows) {
            const int j = ra % g.ny + g.o1, k = ra / g.ny + g.o2;
            sa = data + size_t(H) * (size_t(j) + size_t(g.M1) * k);
        }
        if (ra + 1 < rows) {
            const int j = (ra + 1) % g.ny + g.o1, k = (ra + 1) / g.ny + g.o2;
            sb = data + size_t(H) * (size_t(j) + size_t(g.M1) * k);
        }
        sfor<0, N1>([&](auto c) {
            constexpr int n1 = decltype(c)::value;
            const int i = N2 * n1 + m;
            const bool lo = i < H;
            const int kk = lo ? i : M - i;
            Cx<T> x{T(0), T(0)}, y{T(0), T(0)};
            if (sa) x = sa[kk];
            if (sb) y = sb[kk];
            // lo: x + i y; above H: conj(x) + i conj(y)
            a[n1] = lo ? Cx<T>{x.re - y.im, x.im + y.re} : Cx<T>{x.re + y.im, y.re - x.im};
        });
    }
    __syncthreads();
    if (m < N2) {
        dft<T, N1, 1>(a);
        sfor<0, N1>([&](auto c) {
            constexpr int k1 = decltype(c)::value;
            const Cx<T> w = W[m * k1];
            Xp[k1 * (N2 + 1) + m] = k1 == 0 ? a[0] : cmul(a[k1], Cx<T>{w.re, -w.im});
        });
    }
    __syncthreads();
    if (m < N1 && ra < rows) {
        Cx<T> b[N2];
        sfor<0, N2>([&](auto c) {
            constexpr int n2 = decltype(c)::value;
            b[n2] = Xp[m * (N2 + 1) + n2];
        });
        dft<T, N2, 1>(b);
        const size_t b0 = size_t(g.nx) * ra, b1 = b0 + g.nx;  // rows ra, ra + 1
        sfor<0, N2>([&](auto c) {
            constexpr int k2 = decltype(c)::value;
            const int x = m + N1 * k2;
            if (x >= g.o0 && x < g.o0 + g.nx) {
                out[b0 + (x - g.o0)] = b[k2].re;
                if (ra + 1 < rows)
                    out[b1 + (x - g.o0)] = b[k2].im;
            }
        });
    }
}

template <typename T, int N1, int N2>
hipError_t launch_x_one(bool forward, Cx<T> *data, const T *in, T *out, const RealGeom &g,
                        const Cx<T> *twM, const int *skip, hipStream_t st)
{
    constexpr int P = xpencils<T>(), NT = P * team<N1, N2>();
    const int rows = g.ny * g.nz;
    const int nblk = (rows + 2 * P - 1) / (2 * P);
    if (nblk > 0) {
        if (forward)
            rx_fwd_reg<T, N1, N2, P><<<nblk, NT, 0, st>>>(data, in, g, twM, skip);
        else
            rx_inv_reg<T, N1, N2, P><<<nblk, NT, 0, st>>>(data, out, g, twM, skip);
    }
    return hipGetLastError();
}

// The x pass of length M in the register form; false when M has no split.
template <typename T>
bool launch_x(int M, bool forward, Cx<T> *data, const T *in, T *out, const RealGeom &g,
              const Cx<T> *twM, const int *skip, hipStream_t st, hipError_t &err)
{
    switch (M) {
    case 176: err = launch_x_one<T, 16, 11>(forward, data, in, out, g, twM, skip, st); return true;
    case 192: err = launch_x_one<T, 16, 12>(forward, data, in, out, g, twM, skip, st); return true;
    case 256: err = launch_x_one<T, 16, 16>(forward, data, in, out, g, twM, skip, st); return true;
    case 320: err = launch_x_one<T, 20, 16>(forward, data, in, out, g, twM, skip, st); return true;
    case 360: err = launch_x_one<T, 20, 18>(forward, data, in, out, g, twM, skip, st); return true;
    case 384: err = launch_x_one<T, 24, 16>(forward, data, in, out, g, twM, skip, st); return true;
    case 448: err = launch_x_one<T, 28, 16>(forward, data, in, out, g, twM, skip, st); return true;
    case 512: err = launch_x_one<T, 32, 16>(forward, data, in, out, g, twM, skip, st); return true;
    case 640: err = launch_x_one<T, 32, 20>(forward, data, in, out, g, twM, skip, st); return true;
    default: return false;
    }
}

// The identity op of passes without one.
template <typename T>
struct NoOp {
    __device__ Cx<T> operator()(int, int, int, Cx<T> v) const { return v; }
};

}  // namespace reg
}  // namespace fft
}  // namespace qsm
