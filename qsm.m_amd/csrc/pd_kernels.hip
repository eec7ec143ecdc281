// pd_kernels.hip -- primal-dual TV/TGV kernels for gfx950.
//
// Built with -ffp-contract=off: every expression is evaluated exactly as
// written (no fused multiply-add), which together with the association order
// of pd_math.h reproduces the reference's CPU arithmetic bit for bit.
#include "pd_kernels.h"

namespace qsm {

// ---------------------------------------------------------------------------
// deterministic block reduction of four double sums
// ---------------------------------------------------------------------------

// Sums s[0..3] over the block in a fixed order (wavefront shuffles, then the
// waves in index order) and writes the result from thread 0.
template <int NT>
__device__ __forceinline__ void block_sum4(double s[4], double *dst)
{
    static_assert(NT % 64 == 0, "block must be whole wavefronts");
    __shared__ double red[NT / 64][4];
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 4; ++q)
            s[q] += __shfl_down(s[q], off, 64);
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if ((tid & 63) == 0)
        for (int q = 0; q < 4; ++q)
            red[tid >> 6][q] = s[q];
    __syncthreads();
    if (tid == 0) {
        double t[4] = {0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < NT / 64; ++w)
            for (int q = 0; q < 4; ++q)
                t[q] += red[w][q];
        for (int q = 0; q < 4; ++q)
            dst[q] = t[q];
    }
}

// ---------------------------------------------------------------------------
// two-sweep schedule (reference structure)
// ---------------------------------------------------------------------------

constexpr int SW_X = 64;
constexpr int SW_Y = 4;

template <typename T, bool TGV>
__global__ __launch_bounds__(SW_X *SW_Y) void dual_sweep_kernel(
    Bank<T> s, Bars<T> b, const T *__restrict__ f,
    const uint8_t *__restrict__ mk, uint8_t bit, Geom g, DualCoef<T> c,
    const int *stop)
{
    if (stop && *stop)
        return;
    const int i = blockIdx.x * SW_X + threadIdx.x + 1;
    const int j = blockIdx.y * SW_Y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.nx - 2 || j > g.ny - 2)
        return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const int l = i + sy * j + sk * k;
    if (!(mk[l] & bit))
        return;

    const Star<T> eb{b.eb[l], b.eb[l - 1], b.eb[l + 1], b.eb[l - sy],
                     b.eb[l + sy], b.eb[l - sk], b.eb[l + sk]};
    const Star<T> cb{b.cb[l], b.cb[l - 1], b.cb[l + 1], b.cb[l - sy],
                     b.cb[l + sy], b.cb[l - sk], b.cb[l + sk]};
    DualVox<T> d;
    d.nu = dual_nu(c, s.nu[l], eb, cb, f[l]);
    d.p1 = s.p1[l];
    d.p2 = s.p2[l];
    d.p3 = s.p3[l];
    if constexpr (TGV) {
        d.q1 = s.q1[l]; d.q2 = s.q2[l]; d.q3 = s.q3[l];
        d.q4 = s.q4[l]; d.q5 = s.q5[l]; d.q6 = s.q6[l];
        tgv_dual_pq(c, d, cb,
                    b.b1[l], b.b1[l + 1], b.b1[l + sy], b.b1[l + sk],
                    b.b2[l], b.b2[l + 1], b.b2[l + sy], b.b2[l + sk],
                    b.b3[l], b.b3[l + 1], b.b3[l + sy], b.b3[l + sk]);
        s.q1[l] = d.q1; s.q2[l] = d.q2; s.q3[l] = d.q3;
        s.q4[l] = d.q4; s.q5[l] = d.q5; s.q6[l] = d.q6;
    } else {
        tv_dual_p(c, d, cb);
    }
    s.nu[l] = d.nu;
    s.p1[l] = d.p1;
    s.p2[l] = d.p2;
    s.p3[l] = d.p3;
}

template <typename T, bool TGV, bool CHECK>
__global__ __launch_bounds__(SW_X *SW_Y) void primal_sweep_kernel(
    Bank<T> s, Bars<T> b, const uint8_t *__restrict__ mk, uint8_t bit,
    Geom g, PrimalCoef<T> c, const int *stop, double *partials)
{
    if (stop && *stop)
        return;
    const int i = blockIdx.x * SW_X + threadIdx.x + 1;
    const int j = blockIdx.y * SW_Y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    const int sy = g.nx, sk = g.nx * g.ny;
    const int l = i + sy * j + sk * k;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};

    bool act = (i <= g.nx - 2 && j <= g.ny - 2);
    if (act)
        act = (mk[l] & bit) != 0;
    if (act) {
        PrimalIn<T> in;
        in.nu = Star<T>{s.nu[l], s.nu[l - 1], s.nu[l + 1], s.nu[l - sy],
                        s.nu[l + sy], s.nu[l - sk], s.nu[l + sk]};
        in.p1c = s.p1[l]; in.p1xm = s.p1[l - 1];
        in.p2c = s.p2[l]; in.p2ym = s.p2[l - sy];
        in.p3c = s.p3[l]; in.p3zm = s.p3[l - sk];
        if constexpr (TGV) {
            in.q1c = s.q1[l]; in.q1xm = s.q1[l - 1];
            in.q2c = s.q2[l]; in.q2xm = s.q2[l - 1]; in.q2ym = s.q2[l - sy];
            in.q3c = s.q3[l]; in.q3xm = s.q3[l - 1]; in.q3zm = s.q3[l - sk];
            in.q4c = s.q4[l]; in.q4ym = s.q4[l - sy];
            in.q5c = s.q5[l]; in.q5ym = s.q5[l - sy]; in.q5zm = s.q5[l - sk];
            in.q6c = s.q6[l]; in.q6zm = s.q6[l - sk];
        }
        const T en = s.eta[l], cn = s.chi[l];
        T e, x;
        primal_eta_chi(c, in, en, cn, e, x);
        s.eta[l] = e;
        s.chi[l] = x;
        b.eb[l] = extrap(e, en);
        b.cb[l] = extrap(x, cn);
        if constexpr (TGV) {
            const T o1 = s.w1[l], o2 = s.w2[l], o3 = s.w3[l];
            T n1 = o1, n2 = o2, n3 = o3;
            primal_w(c, in, n1, n2, n3);
            s.w1[l] = n1; s.w2[l] = n2; s.w3[l] = n3;
            b.b1[l] = extrap(n1, o1);
            b.b2[l] = extrap(n2, o2);
            b.b3[l] = extrap(n3, o3);
        }
        if constexpr (CHECK)
            conv_terms(x, cn, e, en, acc);
    }
    if constexpr (CHECK) {
        const int blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        block_sum4<SW_X * SW_Y>(acc, partials + 4 * blk);
    }
}

static dim3 sweep_grid(Geom g)
{
    return dim3((g.nx - 2 + SW_X - 1) / SW_X, (g.ny - 2 + SW_Y - 1) / SW_Y,
                g.nz - 2);
}

int primal_sweep_blocks(Geom g)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3)
        return 0;
    const dim3 gr = sweep_grid(g);
    return int(gr.x * gr.y * gr.z);
}

template <typename T, bool TGV>
hipError_t launch_dual_sweep(const Bank<T> &s, const Bars<T> &b, const T *f,
                             const uint8_t *mk, uint8_t bit, Geom g,
                             const DualCoef<T> &c, const int *stop,
                             hipStream_t st)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3)
        return hipSuccess;
    dual_sweep_kernel<T, TGV><<<sweep_grid(g), dim3(SW_X, SW_Y), 0, st>>>(
        s, b, f, mk, bit, g, c, stop);
    return hipGetLastError();
}

template <typename T, bool TGV>
hipError_t launch_primal_sweep(const Bank<T> &s, const Bars<T> &b,
                               const uint8_t *mk, uint8_t bit, Geom g,
                               const PrimalCoef<T> &c, const int *stop,
                               double *partials, hipStream_t st)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3)
        return hipSuccess;
    if (partials)
        primal_sweep_kernel<T, TGV, true><<<sweep_grid(g), dim3(SW_X, SW_Y), 0, st>>>(
            s, b, mk, bit, g, c, stop, partials);
    else
        primal_sweep_kernel<T, TGV, false><<<sweep_grid(g), dim3(SW_X, SW_Y), 0, st>>>(
            s, b, mk, bit, g, c, stop, partials);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// convergence, masks, scaling
// ---------------------------------------------------------------------------

constexpr int FIN_T = 256;

__global__ __launch_bounds__(FIN_T) void conv_finalize_kernel(
    const double *partials, int nblk, Ctl *ctl, double *hist, int hist_cap,
    double tol, uint32_t iter, uint32_t bank)
{
    if (ctl->stop)
        return;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += FIN_T)
        for (int q = 0; q < 4; ++q)
            s[q] += partials[4 * b + q];
    __shared__ double tot[4];
    block_sum4<FIN_T>(s, tot);
    __syncthreads();
    if (threadIdx.x == 0) {
        // pd_tgv_mex.c:1004-1010
        const double n1 = sqrt(tot[2] / tot[0]);
        const double n2 = sqrt(tot[3] / tot[1]);
        const uint32_t c = ctl->checks;
        if (hist && int(c) < hist_cap) {
            hist[2 * c] = n1;
            hist[2 * c + 1] = n2;
        }
        ctl->checks = c + 1;
        ctl->nr[0] = n1;
        ctl->nr[1] = n2;
        if ((n1 < tol) && (n2 < 2.0 * tol)) {
            ctl->stop = 1;
            ctl->stop_iter = iter;
            ctl->stop_bank = bank;
        }
    }
}

hipError_t launch_conv_finalize(const double *partials, int nblk, Ctl *ctl,
                                double *hist, int hist_cap, double tol,
                                uint32_t iter, uint32_t bank, hipStream_t st)
{
    conv_finalize_kernel<<<1, FIN_T, 0, st>>>(partials, nblk, ctl, hist,
                                              hist_cap, tol, iter, bank);
    return hipGetLastError();
}

constexpr int RED_T = 256;

template <typename T>
__global__ __launch_bounds__(RED_T) void conv_sums_kernel(
    const T *chi, const T *chi0, const T *eta, const T *eta0, size_t N,
    double *partials)
{
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const size_t stride = size_t(gridDim.x) * RED_T;
    for (size_t l = size_t(blockIdx.x) * RED_T + threadIdx.x; l < N; l += stride)
        conv_terms(chi[l], chi0[l], eta[l], eta0[l], s);
    block_sum4<RED_T>(s, partials + 4 * blockIdx.x);
}

template <typename T>
hipError_t launch_conv_sums(const T *chi, const T *chi0, const T *eta,
                            const T *eta0, size_t N, double *partials,
                            int nblk, hipStream_t st)
{
    conv_sums_kernel<T><<<nblk, RED_T, 0, st>>>(chi, chi0, eta, eta0, N,
                                                partials);
    return hipGetLastError();
}

__global__ void mask_init_kernel(uint8_t *out, const uint8_t *m, Geom g,
                                 int raw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    if (i >= g.nx)
        return;
    const int sy = g.nx, sk = g.nx * g.ny;
    const int l = i + sy * j + sk * k;
    // pd_tgv_mex.c:1075-1115: the "both neighbours present" code of all
    // three axes, i.e. all six face neighbours exist and are in m.
    const bool ix = (i > 0 && m[l - 1]) && (i + 1 < g.nx && m[l + 1]);
    const bool iy = (j > 0 && m[l - sy]) && (j + 1 < g.ny && m[l + sy]);
    const bool iz = (k > 0 && m[l - sk]) && (k + 1 < g.nz && m[l + sk]);
    const bool mi = ix && iy && iz;
    if (raw) {
        out[l] = mi ? 1 : 0;
        return;
    }
    const bool interior = i >= 1 && i + 1 < g.nx && j >= 1 && j + 1 < g.ny
                          && k >= 1 && k + 1 < g.nz;
    out[l] = uint8_t((interior && m[l] != 0 ? 1 : 0) | (interior && mi ? 2 : 0));
}

hipError_t launch_mask_init(uint8_t *out, const uint8_t *m, Geom g, int raw,
                            hipStream_t st)
{
    const dim3 grid((g.nx + 255) / 256, g.ny, g.nz);
    mask_init_kernel<<<grid, 256, 0, st>>>(out, m, g, raw);
    return hipGetLastError();
}

__global__ void ctl_reset_kernel(Ctl *ctl, int n)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n)
        return;
    ctl[v].stop = 0;
    ctl[v].stop_iter = 0;
    ctl[v].checks = 0;
    ctl[v].stop_bank = 0;
    ctl[v].nr[0] = 1.0;
    ctl[v].nr[1] = 1.0;
}

hipError_t launch_ctl_reset(Ctl *ctl, int n, hipStream_t st)
{
    ctl_reset_kernel<<<(n + 63) / 64, 64, 0, st>>>(ctl, n);
    return hipGetLastError();
}

template <typename T>
__global__ void select_copy_kernel(T *dst, const T *chi0, const T *chi1,
                                   const Ctl *ctl, uint32_t cur_bank, size_t N)
{
    const uint32_t b = ctl->stop ? ctl->stop_bank : cur_bank;
    const T *src = (b & 1u) ? chi1 : chi0;
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t l = size_t(blockIdx.x) * blockDim.x + threadIdx.x; l < N; l += stride)
        dst[l] = src[l];
}

template <typename T>
hipError_t launch_select_copy(T *dst, const T *chi0, const T *chi1,
                              const Ctl *ctl, uint32_t cur_bank, size_t N,
                              hipStream_t st)
{
    select_copy_kernel<T><<<1024, 256, 0, st>>>(dst, chi0, chi1, ctl, cur_bank, N);
    return hipGetLastError();
}

// Clears n16 16-byte words and then `tail` trailing bytes (a fill kernel,
// so that code a caller captures into a graph records no memset node).
__global__ void zero_bytes_kernel(uint4 *p, size_t n16, uint8_t *tail, int ntail)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (size_t l = g; l < n16; l += stride)
        p[l] = make_uint4(0u, 0u, 0u, 0u);
    if (g < size_t(ntail))
        tail[g] = 0;
}

hipError_t launch_zero_bytes(void *p, size_t bytes, hipStream_t st)
{
    if (bytes == 0)
        return hipSuccess;
    const size_t n16 = bytes / 16;
    uint8_t *tail = static_cast<uint8_t *>(p) + n16 * 16;
    zero_bytes_kernel<<<2048, 256, 0, st>>>(static_cast<uint4 *>(p), n16, tail,
                                            int(bytes - n16 * 16));
    return hipGetLastError();
}

template <typename T>
__global__ void scale_kernel(T *f, const T *src, T sg, size_t N)
{
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t l = size_t(blockIdx.x) * blockDim.x + threadIdx.x; l < N; l += stride)
        f[l] = sg * src[l];
}

template <typename T>
hipError_t launch_scale(T *f, const T *src, T sg, size_t N, hipStream_t st)
{
    scale_kernel<T><<<2048, 256, 0, st>>>(f, src, sg, N);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// explicit instantiations
// ---------------------------------------------------------------------------
#define QSM_INST(T)                                                           \
    template hipError_t launch_dual_sweep<T, true>(                           \
        const Bank<T> &, const Bars<T> &, const T *, const uint8_t *, uint8_t, \
        Geom, const DualCoef<T> &, const int *, hipStream_t);                 \
    template hipError_t launch_dual_sweep<T, false>(                          \
        const Bank<T> &, const Bars<T> &, const T *, const uint8_t *, uint8_t, \
        Geom, const DualCoef<T> &, const int *, hipStream_t);                 \
    template hipError_t launch_primal_sweep<T, true>(                         \
        const Bank<T> &, const Bars<T> &, const uint8_t *, uint8_t, Geom,     \
        const PrimalCoef<T> &, const int *, double *, hipStream_t);           \
    template hipError_t launch_primal_sweep<T, false>(                        \
        const Bank<T> &, const Bars<T> &, const uint8_t *, uint8_t, Geom,     \
        const PrimalCoef<T> &, const int *, double *, hipStream_t);           \
    template hipError_t launch_conv_sums<T>(const T *, const T *, const T *,  \
                                            const T *, size_t, double *, int, \
                                            hipStream_t);                     \
    template hipError_t launch_scale<T>(T *, const T *, T, size_t, hipStream_t); \
    template hipError_t launch_select_copy<T>(T *, const T *, const T *,      \
                                              const Ctl *, uint32_t, size_t,  \
                                              hipStream_t);

QSM_INST(double)
QSM_INST(float)
#undef QSM_INST

}  // namespace qsm
