// bgremove.hip -- background field removal by the spherical mean value
// property, SHARP (src/bgremove/sharp.m) and V-SHARP (src/bgremove/vsharp.m),
// on gfx950.
//
// Every step is a real convolution with a real, even multiplier built from
// the spectrum H of the normalised ball (smvKernel.m): the mask smoothed by
// H and thresholded (the eroded mask), the field high-passed by 1 - H, and
// the deconvolution by 1 - H truncated where |1 - H| < thr.  They run on the
// mask's bounding box (cropIndices.m) with the zero padding of padfastfft.m
// implicit (fft_dev.h real_conv: only data rows and planes are read and
// stored): the eroded mask is zero outside the box, so nothing outside it
// is needed.  The mask is smoothed in single precision whatever the class
// of f, as the reference's fft3(single(mask)) does.  V-SHARP convolves one
// radius at a time (the reference packs two radii into one complex kernel;
// for a real field that is the same pair of real convolutions).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fft_dev.h"

using namespace qsm;
using namespace qsm::fft;

namespace {

constexpr int NT = 256;

// Number of samples of the smvKernel.m ball: grid points (i h0, j h1, k h2)
// with (Y*Y + X*X) + Z*Z <= r*r (the reference's order of the sum).
long ball_count(const double h[3], double r)
{
    const long R0 = long(std::ceil(r / h[0])), R1 = long(std::ceil(r / h[1])),
               R2 = long(std::ceil(r / h[2]));
    long K = 0;
    for (long i = -R0; i <= R0; ++i)
        for (long j = -R1; j <= R1; ++j)
            for (long k = -R2; k <= R2; ++k) {
                const double X = double(i) * h[0], Y = double(j) * h[1], Z = double(k) * h[2];
                if ((Y * Y + X * X) + Z * Z <= r * r) ++K;
            }
    return K;
}

// ifftshift(smvKernel(M, r, vsz)): the ball of value v = 1/K centred on
// sample 0 with periodic wrap (index m of an axis of length n is at offset
// m for m < n - n/2, else m - n), over the whole M0 x M1 x M2 grid.
template <typename T>
__global__ __launch_bounds__(NT) void ball_kernel(T *h, int M0, int M1, int M2, double h0,
                                                  double h1, double h2, double r2, T v)
{
    const size_t n = size_t(M0) * M1 * M2;
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l >= n) return;
    const int i = int(l % M0), j = int((l / M0) % M1), k = int(l / (size_t(M0) * M1));
    const double X = double(i < M0 - M0 / 2 ? i : i - M0) * h0;
    const double Y = double(j < M1 - M1 / 2 ? j : j - M1) * h1;
    const double Z = double(k < M2 - M2 / 2 ? k : k - M2) * h2;
    h[l] = ((Y * Y + X * X) + Z * Z <= r2) ? v : T(0);
}

// The SMV multipliers at half-spectrum sample (x, y, z), from Sb = 2
// FFT3(ifftshift(ball)) as the x pass leaves it (H = Sb.re / 2 in the class
// Tb of f: real(fft3(ifftshift(h))), sharp.m:62):
//   MODE 0  H                (the mask smoothing, sharp.m:63)
//   MODE 1  1 - H            (the high-pass, sharp.m:65-75)
//   MODE 2  1 / (1 - H), 0 where |1 - H| < thr (the deconvolution,
//           sharp.m:76-78; vsharp.m's ih)
// times scale (1/(M0 M1 M2) of ifft3, 1/2 for the x pass's doubling).
template <typename Tb, typename T, int MODE>
struct SmvOp {
    const Cx<Tb> *Sb;
    int H, M1;
    double scale;
    Tb thr;
    __device__ Cx<T> operator()(int x, int y, int z, Cx<T> v) const
    {
        const Tb hb = Tb(0.5) * Sb[size_t(x) + size_t(H) * (size_t(y) + size_t(M1) * z)].re;
        double c;
        if constexpr (MODE == 0) {
            c = double(hb) * scale;
        } else if constexpr (MODE == 1) {
            c = double(Tb(1) - hb) * scale;
        } else {
            const Tb hm = Tb(1) - hb;
            c = (fabs(hm) < thr) ? 0.0 : scale / double(hm);
        }
        const T cc = T(c);
        return Cx<T>{v.re * cc, v.im * cc};
    }
};

template <typename T>
struct KeepOp {
    __device__ Cx<T> operator()(int, int, int, Cx<T> v) const { return v; }
};

__global__ __launch_bounds__(NT) void to_float_kernel(float *mf, const uint8_t *m, size_t n)
{
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l < n) mf[l] = m[l] ? 1.0f : 0.0f;
}

// sharp.m:63 / vsharp.m:103-104: the smoothed mask > 0.999 (single).
__global__ __launch_bounds__(NT) void threshold_kernel(uint8_t *m, const float *s, size_t n)
{
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l < n) m[l] = s[l] > 0.999f ? 1 : 0;
}

// maskn .* g
template <typename T>
__global__ __launch_bounds__(NT) void mask_kernel(T *g, const uint8_t *m, size_t n)
{
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l < n) g[l] = T(m[l] ? 1 : 0) * g[l];
}

// vsharp.m:118-119: f = f + (mask_r - mask_prev) .* real(ifft3(h_r .* F))
template <typename T>
__global__ __launch_bounds__(NT) void accumulate_kernel(T *acc, const T *c, const uint8_t *mr,
                                                        const uint8_t *mp, size_t n)
{
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l < n) acc[l] = acc[l] + (T(mr[l] ? 1 : 0) - T(mp[l] ? 1 : 0)) * c[l];
}

template <typename T>
__global__ __launch_bounds__(NT) void zero_kernel(T *a, size_t n)
{
    const size_t l = size_t(blockIdx.x) * NT + threadIdx.x;
    if (l < n) a[l] = T(0);
}

unsigned grid1(size_t n) { return unsigned((n + NT - 1) / NT); }

size_t al(size_t n) { return (n + 255) / 256 * 256; }

// Work space of one call: the ball image on the transform grid, two ball
// spectra (V-SHARP keeps the first radius's for its deconvolution), the
// convolution's work spectrum, the spectra of the mask and of a single
// echo's f that V-SHARP reuses for every radius, two float volumes and two
// of the class of f on the crop, three byte masks.  The two reused spectra
// are laid out only for the calls that use them (SPEC_* bits): SHARP needs
// neither, multi-echo V-SHARP only the mask's.
template <typename T>
struct Work {
    T *ball;
    Cx<T> *Sb, *Sb1, *W, *Sf;
    Cx<float> *Sm;
    float *mf, *sf;
    T *g, *c;
    uint8_t *m1, *mp, *mo;
};

enum { SPEC_F = 1, SPEC_M = 2 };

// Spectra a call keeps: V-SHARP the mask's, and f's for a single echo.
inline int vsharp_spectra(size_t nech) { return SPEC_M | (nech == 1 ? SPEC_F : 0); }

template <typename T>
size_t layout(const size_t szc[3], const size_t M[3], int spectra, Work<T> *w, char *base)
{
    const size_t n = szc[0] * szc[1] * szc[2], Mt = M[0] * M[1] * M[2];
    const size_t spec = al(real_conv_work_bytes(M, sizeof(T)));
    size_t off = 0;
    auto take = [&](size_t b) { char *p = base ? base + off : nullptr; off += al(b); return p; };
    char *pb = take(Mt * sizeof(T)), *ps = take(spec), *ps1 = take(spec), *pw = take(spec);
    char *psf = (spectra & SPEC_F) ? take(spec) : nullptr;
    char *psm = (spectra & SPEC_M) ? take(real_conv_work_bytes(M, sizeof(float))) : nullptr;
    char *pmf = take(n * sizeof(float)), *pvf = take(n * sizeof(float));
    char *pg = take(n * sizeof(T)), *pc = take(n * sizeof(T));
    char *pm1 = take(n), *pmp = take(n), *pmo = take(n);
    if (w) {
        w->ball = reinterpret_cast<T *>(pb);
        w->Sb = reinterpret_cast<Cx<T> *>(ps);
        w->Sb1 = reinterpret_cast<Cx<T> *>(ps1);
        w->W = reinterpret_cast<Cx<T> *>(pw);
        w->Sf = reinterpret_cast<Cx<T> *>(psf);
        w->Sm = reinterpret_cast<Cx<float> *>(psm);
        w->mf = reinterpret_cast<float *>(pmf);
        w->sf = reinterpret_cast<float *>(pvf);
        w->g = reinterpret_cast<T *>(pg);
        w->c = reinterpret_cast<T *>(pc);
        w->m1 = reinterpret_cast<uint8_t *>(pm1);
        w->mp = reinterpret_cast<uint8_t *>(pmp);
        w->mo = reinterpret_cast<uint8_t *>(pmo);
    }
    return off;
}

void plan_for(const size_t szc[3], const double vsz[3], double rmax, size_t M[3], size_t off[3])
{
    const double hmin = std::min(vsz[0], std::min(vsz[1], vsz[2]));
    const size_t p = size_t(std::ceil(rmax / hmin));  // sharp.m:53, vsharp.m:56
    const size_t pad[3] = {p, p, p};
    pad_plan(szc, pad, M, off);
}

// Sb = 2 FFT3(ifftshift(ball of radius r)) on the transform grid.
template <typename T>
int ball_spectrum(Cx<T> *Sb, T *ball, const size_t M[3], const double vsz[3], double r,
                  hipStream_t st)
{
    const long K = ball_count(vsz, r);
    const size_t Mt = M[0] * M[1] * M[2];
    ball_kernel<T><<<grid1(Mt), NT, 0, st>>>(ball, int(M[0]), int(M[1]), int(M[2]), vsz[0],
                                             vsz[1], vsz[2], r * r, T(1.0 / double(K)));
    QSM_HIP(hipGetLastError());
    const size_t zero[3] = {0, 0, 0};
    return real_conv<T, false>(ball, static_cast<T *>(nullptr), M, M, zero, Sb, KeepOp<T>{}, st);
}

// The eroded mask of one radius: real(ifft3(H .* fft3(single(mask)))) > 0.999;
// from_spectrum: fft3(single(mask)) is already in w.Sm.
template <typename T>
int erode(uint8_t *mout, const Cx<T> *Sb, const size_t szc[3], const size_t M[3],
          const size_t off[3], Work<T> &w, hipStream_t st, bool from_spectrum = false)
{
    const size_t n = szc[0] * szc[1] * szc[2];
    SmvOp<T, float, 0> op{Sb, int(M[0] / 2 + 1), int(M[1]),
                          0.5 / (double(M[0]) * double(M[1]) * double(M[2])), T(0)};
    Cx<float> *W = reinterpret_cast<Cx<float> *>(w.W);
    int rc = from_spectrum ? real_conv_from<float>(w.Sm, w.sf, szc, M, off, W, op, st)
                           : real_conv<float, true>(w.mf, w.sf, szc, M, off, W, op, st);
    if (rc) return rc;
    threshold_kernel<<<grid1(n), NT, 0, st>>>(mout, w.sf, n);
    QSM_HIP(hipGetLastError());
    return QSM_OK;
}

// sharp.m:47-79 on the crop (echoes contiguous): flc per echo, mout the
// eroded mask.
template <typename T>
int sharp_t(const T *fc, const uint8_t *mc, const size_t szc[3], size_t nech,
            const double vsz[3], double r, double thr, T *flc, uint8_t *mout, void *work,
            hipStream_t st)
{
    int rc;
    size_t M[3], off[3];
    plan_for(szc, vsz, r, M, off);
    Work<T> w;
    layout<T>(szc, M, 0, &w, static_cast<char *>(work));
    const size_t n = szc[0] * szc[1] * szc[2];
    const double scale = 0.5 / (double(M[0]) * double(M[1]) * double(M[2]));
    const int H = int(M[0] / 2 + 1), M1 = int(M[1]);
    if ((rc = ball_spectrum<T>(w.Sb, w.ball, M, vsz, r, st))) return rc;
    to_float_kernel<<<grid1(n), NT, 0, st>>>(w.mf, mc, n);
    QSM_HIP(hipGetLastError());
    if ((rc = erode<T>(w.m1, w.Sb, szc, M, off, w, st))) return rc;
    const SmvOp<T, T, 1> hp{w.Sb, H, M1, scale, T(thr)};
    const SmvOp<T, T, 2> dc{w.Sb, H, M1, scale, T(thr)};
    for (size_t e = 0; e < nech; ++e) {
        const T *f = fc + e * n;
        T *o = flc + e * n;
        // f = maskn .* real(ifft3(h .* fft3(f))), h = 1 - H (sharp.m:65-75)
        if ((rc = real_conv<T, true>(f, w.g, szc, M, off, w.W, hp, st))) return rc;
        mask_kernel<T><<<grid1(n), NT, 0, st>>>(w.g, w.m1, n);
        QSM_HIP(hipGetLastError());
        // f = maskn .* real(ifft3(fft3(f) ./ h, 0 where |h| < thr)) (:76-79)
        if ((rc = real_conv<T, true>(w.g, o, szc, M, off, w.W, dc, st))) return rc;
        mask_kernel<T><<<grid1(n), NT, 0, st>>>(o, w.m1, n);
        QSM_HIP(hipGetLastError());
    }
    QSM_HIP(hipMemcpyAsync(mout, w.m1, n, hipMemcpyDeviceToDevice, st));
    return QSM_OK;
}

// vsharp.m:47-146 on the crop; rs sorted descending, at least two radii.
template <typename T>
int vsharp_t(const T *fc, const uint8_t *mc, const size_t szc[3], size_t nech,
             const double vsz[3], const std::vector<double> &rs, double thr, T *flc,
             uint8_t *mout, void *work, hipStream_t st)
{
    int rc;
    size_t M[3], off[3];
    plan_for(szc, vsz, rs.front(), M, off);
    Work<T> w;
    layout<T>(szc, M, vsharp_spectra(nech), &w, static_cast<char *>(work));
    const size_t n = szc[0] * szc[1] * szc[2];
    const double scale = 0.5 / (double(M[0]) * double(M[1]) * double(M[2]));
    const int H = int(M[0] / 2 + 1), M1 = int(M[1]);
    const size_t nrs = rs.size();
    // the reference returns its last mask1: the first radius of the last
    // pair, or the last radius when the count is odd (vsharp.m:87-140)
    const size_t qout = (nrs % 2 == 0) ? nrs - 2 : nrs - 1;
    to_float_kernel<<<grid1(n), NT, 0, st>>>(w.mf, mc, n);
    QSM_HIP(hipGetLastError());
    zero_kernel<uint8_t><<<grid1(n), NT, 0, st>>>(w.mp, n);
    zero_kernel<T><<<grid1(n * nech), NT, 0, st>>>(flc, n * nech);
    QSM_HIP(hipGetLastError());
    // The forward transforms of the mask and (one echo) of f do not depend
    // on the radius: computed once, each radius runs only op + inverse.
    if ((rc = real_conv<float, false>(w.mf, static_cast<float *>(nullptr), szc, M, off, w.Sm,
                                      KeepOp<float>{}, st)))
        return rc;
    const bool one = nech == 1;
    if (one && (rc = real_conv<T, false>(fc, static_cast<T *>(nullptr), szc, M, off, w.Sf,
                                         KeepOp<T>{}, st)))
        return rc;
    for (size_t q = 0; q < nrs; ++q) {
        Cx<T> *Sb = q == 0 ? w.Sb1 : w.Sb;  // the first radius's spectrum is kept for ih
        if ((rc = ball_spectrum<T>(Sb, w.ball, M, vsz, rs[q], st))) return rc;
        if ((rc = erode<T>(w.m1, Sb, szc, M, off, w, st, true))) return rc;
        const SmvOp<T, T, 1> hp{Sb, H, M1, scale, T(thr)};
        for (size_t e = 0; e < nech; ++e) {
            rc = one ? real_conv_from<T>(w.Sf, w.c, szc, M, off, w.W, hp, st)
                     : real_conv<T, true>(fc + e * n, w.c, szc, M, off, w.W, hp, st);
            if (rc) return rc;
            accumulate_kernel<T><<<grid1(n), NT, 0, st>>>(flc + e * n, w.c, w.m1, w.mp, n);
            QSM_HIP(hipGetLastError());
        }
        if (q == qout)
            QSM_HIP(hipMemcpyAsync(w.mo, w.m1, n, hipMemcpyDeviceToDevice, st));
        QSM_HIP(hipMemcpyAsync(w.mp, w.m1, n, hipMemcpyDeviceToDevice, st));
    }
    // f = maskn .* real(ifft3(ih .* fft3(f))), ih from the first radius,
    // maskn the last radius's mask (vsharp.m:142-143)
    const SmvOp<T, T, 2> dc{w.Sb1, H, M1, scale, T(thr)};
    for (size_t e = 0; e < nech; ++e) {
        T *o = flc + e * n;
        if ((rc = real_conv<T, true>(o, o, szc, M, off, w.W, dc, st))) return rc;
        mask_kernel<T><<<grid1(n), NT, 0, st>>>(o, w.mp, n);
        QSM_HIP(hipGetLastError());
    }
    QSM_HIP(hipMemcpyAsync(mout, w.mo, n, hipMemcpyDeviceToDevice, st));
    return QSM_OK;
}

int check_common(int dtype, const size_t sz[3], size_t nech, const double vsz[3], double thr)
{
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    QSM_ARG(sz && vsz, "null pointer");
    int rc = check_grid(sz, vsz, 1);
    if (rc) return rc;
    QSM_ARG(nech >= 1, "at least one echo");
    QSM_ARG(std::isfinite(thr), "thr must be finite");
    return QSM_OK;
}

int sorted_radii(const double *rs, size_t nrs, std::vector<double> &out)
{
    QSM_ARG(rs && nrs >= 2, "vsharp needs at least two radii (the reference sets its inverse "
                            "kernel from the first pair)");
    out.assign(rs, rs + nrs);
    for (double r : out) QSM_ARG(std::isfinite(r) && r > 0.0, "radii must be finite and > 0");
    std::sort(out.begin(), out.end(), [](double a, double b) { return a > b; });
    return QSM_OK;
}

// Host forms: crop to the mask's non-empty slices (sharp.m:51-56), run on
// the device, uncrop into zeros (sharp.m:81-84).  work_bytes(szc) sizes the
// device work space; run(fc, mc, szc, flc, mo, work) is the device form.
// The crop and the uncrop run on the device: the whole field and mask go
// up, the whole results come down into the caller's arrays.
template <typename WorkBytes, typename Run>
int host_form(int dtype, const void *f, const uint8_t *mask, const size_t sz[3], size_t nech,
              void *fl, uint8_t *mask_out, WorkBytes work_bytes, Run run)
{
    const size_t es = dtype_size(dtype), N = sz[0] * sz[1] * sz[2];
    std::vector<char> any[3];
    axis_any(mask, sz, any);
    std::vector<int> ix;  // the x, then y, then z indices of non-empty slices
    size_t szc[3];
    for (int a = 0; a < 3; ++a) {
        const size_t before = ix.size();
        for (size_t i = 0; i < sz[a]; ++i)
            if (any[a][i]) ix.push_back(int(i));
        szc[a] = ix.size() - before;
    }
    if (szc[0] == 0) {  // empty mask: nothing to remove
        std::memset(fl, 0, N * nech * es);
        if (mask_out) std::memset(mask_out, 0, N);
        return QSM_OK;
    }
    const size_t n = szc[0] * szc[1] * szc[2];
    int rc;
    if ((rc = ensure_device())) return rc;
    // dF, dM: the whole field and mask, then the whole results
    DevBuf dF, dM, dix, df, dm, dfl, dmo, dw;
    if ((rc = dF.alloc(N * nech * es)) || (rc = dM.alloc(N)) ||
        (rc = dix.alloc(ix.size() * sizeof(int))) || (rc = df.alloc(n * nech * es)) ||
        (rc = dm.alloc(n)) || (rc = dfl.alloc(n * nech * es)) || (rc = dmo.alloc(n)) ||
        (rc = dw.alloc(work_bytes(szc))))
        return rc;
    QSM_HIP(hipMemcpy(dF.p, f, N * nech * es, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(dM.p, mask, N, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(dix.p, ix.data(), ix.size() * sizeof(int), hipMemcpyHostToDevice));
    const int *dixp = static_cast<const int *>(dix.p);
    const CropMap cm{dixp, dixp + szc[0], dixp + szc[0] + szc[1], int(szc[0]), int(szc[1]),
                     int(szc[2]), sz[0], sz[1], N, n};
    const int ne = int(nech);
    uint8_t *Mb = static_cast<uint8_t *>(dM.p);
    crop_gather_kernel<uint8_t, true><<<crop_blocks(cm, 1), 256>>>(
        static_cast<uint8_t *>(dm.p), Mb, cm, 1);
    if (dtype == QSM_F32)
        crop_gather_kernel<float><<<crop_blocks(cm, ne), 256>>>(
            static_cast<float *>(df.p), static_cast<const float *>(dF.p), cm, ne);
    else
        crop_gather_kernel<double><<<crop_blocks(cm, ne), 256>>>(
            static_cast<double *>(df.p), static_cast<const double *>(dF.p), cm, ne);
    QSM_HIP(hipGetLastError());
    if ((rc = run(df.p, static_cast<const uint8_t *>(dm.p), szc, dfl.p,
                  static_cast<uint8_t *>(dmo.p), dw.p)))
        return rc;
    // uncrop into zeros
    QSM_HIP(hipMemsetAsync(dF.p, 0, N * nech * es, nullptr));
    if (dtype == QSM_F32)
        crop_scatter_kernel<float><<<crop_blocks(cm, ne), 256>>>(
            static_cast<float *>(dF.p), static_cast<const float *>(dfl.p), cm, ne);
    else
        crop_scatter_kernel<double><<<crop_blocks(cm, ne), 256>>>(
            static_cast<double *>(dF.p), static_cast<const double *>(dfl.p), cm, ne);
    if (mask_out) {
        QSM_HIP(hipMemsetAsync(Mb, 0, N, nullptr));
        crop_scatter_kernel<uint8_t><<<crop_blocks(cm, 1), 256>>>(
            Mb, static_cast<const uint8_t *>(dmo.p), cm, 1);
    }
    QSM_HIP(hipGetLastError());
    QSM_HIP(hipMemcpy(fl, dF.p, N * nech * es, hipMemcpyDeviceToHost));
    if (mask_out) QSM_HIP(hipMemcpy(mask_out, Mb, N, hipMemcpyDeviceToHost));
    return QSM_OK;
}

// Sizes the work space only for arguments a call accepts: a known class,
// non-empty crop, finite positive voxel size and radius (the padding is
// ceil(rmax / min(vsz)) voxels, sharp.m:53).
bool work_args_ok(int dtype, const size_t szc[3], const double vsz[3], double rmax)
{
    if (!szc || !vsz || (dtype != QSM_F32 && dtype != QSM_F64)) return false;
    for (int a = 0; a < 3; ++a)
        if (szc[a] == 0 || !std::isfinite(vsz[a]) || !(vsz[a] > 0.0)) return false;
    const double p = std::ceil(rmax / std::min(vsz[0], std::min(vsz[1], vsz[2])));
    return std::isfinite(rmax) && rmax > 0.0 && p < 1e6;
}

}  // namespace

extern "C" {

size_t qsm_sharp_work_bytes(int dtype, const size_t szc[3], const double vsz[3], double rmax)
{
    if (!work_args_ok(dtype, szc, vsz, rmax)) return 0;
    size_t M[3], off[3];
    plan_for(szc, vsz, rmax, M, off);
    return dtype == QSM_F32 ? layout<float>(szc, M, 0, nullptr, nullptr)
                            : layout<double>(szc, M, 0, nullptr, nullptr);
}

size_t qsm_vsharp_work_bytes(int dtype, const size_t szc[3], const double vsz[3], double rmax,
                             size_t nechoes)
{
    if (!work_args_ok(dtype, szc, vsz, rmax) || nechoes < 1) return 0;
    size_t M[3], off[3];
    plan_for(szc, vsz, rmax, M, off);
    const int sp = vsharp_spectra(nechoes);
    return dtype == QSM_F32 ? layout<float>(szc, M, sp, nullptr, nullptr)
                            : layout<double>(szc, M, sp, nullptr, nullptr);
}

int qsm_sharp_dev(int dtype, const void *fc, const uint8_t *mc, const size_t szc[3],
                  size_t nechoes, const double vsz[3], double r, double thr, void *flc,
                  uint8_t *mask_out, void *work, void *stream)
{
    int rc = check_common(dtype, szc, nechoes, vsz, thr);
    if (rc) return rc;
    QSM_ARG(fc && mc && flc && mask_out && work, "null pointer");
    QSM_ARG(std::isfinite(r) && r > 0.0, "radius must be finite and > 0");
    const hipStream_t st = as_stream(stream);
    return dtype == QSM_F64
        ? sharp_t<double>(static_cast<const double *>(fc), mc, szc, nechoes, vsz, r, thr,
                          static_cast<double *>(flc), mask_out, work, st)
        : sharp_t<float>(static_cast<const float *>(fc), mc, szc, nechoes, vsz, r, thr,
                         static_cast<float *>(flc), mask_out, work, st);
}

int qsm_vsharp_dev(int dtype, const void *fc, const uint8_t *mc, const size_t szc[3],
                   size_t nechoes, const double vsz[3], const double *rs, size_t nrs,
                   double thr, void *flc, uint8_t *mask_out, void *work, void *stream)
{
    int rc = check_common(dtype, szc, nechoes, vsz, thr);
    if (rc) return rc;
    QSM_ARG(fc && mc && flc && mask_out && work, "null pointer");
    std::vector<double> r;
    if ((rc = sorted_radii(rs, nrs, r))) return rc;
    const hipStream_t st = as_stream(stream);
    return dtype == QSM_F64
        ? vsharp_t<double>(static_cast<const double *>(fc), mc, szc, nechoes, vsz, r, thr,
                           static_cast<double *>(flc), mask_out, work, st)
        : vsharp_t<float>(static_cast<const float *>(fc), mc, szc, nechoes, vsz, r, thr,
                          static_cast<float *>(flc), mask_out, work, st);
}

int qsm_sharp(int dtype, const void *f, const uint8_t *mask, const size_t sz[3], size_t nechoes,
              const double vsz[3], double r, double thr, void *fl, uint8_t *mask_out)
{
    int rc = check_common(dtype, sz, nechoes, vsz, thr);
    if (rc) return rc;
    QSM_ARG(f && mask && fl, "null pointer");
    QSM_ARG(std::isfinite(r) && r > 0.0, "radius must be finite and > 0");
    return host_form(
        dtype, f, mask, sz, nechoes, fl, mask_out,
        [&](const size_t *szc) { return qsm_sharp_work_bytes(dtype, szc, vsz, r); },
        [&](const void *df, const uint8_t *dm, const size_t *szc, void *dfl, uint8_t *dmo,
            void *dw) {
            return qsm_sharp_dev(dtype, df, dm, szc, nechoes, vsz, r, thr, dfl, dmo, dw, nullptr);
        });
}

int qsm_vsharp(int dtype, const void *f, const uint8_t *mask, const size_t sz[3],
               size_t nechoes, const double vsz[3], const double *rs, size_t nrs, double thr,
               void *fl, uint8_t *mask_out)
{
    int rc = check_common(dtype, sz, nechoes, vsz, thr);
    if (rc) return rc;
    QSM_ARG(f && mask && fl, "null pointer");
    std::vector<double> r;
    if ((rc = sorted_radii(rs, nrs, r))) return rc;
    return host_form(
        dtype, f, mask, sz, nechoes, fl, mask_out,
        [&](const size_t *szc) {
            return qsm_vsharp_work_bytes(dtype, szc, vsz, r.front(), nechoes);
        },
        [&](const void *df, const uint8_t *dm, const size_t *szc, void *dfl, uint8_t *dmo,
            void *dw) {
            return qsm_vsharp_dev(dtype, df, dm, szc, nechoes, vsz, rs, nrs, thr, dfl, dmo, dw,
                                  nullptr);
        });
}

}  // extern "C"
