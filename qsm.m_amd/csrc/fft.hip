// fft.hip -- 3D FFT and the FFT dipole convolution for gfx950.
//
// Reference idioms replaced: fft3.m / ifft3.m (separable c2c transforms;
// ifft includes 1/M), dipoleKernel.m:56-66 (k-space dipole kernel) and
// padfastfft.m / unpadfastfft.m ('both' zero padding to 7-smooth sizes), in
// the composition real(ifft3(D .* fft3(pad(chi)))) cropped back.  The pass
// kernels and the half-spectrum real convolution are in fft_dev.h; this file
// holds the transform plans and the entry points.
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "fft_dev.h"

using namespace qsm;
using namespace qsm::fft;

namespace {

// ---- host-side planning ----------------------------------------------------

bool factor(int M, Plan &pl)
{
    pl.M = M;
    pl.mM = recip(unsigned(M));
    pl.nst = 0;
    int rest = M, Ns = 1, off = 0;
    const int radices[] = {8, 13, 11, 7, 5, 4, 3, 2};
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

}  // namespace

namespace qsm {
namespace fft {

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
    QSM_ARG(factor(int(M), np.pl), "%s %zu has a prime factor above 13", what, M);
    const std::vector<Cx<double>> td = twiddles(np.pl);
    // the stage table, then the full circle exp(-2 pi i j / M), j < M
    const size_t nst = td.size();
    std::vector<Cx<T>> t(nst + M);
    for (size_t i = 0; i < nst; ++i)
        t[i] = Cx<T>{T(td[i].re), T(td[i].im)};
    for (size_t j = 0; j < M; ++j) {
        const Cx<double> w = twiddle(long(j), long(M));
        t[nst + j] = Cx<T>{T(w.re), T(w.im)};
    }
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
    np.dtwM = static_cast<const char *>(p) + head + nst * sizeof(Cx<T>);
    pc.plans.emplace(key, np);
    ap = np;
    return QSM_OK;
}

template int make_axis<float>(AxisPlan &, size_t, const char *);
template int make_axis<double>(AxisPlan &, size_t, const char *);

}  // namespace fft
}  // namespace qsm

namespace {

// Multiplier of the dipole convolution at half-spectrum sample (x, y, z):
// Ds(k) * scale.  The reference keeps real(ifftn(D .* fftn(x))); for a real
// x that is the inverse transform of fftn(x) .* Ds with Ds(k) = (D(k) +
// D(-k))/2, the form a half spectrum can carry.  Ds = D except where an even
// axis is at its Nyquist index (there -k has the same index, with the other
// components negated, and a tilted B0 makes D(k) != D(-k)).
//
// Evaluated in the transform's precision T: with kz = b.k and kz' the same
// sum at -k (Nyquist components kept), |k|^2 is common to both, so
// Ds = 1/3 - (kz^2 + kz'^2) / (2 |k|^2) with one reciprocal (hardware
// reciprocal in FP32, refined by two Newton steps in FP64), and away from
// the Nyquist planes kz' = -kz exactly, so Ds = D with no branch.  Ds(0) = 0
// (dipoleKernel.m:59-66).
template <typename T>
struct DipoleOp {
    int M0, M1, M2;
    double dk0, dk1, dk2, b0, b1, b2, scale;

    __device__ static T coord(int m, int n, double dk)
    {
        const int low = n / 2;
        return T(m < n - low ? m : m - n) * T(dk);
    }

    __device__ static T recip(T a)
    {
        if constexpr (sizeof(T) == 4) {
            return __builtin_amdgcn_rcpf(a);
        } else {
            T r = __builtin_amdgcn_rcp(a);
            r = __builtin_fma(r, __builtin_fma(-a, r, T(1)), r);
            return __builtin_fma(r, __builtin_fma(-a, r, T(1)), r);
        }
    }

    __device__ Cx<T> operator()(int x, int y, int z, Cx<T> v) const
    {
        const T X = coord(x, M0, dk0), Y = coord(y, M1, dk1), Z = coord(z, M2, dk2);
        const T bx = T(b0) * X, by = T(b1) * Y, bz = T(b2) * Z;
        const T kz = (bx + by) + bz;
        const T kzm = ((2 * x == M0 ? bx : -bx) + (2 * y == M1 ? by : -by))
                    + (2 * z == M2 ? bz : -bz);
        const T k2 = (X * X + Y * Y) + Z * Z;
        const T h = T(0.5) * (kz * kz + kzm * kzm);
        const T D = k2 > T(0) ? T(1.0 / 3.0) - h * recip(k2) : T(0);
        const T d = D * T(scale);
        return Cx<T>{v.re * d, v.im * d};
    }
};

template <typename T>
int dipole_conv_dev_t(const void *chi, const size_t sz[3], const double h[3],
                      const double bdir[3], const size_t M[3], const size_t off[3],
                      void *field, void *work, hipStream_t st)
{
    // normalised b (dipoleKernel.m:35-38)
    const double bn = std::sqrt(bdir[0] * bdir[0] + bdir[1] * bdir[1] + bdir[2] * bdir[2]);
    QSM_ARG(bn > 0.0, "B0 direction must be non-zero");
    double b[3] = {bdir[0], bdir[1], bdir[2]};
    if (bn != 1.0)
        for (double &x : b) x /= bn;
    DipoleOp<T> op;
    op.M0 = int(M[0]);
    op.M1 = int(M[1]);
    op.M2 = int(M[2]);
    op.dk0 = 1.0 / (double(M[0]) * h[0]);
    op.dk1 = 1.0 / (double(M[1]) * h[1]);
    op.dk2 = 1.0 / (double(M[2]) * h[2]);
    op.b0 = b[0];
    op.b1 = b[1];
    op.b2 = b[2];
    // 1/(M0 M1 M2) of ifftn, and 1/2 for the forward x pass's doubling
    op.scale = 0.5 / (double(M[0]) * double(M[1]) * double(M[2]));
    return real_conv<T, true>(static_cast<const T *>(chi), static_cast<T *>(field), sz, M, off,
                              static_cast<Cx<T> *>(work), op, st);
}

}  // namespace

extern "C" {

int qsm_fft3_dev(int dtype, void *data, const size_t M[3], int inverse, void *stream)
{
    QSM_ARG(data && M, "null pointer");
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    return dtype == QSM_F64
        ? fft3_inplace<double>(static_cast<Cx<double> *>(data), M, inverse, as_stream(stream))
        : fft3_inplace<float>(static_cast<Cx<float> *>(data), M, inverse, as_stream(stream));
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
    return real_conv_work_bytes(M, dtype_size(dtype));  // half spectrum along x
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
