// common.hip -- error reporting, argument checks and device selection.
#include "common.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>

namespace qsm {

static thread_local char g_err[512] = "no error";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    (void)hipGetLastError();
    return (e == hipErrorOutOfMemory) ? QSM_ENOMEM : QSM_EHIP;
}

int check_grid(const size_t sz[3], const double h[3], size_t min_dim)
{
    QSM_ARG(sz != nullptr && h != nullptr, "null size or voxel-size pointer");
    for (int a = 0; a < 3; ++a) {
        QSM_ARG(sz[a] >= min_dim, "dimension %d is %zu, needs >= %zu", a,
                sz[a], min_dim);
        QSM_ARG(std::isfinite(h[a]) && h[a] > 0.0,
                "voxel size %d must be finite and > 0", a);
    }
    const double n = double(sz[0]) * double(sz[1]) * double(sz[2]);
    QSM_ARG(n * 6.0 < 2147483647.0,
            "volume of %.0f voxels is too large (6*N must fit 32 bits)", n);
    return QSM_OK;
}

int ensure_device()
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        set_error("no HIP device available");
        return QSM_ENODEV;
    }
    return QSM_OK;
}

void axis_any(const uint8_t *m, const size_t sz[3], std::vector<char> any[3])
{
    for (int a = 0; a < 3; ++a) any[a].assign(sz[a], 0);
    std::vector<uint8_t> col(sz[0], 0);
    const size_t nx = sz[0];
    size_t l = 0;
    for (size_t k = 0; k < sz[2]; ++k)
        for (size_t j = 0; j < sz[1]; ++j, l += nx) {
            const uint8_t *row = m + l;
            uint8_t acc = 0;
            for (size_t i = 0; i < nx; ++i) {
                col[i] |= row[i];
                acc |= row[i];
            }
            if (acc) any[1][j] = any[2][k] = 1;
        }
    for (size_t i = 0; i < nx; ++i) any[0][i] = col[i] ? 1 : 0;
}

int DevBuf::alloc(size_t bytes)
{
    if (p) { (void)hipFree(p); p = nullptr; }
    if (bytes == 0) bytes = 1;
    QSM_HIP(hipMalloc(&p, bytes));
    return QSM_OK;
}

}  // namespace qsm

extern "C" {

int qsm_version(void) { return 10000; }

const char *qsm_last_error(void) { return qsm::g_err; }

}
