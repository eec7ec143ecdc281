// common.h -- shared host/device helpers of libqsm_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "qsm_hip.h"

namespace qsm {

// Records the calling thread's last error text (printf-style).
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

// Maps a HIP error to a status code, recording where it happened.
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define QSM_HIP(...)                                                         \
    do {                                                                     \
        hipError_t qsm_e_ = (__VA_ARGS__);                                   \
        if (qsm_e_ != hipSuccess)                                            \
            return ::qsm::hip_fail(qsm_e_, #__VA_ARGS__, __FILE__, __LINE__);\
    } while (0)

#define QSM_ARG(cond, ...)                                                   \
    do {                                                                     \
        if (!(cond)) {                                                       \
            ::qsm::set_error(__VA_ARGS__);                                   \
            return QSM_EARG;                                                 \
        }                                                                    \
    } while (0)

// Checks sz (every dimension >= 1, N*6 addressable with 32-bit voxel
// indices) and h (finite, > 0).  min_dim: smallest allowed dimension.
int check_grid(const size_t sz[3], const double h[3], size_t min_dim);

// Ensures a HIP device is present and current; QSM_ENODEV otherwise.
int ensure_device();

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline size_t dtype_size(int dtype) { return dtype == QSM_F32 ? 4 : 8; }

// Scoped device buffer used by the synchronous host-pointer entry points.
struct DevBuf {
    void *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    int alloc(size_t bytes);
};

// Correctly rounded square root for both storage types (the reference calls
// sqrt on the double path and sqrtf on the single path).
__device__ __forceinline__ double rsqrt_cr(double x) { return sqrt(x); }
__device__ __forceinline__ float rsqrt_cr(float x) { return sqrtf(x); }

// XCD-aware block order: the hardware places block b of a 1-D grid on XCD
// b % 8 (each XCD with its own L2).  Returns a logical block index such that
// each XCD runs one contiguous range of logical blocks, so blocks that share
// cache lines (neighbouring tiles, the two halves of a line) meet in one L2.
__device__ __forceinline__ int xcd_block()
{
    const int nblk = int(gridDim.x), b = int(blockIdx.x);
    const int per = nblk / 8, rem = nblk % 8;
    const int xcd = b % 8, idx = b / 8;
    return (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// xcd_block for a 3-D grid: the (x, y, z) block this block runs as, each
// XCD taking one contiguous range of the grid's blocks in x-fastest order.
// For a grid of (x, y) tiles by planes that is a slab of planes per XCD,
// so a plane's -z and +z neighbours are read through the same L2.
__device__ __forceinline__ void xcd_block3(int &bx, int &by, int &bz)
{
    const int gx = int(gridDim.x), gy = int(gridDim.y);
    const int nblk = gx * gy * int(gridDim.z);
    const int b = int(blockIdx.x) + gx * (int(blockIdx.y) + gy * int(blockIdx.z));
    const int per = nblk / 8, rem = nblk % 8;
    const int xcd = b % 8, idx = b / 8;
    const int lb =
        (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
    bx = lb % gx;
    by = (lb / gx) % gy;
    bz = lb / (gx * gy);
}

}  // namespace qsm
