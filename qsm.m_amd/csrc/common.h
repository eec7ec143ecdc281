// common.h -- shared host/device helpers of libqsm_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

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

// Per-axis occupancy of a mask (sz[0] x sz[1] x sz[2], x fastest):
// any[a][i] = 1 where slice i of axis a holds a nonzero voxel.  One pass: a
// byte OR per x position and a word OR per row, both vectorisable.
void axis_any(const uint8_t *m, const size_t sz[3], std::vector<char> any[3]);

// A crop of a volume by per-axis index lists (device arrays), for the
// host-pointer entry points, which crop and uncrop on the device.
struct CropMap {
    const int *ix, *iy, *iz;  // listed x, y, z indices of the volume
    int cx, cy, cz;           // crop size (list lengths)
    size_t nx, ny;            // the volume's x and y size
    size_t N, n;              // voxels of the volume and of the crop
};

__device__ __forceinline__ size_t crop_source(const CropMap &c, size_t q)
{
    const size_t cxy = size_t(c.cx) * c.cy;
    const size_t k = q / cxy, r = q - k * cxy, j = r / c.cx, i = r - j * c.cx;
    return size_t(c.ix[i]) + c.nx * (size_t(c.iy[j]) + c.ny * size_t(c.iz[k]));
}

// out[e n + q] = in[e N + source(q)] for nech volumes (NORM: a byte mask
// made 0/1).
template <typename T, bool NORM = false>
__global__ __launch_bounds__(256) void crop_gather_kernel(T *out, const T *in, CropMap c,
                                                          int nech)
{
    const size_t t = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= c.n * size_t(nech)) return;
    const size_t e = t / c.n, q = t - e * c.n;
    const T v = in[e * c.N + crop_source(c, q)];
    out[t] = NORM ? T(v ? 1 : 0) : v;
}

// The inverse: out[e N + source(q)] = in[e n + q] (the rest of out is
// left as it is).
template <typename T>
__global__ __launch_bounds__(256) void crop_scatter_kernel(T *out, const T *in, CropMap c,
                                                           int nech)
{
    const size_t t = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (t >= c.n * size_t(nech)) return;
    const size_t e = t / c.n, q = t - e * c.n;
    out[e * c.N + crop_source(c, q)] = in[t];
}

inline unsigned crop_blocks(const CropMap &c, int nech)
{
    return unsigned((c.n * size_t(nech) + 255) / 256);
}

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
