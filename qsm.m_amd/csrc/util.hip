// util.hip -- device utilities: a bandwidth-saturating copy used to move
// volumes between resident solves and to calibrate HBM traffic counters.
#include "common.h"

namespace {

// 16-byte vector copy, one vector per lane in a single pass (no grid
// stride: on gfx950 this form streams ~6.2 TB/s at 2 GiB, a grid-stride
// loop ~4.6); the tail (bytes % 16) by single bytes.
__global__ __launch_bounds__(256) void copy16_kernel(uint4 *dst, const uint4 *src,
                                                     size_t n16)
{
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n16)
        dst[i] = src[i];
}

__global__ void copy1_kernel(unsigned char *dst, const unsigned char *src, size_t n)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        dst[i] = src[i];
}

}  // namespace

extern "C" int qsm_copy_dev(void *dst, const void *src, size_t bytes, void *stream)
{
    QSM_ARG(dst && src, "null pointer");
    QSM_ARG((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 16 == 0,
            "copy pointers must be 16-byte aligned");
    const hipStream_t st = qsm::as_stream(stream);
    const size_t n16 = bytes / 16, tail = bytes % 16;
    if (n16) {
        const size_t blocks = (n16 + 255) / 256;
        QSM_ARG(blocks <= 0x7fffffffu, "copy of %zu bytes too large for one launch", bytes);
        copy16_kernel<<<unsigned(blocks), 256, 0, st>>>(static_cast<uint4 *>(dst),
                                                        static_cast<const uint4 *>(src), n16);
        QSM_HIP(hipGetLastError());
    }
    if (tail) {
        copy1_kernel<<<1, 16, 0, st>>>(static_cast<unsigned char *>(dst) + 16 * n16,
                                       static_cast<const unsigned char *>(src) + 16 * n16, tail);
        QSM_HIP(hipGetLastError());
    }
    return QSM_OK;
}
