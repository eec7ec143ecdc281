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

// NA arrays of n 4-byte words (array a at a*n), each read once and
// written once: out[a][l] = in[a][l] * c + in[(a + 1) % NA][l].  The
// state traffic of one fused primal-dual pass (15 arrays in, 15 out) with
// none of its arithmetic or neighbour reads.
template <int NA>
__global__ __launch_bounds__(256) void stream_kernel(float *__restrict__ out,
                                                     const float *__restrict__ in, size_t n)
{
    for (size_t l = size_t(blockIdx.x) * 256 + threadIdx.x; l < n; l += size_t(gridDim.x) * 256) {
        float v[NA];
#pragma unroll
        for (int a = 0; a < NA; ++a) v[a] = in[size_t(a) * n + l];
#pragma unroll
        for (int a = 0; a < NA; ++a) out[size_t(a) * n + l] = v[a] * 1.0001f + v[(a + 1) % NA];
    }
}

template <int NA>
void launch_stream(float *out, const float *in, size_t n, hipStream_t st)
{
    stream_kernel<NA><<<8192, 256, 0, st>>>(out, in, n);
}

}  // namespace

extern "C" int qsm_stream_dev(void *out, const void *in, size_t n, int narr, void *stream)
{
    QSM_ARG(out && in, "null pointer");
    QSM_ARG(narr >= 1 && narr <= 16, "narr must be 1..16");
    const hipStream_t st = qsm::as_stream(stream);
    if (n == 0) return QSM_OK;
    float *o = static_cast<float *>(out);
    const float *i = static_cast<const float *>(in);
    switch (narr) {
#define QSM_STREAM_CASE(k) case k: launch_stream<k>(o, i, n, st); break;
    QSM_STREAM_CASE(1) QSM_STREAM_CASE(2) QSM_STREAM_CASE(3) QSM_STREAM_CASE(4)
    QSM_STREAM_CASE(5) QSM_STREAM_CASE(6) QSM_STREAM_CASE(7) QSM_STREAM_CASE(8)
    QSM_STREAM_CASE(9) QSM_STREAM_CASE(10) QSM_STREAM_CASE(11) QSM_STREAM_CASE(12)
    QSM_STREAM_CASE(13) QSM_STREAM_CASE(14) QSM_STREAM_CASE(15) QSM_STREAM_CASE(16)
#undef QSM_STREAM_CASE
    }
    QSM_HIP(hipGetLastError());
    return QSM_OK;
}

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
