// fused2.hip -- the single-volume kernel of the two-iteration pass
// (fused2_dev.h) and its launchers.
//
// Pass (n, n+1) reads the state of bank A (primal iterates of n-1, dual of n)
// once and writes the state two iterations later to bank B (primal of n+1,
// dual of n+2): per iteration a voxel's state crosses HBM half as often as in
// the one-iteration pass (fused.hip), bitwise the same iterates.
#include "fused2_dev.h"

namespace qsm {
namespace {

using namespace fdev;

template <typename T, bool TGV, int BX, int BY, bool CHECK>
__global__ __launch_bounds__((BX * BY + 63) / 64 * 64) void fused2_kernel(
    const T *in_base, T *out_base, uint32_t N, const T *f, const uint8_t *mk, Geom g,
    int kz, int ntx, int nty, int nchunk, DualCoef<T> dc, PrimalCoef<T> pc,
    const int *stop, double *partials, const F2Item *items)
{
    using I = Ix<TGV>;
    constexpr uint32_t S = sizeof(T);
    constexpr int NC = threads2<BX, BY>();
    __shared__ Lds2<T, TGV, NC> L;

    if (stop && *stop)
        return;

    // XCD-aware order, as the one-iteration pass; a work list gives each
    // block its tile and owned planes (consecutive items: neighbouring
    // tiles, one XCD's L2)
    const int lb = xcd_block();
    int bx, by, k0, k1;
    if (items) {
        const F2Item it = items[lb];
        bx = it.bx;
        by = it.by;
        k0 = it.k0;
        k1 = it.k1;
    } else {
        bx = lb % ntx;
        by = (lb / ntx) % nty;
        k0 = 1 + (lb / (ntx * nty)) * kz;
        k1 = min(k0 + kz, g.nz - 1);
    }

    Ctx<T> c;
    c.nb = N * S;
    c.in = make_rsrc(in_base, uint32_t(I::NARR) * c.nb);
    c.out = make_rsrc(out_base, uint32_t(I::NARR) * c.nb);
    c.f = make_rsrc(f, c.nb);
    c.mk = make_rsrc(mk, N);
    c.sk = g.nx * g.ny;
    c.syb = uint32_t(g.nx) * S;
    c.skb = uint32_t(c.sk) * S;
    c.nz = g.nz;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    fused2_tile<T, TGV, BX, BY, CHECK, 2>(L, c, dc, pc, g.nx, g.ny, g.nz, bx, by, k0, k1, acc);
    if constexpr (CHECK)
        block_sum4_f<NC>(acc, partials + 4 * lb);
}

void fused2_tiles(Geom g, int tx, int ty, int kz, int &ntx, int &nty, int &nch)
{
    ntx = (g.nx + tx - 1) / tx;
    nty = (g.ny + ty - 1) / ty;
    nch = (g.nz - 2 + kz - 1) / kz;
}

template <typename T, bool TGV, int BX, int BY>
hipError_t launch2_shape(const T *in, T *out, size_t N, const T *f, const uint8_t *mk, Geom g,
                         int kz, const DualCoef<T> &dc, const PrimalCoef<T> &pc,
                         const int *stop, double *partials, const F2Item *items, int nitems,
                         hipStream_t st)
{
    int ntx, nty, nch;
    fused2_tiles(g, BX - 6, BY - 6, kz, ntx, nty, nch);
    const int nblk = items ? nitems : ntx * nty * nch;
    if (nblk <= 0)
        return hipSuccess;
    constexpr int NT = threads2<BX, BY>();
    if (partials)
        fused2_kernel<T, TGV, BX, BY, true><<<nblk, NT, 0, st>>>(
            in, out, uint32_t(N), f, mk, g, kz, ntx, nty, nch, dc, pc, stop, partials, items);
    else
        fused2_kernel<T, TGV, BX, BY, false><<<nblk, NT, 0, st>>>(
            in, out, uint32_t(N), f, mk, g, kz, ntx, nty, nch, dc, pc, stop, partials, items);
    return hipGetLastError();
}

// Tile (bx, by) of the two-iteration pass: [lo, hi) = the planes in
// [1, nz - 1) where one of its owned voxels has a mask bit (hi = lo when
// none).  One block per tile; each thread scans its columns from both ends.
__global__ __launch_bounds__(256) void tile_zrange_kernel(const uint8_t *mask, Geom g, int tx,
                                                          int ty, int ntx, int *out)
{
    __shared__ int slo, shi;
    const int bx = int(blockIdx.x) % ntx, by = int(blockIdx.x) / ntx;
    if (threadIdx.x == 0) {
        slo = g.nz;
        shi = 0;
    }
    __syncthreads();
    int lo = g.nz, hi = 0;
    const size_t sk = size_t(g.nx) * g.ny;
    for (int c = int(threadIdx.x); c < tx * ty; c += int(blockDim.x)) {
        const int x = bx * tx + c % tx, y = by * ty + c / tx;
        if (x >= g.nx || y >= g.ny)
            continue;
        const uint8_t *col = mask + x + size_t(g.nx) * y;
        for (int k = 1; k < g.nz - 1 && k < lo; ++k)
            if (col[sk * k]) {
                lo = k;
                break;
            }
        for (int k = g.nz - 2; k >= 1 && k + 1 > hi; --k)
            if (col[sk * k]) {
                hi = k + 1;
                break;
            }
    }
    if (lo < hi) {
        atomicMin(&slo, lo);
        atomicMax(&shi, hi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int l = slo, h = shi;
        out[2 * blockIdx.x] = l < h ? l : 1;
        out[2 * blockIdx.x + 1] = l < h ? h : 1;
    }
}

}  // namespace

void fused2_tile(int dtype, int &tx, int &ty)
{
    if (dtype == QSM_F32) { tx = QSM_F2_F32_BX - 6; ty = QSM_F2_F32_BY - 6; }
    else { tx = QSM_F2_F64_BX - 6; ty = QSM_F2_F64_BY - 6; }
}

int fused2_blocks(Geom g, int dtype, int kz)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3 || kz < 1)
        return 0;
    int tx, ty, ntx, nty, nch;
    fused2_tile(dtype, tx, ty);
    fused2_tiles(g, tx, ty, kz, ntx, nty, nch);
    return ntx * nty * nch;
}

template <typename T, bool TGV>
hipError_t launch_fused2(const T *in, T *out, size_t N, const T *f, const uint8_t *mk, Geom g,
                         int kz, const DualCoef<T> &dc, const PrimalCoef<T> &pc,
                         const int *stop, double *partials, const F2Item *items, int nitems,
                         hipStream_t st)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3)
        return hipSuccess;
    if (kz < 1)
        return hipErrorInvalidValue;
    if constexpr (sizeof(T) == 4)
        return launch2_shape<T, TGV, QSM_F2_F32_BX, QSM_F2_F32_BY>(
            in, out, N, f, mk, g, kz, dc, pc, stop, partials, items, nitems, st);
    else
        return launch2_shape<T, TGV, QSM_F2_F64_BX, QSM_F2_F64_BY>(
            in, out, N, f, mk, g, kz, dc, pc, stop, partials, items, nitems, st);
}

hipError_t launch_tile_zrange(const uint8_t *mask, Geom g, int dtype, int *out, hipStream_t st)
{
    int tx, ty, ntx, nty, nch;
    fused2_tile(dtype, tx, ty);
    fused2_tiles(g, tx, ty, 1, ntx, nty, nch);
    if (ntx * nty > 0)
        tile_zrange_kernel<<<ntx * nty, 256, 0, st>>>(mask, g, tx, ty, ntx, out);
    return hipGetLastError();
}

#define QSM_INST(T, B)                                                          \
    template hipError_t launch_fused2<T, B>(const T *, T *, size_t, const T *,  \
                                            const uint8_t *, Geom, int,         \
                                            const DualCoef<T> &,                \
                                            const PrimalCoef<T> &, const int *, \
                                            double *, const F2Item *, int,      \
                                            hipStream_t);
QSM_INST(double, true)
QSM_INST(double, false)
QSM_INST(float, true)
QSM_INST(float, false)
#undef QSM_INST

}  // namespace qsm
