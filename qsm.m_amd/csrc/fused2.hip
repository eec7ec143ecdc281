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
    const int *stop, double *partials)
{
    using I = Ix<TGV>;
    constexpr uint32_t S = sizeof(T);
    constexpr int NC = threads2<BX, BY>();
    __shared__ Lds2<T, TGV, NC> L;

    if (stop && *stop)
        return;

    // XCD-aware order, as the one-iteration pass
    const int lb = xcd_block();
    const int bx = lb % ntx;
    const int by = (lb / ntx) % nty;
    const int bz = lb / (ntx * nty);

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
    fused2_tile<T, TGV, BX, BY, CHECK, 2>(L, c, dc, pc, g.nx, g.ny, g.nz, bx, by, bz, kz, acc);
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
                         const int *stop, double *partials, hipStream_t st)
{
    int ntx, nty, nch;
    fused2_tiles(g, BX - 6, BY - 6, kz, ntx, nty, nch);
    const int nblk = ntx * nty * nch;
    constexpr int NT = threads2<BX, BY>();
    if (partials)
        fused2_kernel<T, TGV, BX, BY, true><<<nblk, NT, 0, st>>>(
            in, out, uint32_t(N), f, mk, g, kz, ntx, nty, nch, dc, pc, stop, partials);
    else
        fused2_kernel<T, TGV, BX, BY, false><<<nblk, NT, 0, st>>>(
            in, out, uint32_t(N), f, mk, g, kz, ntx, nty, nch, dc, pc, stop, partials);
    return hipGetLastError();
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
                         const int *stop, double *partials, hipStream_t st)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3)
        return hipSuccess;
    if (kz < 1)
        return hipErrorInvalidValue;
    if constexpr (sizeof(T) == 4)
        return launch2_shape<T, TGV, QSM_F2_F32_BX, QSM_F2_F32_BY>(in, out, N, f, mk, g, kz, dc,
                                                                   pc, stop, partials, st);
    else
        return launch2_shape<T, TGV, QSM_F2_F64_BX, QSM_F2_F64_BY>(in, out, N, f, mk, g, kz, dc,
                                                                   pc, stop, partials, st);
}

#define QSM_INST(T, B)                                                          \
    template hipError_t launch_fused2<T, B>(const T *, T *, size_t, const T *,  \
                                            const uint8_t *, Geom, int,         \
                                            const DualCoef<T> &,                \
                                            const PrimalCoef<T> &, const int *, \
                                            double *, hipStream_t);
QSM_INST(double, true)
QSM_INST(double, false)
QSM_INST(float, true)
QSM_INST(float, false)
#undef QSM_INST

}  // namespace qsm
