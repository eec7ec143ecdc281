// fused.hip -- the single-pass primal-dual kernel for gfx950.
//
// Pass n applies the primal update of iteration n and then the dual update
// of iteration n+1 (see pd_kernels.h).  Layout: every state array lives in
// one bank allocation, array a of voxel l at element a*N + l.  All bank
// accesses are buffer instructions on one 128-bit resource per bank: the
// voxel's byte offset is a 32-bit VGPR shared by every array, the array's
// offset (a*N*sizeof(T)) is an SGPR, and the +/-1 neighbours fold into the
// instruction's immediate offset.  This keeps the 30 array bases of the two
// banks out of the scalar register file (no spills) and removes the 64-bit
// address arithmetic per load.
//
// Block = TX x TY voxels of one z-chunk; each thread owns one (x, y) column
// and marches up the chunk keeping its z-neighbours in registers.  The dual
// of the tile needs the primal extrapolations of its x/y neighbours, which
// the block recomputes on a one-cell ring (-y row, +y row, -x column,
// +x column; the -x/-y cells need only chi_ and eta_).  Extrapolations are
// exchanged through LDS (two plane slots, one barrier per plane).
#include "fused_dev.h"

namespace qsm {
namespace {

using namespace fdev;

template <typename T, bool TGV, int TX, int TY, bool CHECK>
__global__ __launch_bounds__(TX *TY) void fused_kernel(
    const T *in_base, T *out_base, uint32_t N, const T *f,
    const uint8_t *mk, Geom g, int kz, int ntx, int nty, int nchunk,
    DualCoef<T> dc, PrimalCoef<T> pc, const int *stop, double *partials)
{
    using I = Ix<TGV>;
    constexpr uint32_t S = sizeof(T);
    __shared__ FusedLds<T, TGV, TX, TY> L;

    if (stop && *stop)
        return;

    // each XCD takes a contiguous run of logical tiles, so ring neighbours
    // share its L2 (the grid is ntx * nty * nchunk blocks)
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
    fused_tile<T, TGV, TX, TY, CHECK>(L, c, dc, pc, g.nx, g.ny, g.nz, bx, by, bz, kz, acc);
    if constexpr (CHECK)
        block_sum4_f<TX * TY>(acc, partials + 4 * lb);
}

void fused_tiles(Geom g, FusedShape sh, int &ntx, int &nty, int &nch)
{
    ntx = (g.nx + sh.tx - 1) / sh.tx;
    nty = (g.ny + sh.ty - 1) / sh.ty;
    nch = (g.nz - 2 + sh.kz - 1) / sh.kz;
}

template <typename T, bool TGV, int TX, int TY>
hipError_t launch_shape(const T *in, T *out, size_t N, const T *f,
                        const uint8_t *mk, Geom g, FusedShape sh,
                        const DualCoef<T> &dc, const PrimalCoef<T> &pc,
                        const int *stop, double *partials, hipStream_t st)
{
    int ntx, nty, nch;
    fused_tiles(g, sh, ntx, nty, nch);
    const int nblk = ntx * nty * nch;
    if (partials)
        fused_kernel<T, TGV, TX, TY, true><<<nblk, TX * TY, 0, st>>>(
            in, out, uint32_t(N), f, mk, g, sh.kz, ntx, nty, nch, dc, pc, stop, partials);
    else
        fused_kernel<T, TGV, TX, TY, false><<<nblk, TX * TY, 0, st>>>(
            in, out, uint32_t(N), f, mk, g, sh.kz, ntx, nty, nch, dc, pc, stop, partials);
    return hipGetLastError();
}

}  // namespace

bool fused_shape_supported(int tx, int ty)
{
    return (tx == 64 && (ty == 4 || ty == 8 || ty == 16))
           || (tx == 32 && (ty == 8 || ty == 16 || ty == 32));
}

bool fused_size_supported(int kind, int dtype, size_t N)
{
    const double bank = double(kind == QSM_TGV ? 15 : 6) * double(N) * double(dtype_size(dtype));
    return bank < 4294967295.0;
}

int fused_blocks(Geom g, FusedShape sh)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3)
        return 0;
    int ntx, nty, nch;
    fused_tiles(g, sh, ntx, nty, nch);
    return ntx * nty * nch;
}

template <typename T, bool TGV>
hipError_t launch_fused(const T *in, T *out, size_t N, const T *f,
                        const uint8_t *mk, Geom g, FusedShape sh,
                        const DualCoef<T> &dc, const PrimalCoef<T> &pc,
                        const int *stop, double *partials, hipStream_t st)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3)
        return hipSuccess;
    if (sh.kz < 1 || !fused_shape_supported(sh.tx, sh.ty))
        return hipErrorInvalidValue;
#define QSM_FUSED_CASE(X, Y)                                                   \
    if (sh.tx == X && sh.ty == Y)                                              \
        return launch_shape<T, TGV, X, Y>(in, out, N, f, mk, g, sh, dc, pc,    \
                                          stop, partials, st);
    QSM_FUSED_CASE(64, 4)
    QSM_FUSED_CASE(64, 8)
    QSM_FUSED_CASE(64, 16)
    QSM_FUSED_CASE(32, 8)
    QSM_FUSED_CASE(32, 16)
    QSM_FUSED_CASE(32, 32)
#undef QSM_FUSED_CASE
    return hipErrorInvalidValue;
}

#define QSM_INST(T, B)                                                         \
    template hipError_t launch_fused<T, B>(const T *, T *, size_t, const T *,  \
                                           const uint8_t *, Geom, FusedShape,  \
                                           const DualCoef<T> &,                \
                                           const PrimalCoef<T> &, const int *, \
                                           double *, hipStream_t);
QSM_INST(double, true)
QSM_INST(double, false)
QSM_INST(float, true)
QSM_INST(float, false)
#undef QSM_INST

}  // namespace qsm
