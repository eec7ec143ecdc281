// batch.hip -- many uneven primal-dual inversions advanced together.
//
// Every live volume has its own resident state (two banks, f, packed mask,
// control block, partial sums) and a descriptor in device memory.  One
// iteration of the whole batch is two launches:
//
//  * batch_fused_kernel: a persistent grid (a fixed number of blocks, a
//    multiple of the 8 XCDs) strides over a device-side list of tiles that
//    spans all live volumes; each tile runs the single-pass primal(n) +
//    dual(n+1) march of fused_dev.h with its volume's banks (chosen by that
//    volume's own iteration parity), coefficients and convergence schedule;
//  * batch_advance_kernel: one block per volume slot reduces that volume's
//    partial sums when it checked convergence (fixed order), applies its
//    tolerance test, and advances its iteration counter.
//
// Everything a launch needs is read from device memory, so the launch
// arguments never change: a chunk of iterations is captured once into a
// hipGraph and replayed; volumes join and leave between chunks by updating
// the descriptors and the tile list.  Each volume's arithmetic is that of
// the single-volume solver (same per-voxel updates in the same order, the
// convergence sums reduced in a fixed order per volume), so a result is
// bitwise the same alone and in any mix.
//
// A batch created with tile shape 0x0 runs the two-iteration march of
// fused2_dev.h instead (batch_fused2_kernel): each launch advances a volume
// by two iterations, or by one where its next iteration is a convergence
// check (which must see the state after it) or its chunk or budget has one
// iteration left.  The choice is made on the device per volume
// (batch_chunk2_kernel at the start of a chunk, batch_advance2_kernel after
// each launch), so a captured chunk holds a fixed number of launches, the
// last ones doing nothing for volumes already at the chunk's end.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "fused2_dev.h"

namespace qsm {
namespace {

using namespace fdev;

template <typename T>
struct VolDesc {
    // set by the host when the volume is admitted
    int nx, ny, nz, kind;
    uint32_t N;
    int ntx, nty, nch, kz;
    int tile0;               // first entry of this volume in the tile list
    const T *bank[2];
    const T *f;
    const uint8_t *mk;
    Ctl *ctl;
    double *partials;        // 4 doubles per tile
    double *hist;
    int hist_cap;
    uint32_t target;         // iterations to run
    double tol;
    DualCoef<T> dc;
    PrimalCoef<T> pc;
    // advanced on the device
    uint32_t iter;           // iterations completed
    int live;                // 1 while iterating
    uint32_t cur;            // bank holding the iterates of `iter`
    int mode;                // two-iteration march: iterations of the next launch
    uint32_t chunk_end;      // two-iteration march: iteration ending this chunk
    // two-iteration march: the volume's entries of the tile list as work
    // items (tile, owned planes) when its mask leaves planes to skip, else
    // null (the uniform tiles x chunks); nitems entries
    int nitems;
    const F2Item *items;
};

struct TileList {
    int ntiles[2];   // tiles of TV volumes, of TGV volumes (list order)
    int pad[2];
    // two-iteration march: tiles claimed so far from each XCD's range of
    // the list, per kind (TV 0-7, TGV 8-15); zeroed before every launch
    int claimed[16];
};

// One persistent launch per kind: kind KIND's tiles are a contiguous part
// of the tile list (TV tiles first).  Specialising per kind keeps the
// register count that of one march (two kinds inlined in one kernel would
// cost occupancy).
template <typename T, bool TGV, int TX, int TY>
__global__ __launch_bounds__(TX *TY) void batch_fused_kernel(
    const VolDesc<T> *__restrict__ vols, const int *__restrict__ tile_vol,
    const TileList *__restrict__ tl)
{
    __shared__ FusedLds<T, TGV, TX, TY> L;
    const int T_ = tl->ntiles[TGV ? 1 : 0];
    const int *tv = tile_vol + (TGV ? tl->ntiles[0] : 0);
    // XCD-aware striding: XCD x (= block % 8) walks a contiguous range of
    // the tile list, which keeps a volume's neighbouring tiles on one L2.
    const int xcd = int(blockIdx.x) % 8;
    const int nper = int(gridDim.x) / 8;
    const int bl = int(blockIdx.x) / 8;
    const int span = (T_ + 7) / 8;
    const int t_end = min(T_, (xcd + 1) * span);
    for (int t = xcd * span + bl; t < t_end; t += nper) {
        // the descriptor is read-only here (restrict): its fields stay in
        // registers through the march
        const VolDesc<T> *dp = vols + tv[t];
        if (!dp->live || __builtin_amdgcn_readfirstlane(dp->ctl->stop))
            continue;
        const int lt = t - dp->tile0;
        const int ntx = dp->ntx, nty = dp->nty;
        const int bx = lt % ntx;
        const int by = (lt / ntx) % nty;
        const int bz = lt / (ntx * nty);
        const uint32_t it = dp->iter;
        const bool check = ((it + 1) % 10) == 0;  // pd_tgv_mex.c:298,309
        const T *b0 = dp->bank[0], *b1 = dp->bank[1];
        const DualCoef<T> dc = dp->dc;
        const PrimalCoef<T> pc = dp->pc;
        const int nx = dp->nx, ny = dp->ny, nz = dp->nz, kz = dp->kz;
        constexpr uint32_t S = sizeof(T);
        constexpr uint32_t NARR = TGV ? 15u : 6u;
        Ctx<T> c;
        c.nb = dp->N * S;
        c.in = make_rsrc((it & 1u) ? b1 : b0, NARR * c.nb);
        c.out = make_rsrc((it & 1u) ? b0 : b1, NARR * c.nb);
        c.f = make_rsrc(dp->f, c.nb);
        c.mk = make_rsrc(dp->mk, dp->N);
        c.sk = nx * ny;
        c.syb = uint32_t(nx) * S;
        c.skb = uint32_t(c.sk) * S;
        c.nz = nz;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        if (check) {
            fused_tile<T, TGV, TX, TY, true>(L, c, dc, pc, nx, ny, nz, bx, by, bz, kz, acc);
            block_sum4_f<TX * TY>(acc, dp->partials + 4 * lt);
        } else {
            fused_tile<T, TGV, TX, TY, false>(L, c, dc, pc, nx, ny, nz, bx, by, bz, kz, acc);
        }
    }
}

constexpr int ADV_T = 256;

// One block per slot: finalize the convergence check of volumes that
// checked in the iteration just run, then advance their counters.
template <typename T>
__global__ __launch_bounds__(ADV_T) void batch_advance_kernel(VolDesc<T> *vols, int nslots)
{
    VolDesc<T> &d = vols[blockIdx.x];
    if (!d.live)
        return;
    Ctl *ctl = d.ctl;
    if (ctl->stop)
        return;
    const uint32_t n = d.iter + 1;
    if (n % 10 == 0) {
        const int nblk = d.ntx * d.nty * d.nch;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int b = threadIdx.x; b < nblk; b += ADV_T)
            for (int q = 0; q < 4; ++q)
                s[q] += d.partials[4 * b + q];
        __shared__ double tot[4];
        block_sum4_f<ADV_T>(s, tot);
        __syncthreads();
        if (threadIdx.x == 0) {
            const double n1 = sqrt(tot[2] / tot[0]);  // pd_tgv_mex.c:1004-1010
            const double n2 = sqrt(tot[3] / tot[1]);
            const uint32_t c = ctl->checks;
            if (d.hist && int(c) < d.hist_cap) {
                d.hist[2 * c] = n1;
                d.hist[2 * c + 1] = n2;
            }
            ctl->checks = c + 1;
            ctl->nr[0] = n1;
            ctl->nr[1] = n2;
            if ((n1 < d.tol) && (n2 < 2.0 * d.tol)) {
                ctl->stop = 1;
                ctl->stop_iter = n;
                ctl->stop_bank = n & 1u;  // iteration n wrote bank n & 1
            }
        }
    }
    if (threadIdx.x == 0) {
        d.iter = n;
        d.cur = n & 1u;
        if (n >= d.target || ctl->stop)
            d.live = 0;
    }
}

// ---- the two-iteration march ----------------------------------------------

// One persistent launch per kind over the tile list, as batch_fused_kernel,
// each tile advancing its volume by the volume's mode (0: nothing).
// Claims the next tile of the list for the calling block (thread 0) and
// broadcasts it: first from the block's own XCD's contiguous range (so a
// volume's neighbouring tiles meet in one L2), then from the other XCDs'
// ranges; -1 when all are taken.  Tiles of a march differ in cost (sizes,
// chunk depths, kinds), so claiming balances what a fixed share would not.
__device__ __forceinline__ int claim_tile(int *claimed, int ntiles, int xcd, int *slot)
{
    if (threadIdx.x == 0) {
        const int span = (ntiles + 7) / 8;
        int got = -1;
        for (int k = 0; k < 8 && got < 0; ++k) {
            const int x = (xcd + k) % 8;
            const int len = min(ntiles, (x + 1) * span) - x * span;
            if (len <= 0)
                continue;
            if (k > 0 && __hip_atomic_load(&claimed[x], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT) >= len)
                continue;
            const int i = atomicAdd(&claimed[x], 1);
            if (i < len)
                got = x * span + i;
        }
        *slot = got;
    }
    __syncthreads();
    // the same value in every lane: read as a wave-uniform (scalar) value,
    // so the tile loop and every branch on the tile's descriptor are uniform
    // control flow around the march's barriers (a value the compiler must
    // treat as divergent turns the loop into per-lane exec-mask loops, in
    // which lane 0's next claim and the other lanes' barriers separate)
    const int t = __builtin_amdgcn_readfirstlane(*slot);
    __syncthreads();  // every thread has read the slot before it is reused
    return t;
}

// The tile t of volume *dp advanced by `mode` iterations (uniform over the
// block).
template <typename T, bool TGV, int BX, int BY>
__device__ __forceinline__ void march2_tile(Lds2<T, TGV, threads2<BX, BY>()> &L,
                                            const VolDesc<T> *dp, int t, int mode)
{
    constexpr int NC = threads2<BX, BY>();
    const int lt = t - dp->tile0;
    const int ntx = dp->ntx, nty = dp->nty;
    const bool check = ((dp->iter + uint32_t(mode)) % 10) == 0;  // pd_tgv_mex.c:298,309
    const T *bi = dp->bank[dp->cur & 1u], *bo = dp->bank[(dp->cur & 1u) ^ 1u];
    const DualCoef<T> dc = dp->dc;
    const PrimalCoef<T> pc = dp->pc;
    const int nx = dp->nx, ny = dp->ny, nz = dp->nz, kz = dp->kz;
    int bx, by, k0, k1;  // tile and owned planes
    if (dp->items) {
        const F2Item it = dp->items[lt];
        bx = it.bx;
        by = it.by;
        k0 = it.k0;
        k1 = it.k1;
    } else {
        bx = lt % ntx;
        by = (lt / ntx) % nty;
        k0 = 1 + (lt / (ntx * nty)) * kz;
        k1 = min(k0 + kz, nz - 1);
    }
    constexpr uint32_t S = sizeof(T);
    constexpr uint32_t NARR = TGV ? 15u : 6u;
    Ctx<T> c;
    c.nb = dp->N * S;
    c.in = make_rsrc(bi, NARR * c.nb);
    c.out = make_rsrc(bo, NARR * c.nb);
    c.f = make_rsrc(dp->f, c.nb);
    c.mk = make_rsrc(dp->mk, dp->N);
    c.sk = nx * ny;
    c.syb = uint32_t(nx) * S;
    c.skb = uint32_t(c.sk) * S;
    c.nz = nz;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // one specialised march per (mode, check)
    if (mode == 2) {
        if (check) {
            fused2_tile<T, TGV, BX, BY, true, 2>(L, c, dc, pc, nx, ny, nz, bx, by, k0, k1, acc);
            block_sum4_f<NC>(acc, dp->partials + 4 * lt);
        } else {
            fused2_tile<T, TGV, BX, BY, false, 2>(L, c, dc, pc, nx, ny, nz, bx, by, k0, k1, acc);
        }
    } else if (check) {
        fused2_tile<T, TGV, BX, BY, true, 1>(L, c, dc, pc, nx, ny, nz, bx, by, k0, k1, acc);
        block_sum4_f<NC>(acc, dp->partials + 4 * lt);
    } else {
        fused2_tile<T, TGV, BX, BY, false, 1>(L, c, dc, pc, nx, ny, nz, bx, by, k0, k1, acc);
    }
}

template <typename T, bool TGV, int BX, int BY>
__global__ __launch_bounds__((BX * BY + 63) / 64 * 64) void batch_fused2_kernel(
    const VolDesc<T> *__restrict__ vols, const int *__restrict__ tile_vol,
    TileList *tl)
{
    constexpr int NC = threads2<BX, BY>();
    __shared__ Lds2<T, TGV, NC> L;
    __shared__ int slot;
    const int T_ = tl->ntiles[TGV ? 1 : 0];
    const int *tv = tile_vol + (TGV ? tl->ntiles[0] : 0);
    int *claimed = tl->claimed + (TGV ? 8 : 0);
    const int xcd = int(blockIdx.x) % 8;
    // every branch below is on wave-uniform values (the claimed index, the
    // descriptor's scalar fields, the stop flag read through readfirstlane)
    // and the loop has one exit, at its head: the march's barriers then sit
    // in uniform control flow
    for (int t = claim_tile(claimed, T_, xcd, &slot); t >= 0;
         t = claim_tile(claimed, T_, xcd, &slot)) {
        const VolDesc<T> *dp = vols + tv[t];
        const int mode = dp->mode;
        const bool run = dp->live && mode != 0
                         && !__builtin_amdgcn_readfirstlane(dp->ctl->stop);
        if (run)
            march2_tile<T, TGV, BX, BY>(L, dp, t, mode);
    }
}

// Iterations of a volume's next launch: two, unless the first of them is a
// convergence check or only one is left before min(chunk_end, target).
__device__ __forceinline__ int next_mode(uint32_t iter, uint32_t chunk_end, uint32_t target)
{
    const uint32_t end = min(chunk_end, target);
    if (iter >= end)
        return 0;
    if (end - iter >= 2 && ((iter + 1) % 10) != 0)
        return 2;
    return 1;
}

// Start of a chunk of `chunk` iterations: every live volume may advance
// that many.
template <typename T>
__global__ void batch_chunk2_kernel(VolDesc<T> *vols, int nslots, uint32_t chunk, TileList *tl)
{
    const int v = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (v < 16)
        tl->claimed[v] = 0;
    if (v >= nslots)
        return;
    VolDesc<T> &d = vols[v];
    if (!d.live)
        return;
    d.chunk_end = d.iter + chunk;
    d.mode = next_mode(d.iter, d.chunk_end, d.target);
}

// After a launch: finalize the check of volumes that checked (the last
// iteration the launch ran), advance counters and banks, plan the next.
template <typename T>
__global__ __launch_bounds__(ADV_T) void batch_advance2_kernel(VolDesc<T> *vols, int nslots,
                                                               TileList *tl)
{
    if (blockIdx.x == 0 && threadIdx.x < 16)
        tl->claimed[threadIdx.x] = 0;  // for the next launch
    VolDesc<T> &d = vols[blockIdx.x];
    if (!d.live)
        return;
    Ctl *ctl = d.ctl;
    const int mode = d.mode;
    if (ctl->stop || mode == 0)
        return;
    const uint32_t n = d.iter + uint32_t(mode);
    const uint32_t bank = (d.cur & 1u) ^ 1u;
    if (n % 10 == 0) {
        const int nblk = d.items ? d.nitems : d.ntx * d.nty * d.nch;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int b = threadIdx.x; b < nblk; b += ADV_T)
            for (int q = 0; q < 4; ++q)
                s[q] += d.partials[4 * b + q];
        __shared__ double tot[4];
        block_sum4_f<ADV_T>(s, tot);
        __syncthreads();
        if (threadIdx.x == 0) {
            const double n1 = sqrt(tot[2] / tot[0]);  // pd_tgv_mex.c:1004-1010
            const double n2 = sqrt(tot[3] / tot[1]);
            const uint32_t c = ctl->checks;
            if (d.hist && int(c) < d.hist_cap) {
                d.hist[2 * c] = n1;
                d.hist[2 * c + 1] = n2;
            }
            ctl->checks = c + 1;
            ctl->nr[0] = n1;
            ctl->nr[1] = n2;
            if ((n1 < d.tol) && (n2 < 2.0 * d.tol)) {
                ctl->stop = 1;
                ctl->stop_iter = n;
                ctl->stop_bank = bank;
            }
        }
    }
    if (threadIdx.x == 0) {
        d.iter = n;
        d.cur = bank;
        const bool done = n >= d.target || ctl->stop;
        if (done)
            d.live = 0;
        d.mode = done ? 0 : next_mode(n, d.chunk_end, d.target);
    }
}

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace
}  // namespace qsm

using namespace qsm;

struct BatchVol {
    bool used = false;
    int kind = 0;
    size_t sz[3] = {0, 0, 0};
    size_t N = 0, bytes = 0;
    char *mem = nullptr;
    int ntiles = 0;
    int hist_cap = 0;
    bool uploaded = false;   // its descriptor is on the device (the device
                             // then owns iter and live)
};

struct qsm_batch {
    int dtype, tx, ty;
    bool two = false;            // tiles 0x0: the two-iteration march
    size_t budget, used;
    int max_vols;
    hipStream_t stream;
    bool own_stream = false;     // created here (the caller passed none)
    void *d_vols = nullptr;      // VolDesc<T>[max_vols]
    int *d_tilevol = nullptr;    // tile -> slot
    void *d_tl = nullptr;        // TileList
    int tile_cap = 0;
    std::vector<BatchVol> vols;
    std::vector<unsigned char> h_desc;  // host mirror of the descriptors
    size_t desc_size = 0;
    int grid[2] = {0, 0};         // persistent grid of the TV / TGV kernel
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint32_t graph_chunk = 0;
    bool use_graph = true;
    bool dirty = true;            // tile list needs upload
};

namespace {

template <typename T>
VolDesc<T> &desc(qsm_batch *b, int v)
{
    return reinterpret_cast<VolDesc<T> *>(b->h_desc.data())[v];
}

template <typename T, bool TGV, int TX, int TY>
int persistent_grid(int *out)
{
    int dev = 0, ncu = 0, occ = 0;
    QSM_HIP(hipGetDevice(&dev));
    QSM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    QSM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, reinterpret_cast<const void *>(&batch_fused_kernel<T, TGV, TX, TY>), TX * TY, 0));
    if (occ < 1) occ = 1;
    int g = ncu * occ;
    g = std::max(8, g / 8 * 8);
    *out = g;
    return QSM_OK;
}

template <typename T, int TX, int TY>
int persistent_grids(int out[2])
{
    int rc = persistent_grid<T, false, TX, TY>(&out[0]);
    return rc ? rc : persistent_grid<T, true, TX, TY>(&out[1]);
}

template <typename T, bool TGV, int BX, int BY>
int persistent_grid2(int *out)
{
    int dev = 0, ncu = 0, occ = 0;
    QSM_HIP(hipGetDevice(&dev));
    QSM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    QSM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, reinterpret_cast<const void *>(&batch_fused2_kernel<T, TGV, BX, BY>),
        threads2<BX, BY>(), 0));
    if (occ < 1) occ = 1;
    *out = std::max(8, ncu * occ / 8 * 8);
    return QSM_OK;
}

template <typename T, int BX, int BY>
int persistent_grids2(int out[2])
{
    int rc = persistent_grid2<T, false, BX, BY>(&out[0]);
    return rc ? rc : persistent_grid2<T, true, BX, BY>(&out[1]);
}

// Launches of a captured chunk of `chunk` iterations in the two-iteration
// march: a volume runs (chunk + singles) / 2 of them, and its singles (a
// check falling on the first of two iterations, at most one per ten, and
// the chunk's last odd one) number at most ceil(chunk / 10) + 1.
inline uint32_t launches2(uint32_t chunk)
{
    return (chunk + (chunk + 9) / 10 + 1 + 1) / 2;
}

template <typename T>
hipError_t launch_iteration2(qsm_batch *b)
{
    auto *vols = static_cast<VolDesc<T> *>(b->d_vols);
    auto *tl = static_cast<TileList *>(b->d_tl);
    if constexpr (sizeof(T) == 4) {
        constexpr int BX = QSM_F2_F32_BX, BY = QSM_F2_F32_BY;
        batch_fused2_kernel<T, false, BX, BY><<<b->grid[0], threads2<BX, BY>(), 0, b->stream>>>(
            vols, b->d_tilevol, tl);
        batch_fused2_kernel<T, true, BX, BY><<<b->grid[1], threads2<BX, BY>(), 0, b->stream>>>(
            vols, b->d_tilevol, tl);
    } else {
        constexpr int BX = QSM_F2_F64_BX, BY = QSM_F2_F64_BY;
        batch_fused2_kernel<T, false, BX, BY><<<b->grid[0], threads2<BX, BY>(), 0, b->stream>>>(
            vols, b->d_tilevol, tl);
        batch_fused2_kernel<T, true, BX, BY><<<b->grid[1], threads2<BX, BY>(), 0, b->stream>>>(
            vols, b->d_tilevol, tl);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    batch_advance2_kernel<T><<<b->max_vols, ADV_T, 0, b->stream>>>(vols, b->max_vols, tl);
    return hipGetLastError();
}

// One chunk of `chunk` iterations (both marches).
template <typename T>
hipError_t launch_chunk(qsm_batch *b, uint32_t chunk);

template <typename T>
hipError_t launch_iteration(qsm_batch *b)
{
    auto *vols = static_cast<VolDesc<T> *>(b->d_vols);
    const auto *tl = static_cast<const TileList *>(b->d_tl);
    // both kinds every iteration (a kind with no tiles exits at once), so
    // a captured graph stays valid while the mix of kinds changes
#define QSM_B(X, Y)                                                            \
    if (b->tx == X && b->ty == Y) {                                            \
        batch_fused_kernel<T, false, X, Y><<<b->grid[0], X * Y, 0, b->stream>>>( \
            vols, b->d_tilevol, tl);                                           \
        batch_fused_kernel<T, true, X, Y><<<b->grid[1], X * Y, 0, b->stream>>>(  \
            vols, b->d_tilevol, tl);                                           \
    }
    QSM_B(64, 4) QSM_B(64, 8) QSM_B(64, 16) QSM_B(32, 8) QSM_B(32, 16) QSM_B(32, 32)
#undef QSM_B
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    batch_advance_kernel<T><<<b->max_vols, ADV_T, 0, b->stream>>>(vols, b->max_vols);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_chunk(qsm_batch *b, uint32_t chunk)
{
    if (!b->two) {
        for (uint32_t i = 0; i < chunk; ++i) {
            const hipError_t e = launch_iteration<T>(b);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    auto *vols = static_cast<VolDesc<T> *>(b->d_vols);
    batch_chunk2_kernel<T><<<(b->max_vols + 255) / 256, 256, 0, b->stream>>>(
        vols, b->max_vols, chunk, static_cast<TileList *>(b->d_tl));
    hipError_t e = hipGetLastError();
    for (uint32_t i = 0; i < launches2(chunk) && e == hipSuccess; ++i)
        e = launch_iteration2<T>(b);
    return e;
}

int grid_for(qsm_batch *b)
{
    if (b->two)
        return b->dtype == QSM_F64
            ? persistent_grids2<double, QSM_F2_F64_BX, QSM_F2_F64_BY>(b->grid)
            : persistent_grids2<float, QSM_F2_F32_BX, QSM_F2_F32_BY>(b->grid);
#define QSM_G(X, Y)                                                            \
    if (b->tx == X && b->ty == Y)                                              \
        return b->dtype == QSM_F64 ? persistent_grids<double, X, Y>(b->grid)   \
                                   : persistent_grids<float, X, Y>(b->grid);
    QSM_G(64, 4) QSM_G(64, 8) QSM_G(64, 16) QSM_G(32, 8) QSM_G(32, 16) QSM_G(32, 32)
#undef QSM_G
    set_error("unsupported tile shape");
    return QSM_EARG;
}

int drop_graph(qsm_batch *b)
{
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    b->graph_chunk = 0;
    return QSM_OK;
}

// Rebuilds the tile list from the live slots and uploads it with the
// descriptors (the stream is idle when this runs).
template <typename T>
int upload(qsm_batch *b)
{
    // TV volumes' tiles first, then TGV volumes' (one kernel per kind);
    // tile0 is relative to the kind's part of the list
    std::vector<int> map;
    int nk[2] = {0, 0};
    // the two-iteration march claims tiles in list order: costly tiles
    // (long marches of large planes) first, so cheap ones fill the end
    std::vector<int> order;
    for (int v = 0; v < b->max_vols; ++v)
        order.push_back(v);
    if (b->two)
        std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
            const VolDesc<T> &da = desc<T>(b, a), &dc = desc<T>(b, c);
            return double(std::min(da.kz, std::max(da.nz - 2, 1)) + 6)
                       > double(std::min(dc.kz, std::max(dc.nz - 2, 1)) + 6);
        });
    for (int kind = 0; kind < 2; ++kind) {
        const size_t start = map.size();
        for (int v : order) {
            if (!b->vols[v].used || (b->vols[v].kind == QSM_TGV) != (kind == 1)) continue;
            b->vols[v].uploaded = true;
            VolDesc<T> &d = desc<T>(b, v);
            d.tile0 = int(map.size() - start);
            for (int t = 0; t < b->vols[v].ntiles; ++t)
                map.push_back(v);
        }
        nk[kind] = int(map.size() - start);
    }
    if (int(map.size()) > b->tile_cap) {
        if (b->d_tilevol) (void)hipFree(b->d_tilevol);
        b->d_tilevol = nullptr;
        b->tile_cap = std::max<int>(int(map.size()), 2 * b->tile_cap);
        QSM_HIP(hipMalloc(reinterpret_cast<void **>(&b->d_tilevol), size_t(b->tile_cap) * sizeof(int)));
        drop_graph(b);  // the graph captured the old pointer
    }
    if (!map.empty())
        QSM_HIP(hipMemcpy(b->d_tilevol, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
    TileList tl{};
    tl.ntiles[0] = nk[0];
    tl.ntiles[1] = nk[1];
    QSM_HIP(hipMemcpy(b->d_tl, &tl, sizeof tl, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(b->d_vols, b->h_desc.data(), b->h_desc.size(), hipMemcpyHostToDevice));
    b->dirty = false;
    return QSM_OK;
}

// Pulls the device-advanced fields (iter, live) into the host mirror.
template <typename T>
int pull(qsm_batch *b)
{
    std::vector<unsigned char> tmp(b->h_desc.size());
    QSM_HIP(hipMemcpy(tmp.data(), b->d_vols, tmp.size(), hipMemcpyDeviceToHost));
    for (int v = 0; v < b->max_vols; ++v) {
        // a volume added since the last upload is not on the device yet:
        // its host descriptor is the current one
        if (!b->vols[v].used || !b->vols[v].uploaded)
            continue;
        const VolDesc<T> &s = reinterpret_cast<const VolDesc<T> *>(tmp.data())[v];
        VolDesc<T> &d = desc<T>(b, v);
        d.iter = s.iter;
        d.live = s.live;
        d.cur = s.cur;
        d.mode = s.mode;
        d.chunk_end = s.chunk_end;
    }
    return QSM_OK;
}

template <typename T>
int add_t(qsm_batch *b, int kind, const size_t sz[3], const double h[3], double lam,
          double alpha, const void *d2f, const uint8_t *mask, int on_device,
          uint32_t maxiter, double tol, int kz, int *slot)
{
    int v = -1;
    for (int i = 0; i < b->max_vols; ++i)
        if (!b->vols[i].used) { v = i; break; }
    if (v < 0) {
        set_error("batch has no free slot (%d volumes)", b->max_vols);
        return QSM_EFULL;
    }
    const size_t N = sz[0] * sz[1] * sz[2];
    QSM_ARG(fused_size_supported(kind, b->dtype, N), "volume too large for the batched pass");
    const Geom g{int(sz[0]), int(sz[1]), int(sz[2])};
    const FusedShape sh{b->tx, b->ty, kz};
    int ttx = b->tx, tty = b->ty;
    if (b->two)
        fused2_tile(b->dtype, ttx, tty);
    const int ntiles = b->two ? fused2_blocks(g, b->dtype, kz) : fused_blocks(g, sh);
    const int narr = kind == QSM_TGV ? 15 : 6;
    const size_t bank = align_up(size_t(narr) * N * sizeof(T), 256);
    const int hist_cap = int(maxiter / 10) + 1;
    const size_t off_f = 2 * bank;
    const size_t off_m = off_f + align_up(N * sizeof(T), 256);
    const size_t off_c = off_m + align_up(N, 256);
    const size_t off_p = off_c + 256;
    const size_t off_h = off_p + align_up(size_t(std::max(ntiles, 1)) * 4 * sizeof(double), 256);
    const size_t off_i = off_h + align_up(size_t(hist_cap) * 2 * sizeof(double), 256);
    const size_t bytes = off_i + (b->two ? align_up(size_t(std::max(ntiles, 1)) * sizeof(F2Item), 256) : 0);
    if (b->used + bytes > b->budget) {
        set_error("admission refused: %zu bytes needed, %zu of %zu in use", bytes, b->used, b->budget);
        return QSM_EFULL;
    }
    // the device owns iter/live of running volumes: refresh the host mirror
    // before it is edited and uploaded again
    QSM_HIP(hipStreamSynchronize(b->stream));
    int prc = pull<T>(b);
    if (prc) return prc;
    char *mem = nullptr;
    QSM_HIP(hipMalloc(reinterpret_cast<void **>(&mem), bytes));
    BatchVol &bv = b->vols[v];
    bv.used = true;
    bv.uploaded = false;
    bv.kind = kind;
    for (int a = 0; a < 3; ++a) bv.sz[a] = sz[a];
    bv.N = N;
    bv.bytes = bytes;
    bv.mem = mem;
    bv.ntiles = ntiles;
    bv.hist_cap = hist_cap;
    b->used += bytes;

    T *b0 = reinterpret_cast<T *>(mem), *b1 = reinterpret_cast<T *>(mem + bank);
    T *f = reinterpret_cast<T *>(mem + off_f);
    uint8_t *mk = reinterpret_cast<uint8_t *>(mem + off_m);
    Ctl *ctl = reinterpret_cast<Ctl *>(mem + off_c);
    const hipMemcpyKind kd = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    // load (as qsm_pd_load): mask -> packed mask via bank 1 staging, zero the
    // banks, f = sigma * d2f, reset the control block
    double tau, sigma;
    qsm_step_sizes(h, &tau, &sigma);
    QSM_HIP(hipMemcpy(b1, mask, N, kd));
    QSM_HIP(launch_mask_init(mk, reinterpret_cast<const uint8_t *>(b1), g, 0, b->stream));
    QSM_HIP(hipMemsetAsync(mem, 0, 2 * bank, b->stream));
    QSM_HIP(hipMemcpyAsync(f, d2f, N * sizeof(T), kd, b->stream));
    QSM_HIP(launch_scale<T>(f, f, T(sigma), N, b->stream));
    QSM_HIP(launch_ctl_reset(ctl, 1, b->stream));
    // dual of iteration 1 from the zero primal iterates (bank 1)
    const DualCoef<T> dc = make_dual_coef<T>(h, sigma, kind == QSM_TGV ? alpha * lam : lam, lam);
    const PrimalCoef<T> pc = make_primal_coef<T>(h, tau);
    Bank<T> k0{}, k1{};
    const size_t nn = N;
    auto mkbank = [&](T *a, Bank<T> &bk) {
        if (kind == QSM_TGV) {
            bk.chi = a; bk.eta = a + nn; bk.w1 = a + 2 * nn; bk.w2 = a + 3 * nn; bk.w3 = a + 4 * nn;
            bk.nu = a + 5 * nn; bk.p1 = a + 6 * nn; bk.p2 = a + 7 * nn; bk.p3 = a + 8 * nn;
            bk.q1 = a + 9 * nn; bk.q2 = a + 10 * nn; bk.q3 = a + 11 * nn;
            bk.q4 = a + 12 * nn; bk.q5 = a + 13 * nn; bk.q6 = a + 14 * nn;
        } else {
            bk.chi = a; bk.eta = a + nn; bk.nu = a + 2 * nn;
            bk.p1 = a + 3 * nn; bk.p2 = a + 4 * nn; bk.p3 = a + 5 * nn;
        }
    };
    mkbank(b0, k0);
    mkbank(b1, k1);
    const Bars<T> bars{k1.chi, k1.eta, k1.w1, k1.w2, k1.w3};
    if (kind == QSM_TGV)
        QSM_HIP(launch_dual_sweep<T, true>(k0, bars, f, mk, 2, g, dc, nullptr, b->stream));
    else
        QSM_HIP(launch_dual_sweep<T, false>(k0, bars, f, mk, 2, g, dc, nullptr, b->stream));
    const int ntx = (g.nx + ttx - 1) / ttx, nty = (g.ny + tty - 1) / tty;
    const int nch = (g.nz - 2 + kz - 1) / kz;
    // Two-iteration march: each (tile, chunk) of the uniform list trimmed to
    // the planes where the tile holds active voxels, empty ones dropped.  A
    // tile's voxels outside those planes are never stored by the march, so
    // the iterates are the same; the list keeps its order (chunk, row,
    // column), so neighbours still march the same planes together.
    F2Item *items = b->two ? reinterpret_cast<F2Item *>(mem + off_i) : nullptr;
    int nitems = 0;
    if (items && ntiles > 0) {
        int *dr = reinterpret_cast<int *>(mem + off_p);  // 2 ntx nty ints fit the partials
        QSM_HIP(launch_tile_zrange(mk, g, b->dtype, dr, b->stream));
        std::vector<int> r(2 * size_t(ntx) * nty);
        QSM_HIP(hipMemcpyAsync(r.data(), dr, r.size() * sizeof(int), hipMemcpyDeviceToHost,
                               b->stream));
        QSM_HIP(hipStreamSynchronize(b->stream));
        std::vector<F2Item> it;
        for (int j = 0; j < nch; ++j) {
            const int c0 = 1 + j * kz, c1 = std::min(c0 + kz, g.nz - 1);
            for (int by = 0; by < nty; ++by)
                for (int bx = 0; bx < ntx; ++bx) {
                    const int t = bx + ntx * by;
                    const int a = std::max(c0, r[2 * t]), e = std::min(c1, r[2 * t + 1]);
                    if (a < e)
                        it.push_back(F2Item{bx, by, a, e});
                }
        }
        if (!it.empty() && int(it.size()) <= ntiles) {
            QSM_HIP(hipMemcpyAsync(items, it.data(), it.size() * sizeof(F2Item),
                                   hipMemcpyHostToDevice, b->stream));
            nitems = int(it.size());
        }
    }
    QSM_HIP(hipStreamSynchronize(b->stream));
    if (nitems)
        bv.ntiles = nitems;

    VolDesc<T> &d = desc<T>(b, v);
    std::memset(&d, 0, sizeof d);
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz; d.kind = kind;
    d.N = uint32_t(N);
    d.ntx = ntx; d.nty = nty; d.nch = ntiles ? nch : 0; d.kz = kz;
    d.nitems = nitems;
    d.items = nitems ? items : nullptr;
    d.bank[0] = b0; d.bank[1] = b1;
    d.f = f; d.mk = mk; d.ctl = ctl;
    d.partials = reinterpret_cast<double *>(mem + off_p);
    d.hist = reinterpret_cast<double *>(mem + off_h);
    d.hist_cap = hist_cap;
    d.target = maxiter;
    d.tol = tol;
    d.dc = dc;
    d.pc = pc;
    d.iter = 0;
    d.cur = 0;
    d.mode = 0;
    d.chunk_end = 0;
    d.live = (maxiter > 0 && ntiles > 0) ? 1 : 0;
    if (ntiles == 0)
        d.ntx = d.nty = d.nch = 0;
    b->dirty = true;
    *slot = v;
    return QSM_OK;
}

template <typename T>
int result_t(qsm_batch *b, int v, void *chi, uint32_t *iters, int *stopped, double nr[2])
{
    QSM_HIP(hipStreamSynchronize(b->stream));
    int rc = pull<T>(b);
    if (rc) return rc;
    VolDesc<T> &d = desc<T>(b, v);
    Ctl c;
    QSM_HIP(hipMemcpy(&c, d.ctl, sizeof c, hipMemcpyDeviceToHost));
    const uint32_t it = c.stop ? c.stop_iter : d.iter;
    const uint32_t bank = c.stop ? c.stop_bank : d.cur;
    if (iters) *iters = it;
    if (stopped) *stopped = c.stop;
    if (nr) { nr[0] = c.nr[0]; nr[1] = c.nr[1]; }
    if (chi)
        QSM_HIP(hipMemcpy(chi, d.bank[bank & 1u], b->vols[v].N * sizeof(T),
                          hipMemcpyDeviceToHost));
    return QSM_OK;
}

template <typename T>
int step_t(qsm_batch *b, uint32_t chunk)
{
    if (b->dirty) {
        QSM_HIP(hipStreamSynchronize(b->stream));
        int rc = upload<T>(b);
        if (rc) return rc;
    }
    if (!b->use_graph) {
        QSM_HIP(launch_chunk<T>(b, chunk));
        return QSM_OK;
    }
    if (!b->gexec || b->graph_chunk != chunk) {
        drop_graph(b);
        QSM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = launch_chunk<T>(b, chunk);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(b->stream, &g);
        QSM_HIP(le);
        QSM_HIP(ee);
        b->graph = g;
        QSM_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
        b->graph_chunk = chunk;
    }
    QSM_HIP(hipGraphLaunch(b->gexec, b->stream));
    return QSM_OK;
}

}  // namespace

extern "C" {

int qsm_batch_create(qsm_batch **out, int dtype, int tx, int ty, size_t hbm_budget,
                     int max_volumes, void *stream)
{
    QSM_ARG(out, "null handle pointer");
    *out = nullptr;
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    QSM_ARG((tx == 0 && ty == 0) || fused_shape_supported(tx, ty),
            "unsupported tile shape %dx%d", tx, ty);
    QSM_ARG(max_volumes >= 1 && max_volumes <= 4096, "max_volumes out of range");
    int rc = ensure_device();
    if (rc) return rc;
    qsm_batch *b = new (std::nothrow) qsm_batch();
    QSM_ARG(b, "host allocation failed");
    b->dtype = dtype;
    b->tx = tx;
    b->ty = ty;
    b->two = tx == 0 && ty == 0;
    b->budget = hbm_budget;
    b->used = 0;
    b->max_vols = max_volumes;
    b->stream = as_stream(stream);
    if (!stream) {
        // graph capture needs a stream of its own (not the legacy null
        // stream); a blocking one still orders after null-stream copies
        const hipError_t se = hipStreamCreate(&b->stream);
        if (se != hipSuccess) {
            delete b;
            return hip_fail(se, "hipStreamCreate", __FILE__, __LINE__);
        }
        b->own_stream = true;
    }
    b->vols.resize(size_t(max_volumes));
    b->desc_size = dtype == QSM_F64 ? sizeof(VolDesc<double>) : sizeof(VolDesc<float>);
    b->h_desc.assign(b->desc_size * size_t(max_volumes), 0);
    hipError_t e = hipMalloc(&b->d_vols, b->h_desc.size());
    if (e == hipSuccess) e = hipMalloc(&b->d_tl, 256);
    if (e == hipSuccess) e = hipMemset(b->d_vols, 0, b->h_desc.size());
    if (e != hipSuccess) {
        qsm_batch_destroy(b);
        return hip_fail(e, "batch allocation", __FILE__, __LINE__);
    }
    if ((rc = grid_for(b))) {
        qsm_batch_destroy(b);
        return rc;
    }
    *out = b;
    return QSM_OK;
}

size_t qsm_batch_volume_bytes(int kind, int dtype, const size_t sz[3], int tx, int ty,
                              int kz, uint32_t maxiter)
{
    if (!sz || kz < 1) return 0;
    const size_t N = sz[0] * sz[1] * sz[2];
    const size_t s = dtype_size(dtype);
    const Geom g{int(sz[0]), int(sz[1]), int(sz[2])};
    const int ntiles = (tx == 0 && ty == 0) ? fused2_blocks(g, dtype, kz)
                                            : fused_blocks(g, FusedShape{tx, ty, kz});
    const size_t bank = align_up(size_t(kind == QSM_TGV ? 15 : 6) * N * s, 256);
    const bool two = tx == 0 && ty == 0;
    return 2 * bank + align_up(N * s, 256) + align_up(N, 256) + 256
           + align_up(size_t(std::max(ntiles, 1)) * 32, 256)
           + align_up(size_t(maxiter / 10 + 1) * 16, 256)
           + (two ? align_up(size_t(std::max(ntiles, 1)) * sizeof(F2Item), 256) : 0);
}

int qsm_batch_add(qsm_batch *b, int kind, const size_t sz[3], const double h[3],
                  double lam, double alpha, const void *d2f, const uint8_t *mask,
                  int on_device, uint32_t maxiter, double tol, int kz, int *slot)
{
    QSM_ARG(b && d2f && mask && slot, "null pointer");
    QSM_ARG(kind == QSM_TV || kind == QSM_TGV, "bad kind");
    QSM_ARG(kz >= 1, "kz must be >= 1");
    int rc = check_grid(sz, h, 1);
    if (rc) return rc;
    return b->dtype == QSM_F64
        ? add_t<double>(b, kind, sz, h, lam, alpha, d2f, mask, on_device, maxiter, tol, kz, slot)
        : add_t<float>(b, kind, sz, h, lam, alpha, d2f, mask, on_device, maxiter, tol, kz, slot);
}

int qsm_batch_step(qsm_batch *b, uint32_t chunk)
{
    QSM_ARG(b, "null handle");
    if (chunk == 0) return QSM_OK;
    return b->dtype == QSM_F64 ? step_t<double>(b, chunk) : step_t<float>(b, chunk);
}

int qsm_batch_sync(qsm_batch *b)
{
    QSM_ARG(b, "null handle");
    QSM_HIP(hipStreamSynchronize(b->stream));
    return b->dtype == QSM_F64 ? pull<double>(b) : pull<float>(b);
}

int qsm_batch_state(qsm_batch *b, int slot, int *used, int *live, uint32_t *iter)
{
    QSM_ARG(b && slot >= 0 && slot < b->max_vols, "bad slot");
    if (used) *used = b->vols[slot].used;
    if (b->dtype == QSM_F64) {
        const VolDesc<double> &d = desc<double>(b, slot);
        if (live) *live = b->vols[slot].used ? d.live : 0;
        if (iter) *iter = d.iter;
    } else {
        const VolDesc<float> &d = desc<float>(b, slot);
        if (live) *live = b->vols[slot].used ? d.live : 0;
        if (iter) *iter = d.iter;
    }
    return QSM_OK;
}

int qsm_batch_result(qsm_batch *b, int slot, void *chi, uint32_t *iters, int *stopped,
                     double nr[2])
{
    QSM_ARG(b && slot >= 0 && slot < b->max_vols && b->vols[slot].used, "bad slot");
    return b->dtype == QSM_F64 ? result_t<double>(b, slot, chi, iters, stopped, nr)
                               : result_t<float>(b, slot, chi, iters, stopped, nr);
}

int qsm_batch_remove(qsm_batch *b, int slot)
{
    QSM_ARG(b && slot >= 0 && slot < b->max_vols && b->vols[slot].used, "bad slot");
    QSM_HIP(hipStreamSynchronize(b->stream));
    const int prc = b->dtype == QSM_F64 ? pull<double>(b) : pull<float>(b);
    if (prc) return prc;
    BatchVol &bv = b->vols[slot];
    QSM_HIP(hipFree(bv.mem));
    b->used -= bv.bytes;
    bv = BatchVol();
    std::memset(b->h_desc.data() + b->desc_size * size_t(slot), 0, b->desc_size);
    b->dirty = true;
    return QSM_OK;
}

int qsm_batch_set_graph(qsm_batch *b, int use_graph)
{
    QSM_ARG(b, "null handle");
    b->use_graph = use_graph != 0;
    if (!b->use_graph) drop_graph(b);
    return QSM_OK;
}

size_t qsm_batch_used_bytes(qsm_batch *b) { return b ? b->used : 0; }

void qsm_batch_destroy(qsm_batch *b)
{
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (auto &v : b->vols)
        if (v.used && v.mem) (void)hipFree(v.mem);
    if (b->d_vols) (void)hipFree(b->d_vols);
    if (b->d_tilevol) (void)hipFree(b->d_tilevol);
    if (b->d_tl) (void)hipFree(b->d_tl);
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
