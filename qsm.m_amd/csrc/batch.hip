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
// the single-volume solver (same tiles, same reduction order), so a result
// is bitwise the same alone and in any mix.
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "fused_dev.h"

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
};

struct TileList {
    int ntiles[2];   // tiles of TV volumes, of TGV volumes (list order)
    int pad[2];
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
        if (!dp->live || dp->ctl->stop)
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
            }
        }
    }
    if (threadIdx.x == 0) {
        d.iter = n;
        if (n >= d.target || ctl->stop)
            d.live = 0;
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

int grid_for(qsm_batch *b)
{
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
    for (int kind = 0; kind < 2; ++kind) {
        const size_t start = map.size();
        for (int v = 0; v < b->max_vols; ++v) {
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
    TileList tl{{nk[0], nk[1]}, {0, 0}};
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
    const int ntiles = fused_blocks(g, sh);
    const int narr = kind == QSM_TGV ? 15 : 6;
    const size_t bank = align_up(size_t(narr) * N * sizeof(T), 256);
    const int hist_cap = int(maxiter / 10) + 1;
    const size_t off_f = 2 * bank;
    const size_t off_m = off_f + align_up(N * sizeof(T), 256);
    const size_t off_c = off_m + align_up(N, 256);
    const size_t off_p = off_c + 256;
    const size_t off_h = off_p + align_up(size_t(std::max(ntiles, 1)) * 4 * sizeof(double), 256);
    const size_t bytes = off_h + align_up(size_t(hist_cap) * 2 * sizeof(double), 256);
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
    QSM_HIP(hipStreamSynchronize(b->stream));

    VolDesc<T> &d = desc<T>(b, v);
    std::memset(&d, 0, sizeof d);
    d.nx = g.nx; d.ny = g.ny; d.nz = g.nz; d.kind = kind;
    d.N = uint32_t(N);
    int ntx, nty, nch;
    ntx = (g.nx + b->tx - 1) / b->tx;
    nty = (g.ny + b->ty - 1) / b->ty;
    nch = (g.nz - 2 + kz - 1) / kz;
    d.ntx = ntx; d.nty = nty; d.nch = ntiles ? nch : 0; d.kz = kz;
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
    if (iters) *iters = it;
    if (stopped) *stopped = c.stop;
    if (nr) { nr[0] = c.nr[0]; nr[1] = c.nr[1]; }
    if (chi)
        QSM_HIP(hipMemcpy(chi, d.bank[it & 1u], b->vols[v].N * sizeof(T), hipMemcpyDeviceToHost));
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
        for (uint32_t i = 0; i < chunk; ++i)
            QSM_HIP(launch_iteration<T>(b));
        return QSM_OK;
    }
    if (!b->gexec || b->graph_chunk != chunk) {
        drop_graph(b);
        QSM_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t le = hipSuccess;
        for (uint32_t i = 0; i < chunk && le == hipSuccess; ++i)
            le = launch_iteration<T>(b);
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
    QSM_ARG(fused_shape_supported(tx, ty), "unsupported tile shape %dx%d", tx, ty);
    QSM_ARG(max_volumes >= 1 && max_volumes <= 4096, "max_volumes out of range");
    int rc = ensure_device();
    if (rc) return rc;
    qsm_batch *b = new (std::nothrow) qsm_batch();
    QSM_ARG(b, "host allocation failed");
    b->dtype = dtype;
    b->tx = tx;
    b->ty = ty;
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
    const size_t N = sz[0] * sz[1] * sz[2];
    const size_t s = dtype_size(dtype);
    const int ntiles = fused_blocks(Geom{int(sz[0]), int(sz[1]), int(sz[2])}, FusedShape{tx, ty, kz});
    const size_t bank = align_up(size_t(kind == QSM_TGV ? 15 : 6) * N * s, 256);
    return 2 * bank + align_up(N * s, 256) + align_up(N, 256) + 256
           + align_up(size_t(std::max(ntiles, 1)) * 32, 256)
           + align_up(size_t(maxiter / 10 + 1) * 16, 256);
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
