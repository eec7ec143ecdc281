// pd_api.hip -- C ABI of the primal-dual solver: resident handles, the
// drop-in whole-solve entry points and the single-sweep device entry points.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <new>
#include <vector>

#include "pd_kernels.h"

using namespace qsm;

// The allocation-free solve's device work space (qsm_pd_solve_static): one
// copy per GPU, its address looked up per device with hipGetSymbolAddress.
static __device__ __align__(256) char scratch[QSM_PD_STATIC_BYTES];

namespace {

constexpr int NB_TGV = 15;  // chi, eta, w[3], nu, p[3], q[6]
constexpr int NB_TV = 6;    // chi, eta, nu, p[3]

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct Layout {
    size_t bank_bytes, f_off, mask_off, ctl_off, part_off, hist_off, items_off, total;
};

// items: room for a work list of the two-iteration pass (npart items; the
// allocation-free solve has none)
Layout layout(int kind, int dtype, const size_t sz[3], int npart, int hist_cap,
              bool items = true)
{
    const size_t N = sz[0] * sz[1] * sz[2];
    const size_t s = dtype_size(dtype);
    const int nb = kind == QSM_TGV ? NB_TGV : NB_TV;
    Layout L;
    L.bank_bytes = align_up(size_t(nb) * N * s, 256);
    L.f_off = 2 * L.bank_bytes;
    L.mask_off = L.f_off + align_up(N * s, 256);
    L.ctl_off = L.mask_off + align_up(N, 256);
    L.part_off = L.ctl_off + 256;
    L.hist_off = L.part_off + align_up(size_t(npart) * 4 * sizeof(double), 256);
    L.items_off = L.hist_off + align_up(size_t(hist_cap) * 2 * sizeof(double), 256);
    L.total = L.items_off + (items ? align_up(size_t(npart) * sizeof(F2Item), 256) : 0);
    return L;
}

// Partial-sum slots that cover every supported fused shape (at kz = 1, the
// most blocks a shape can have), the two-iteration pass and the two-sweep
// primal.
int max_partials(Geom g)
{
    static const int shapes[][2] = {{64, 4}, {64, 8}, {64, 16}, {32, 8}, {32, 16}, {32, 32}};
    int n = std::max(primal_sweep_blocks(g), 1);
    for (const auto &sh : shapes)
        n = std::max(n, fused_blocks(g, FusedShape{sh[0], sh[1], 1}));
    n = std::max(n, fused2_blocks(g, QSM_F32, 1));
    n = std::max(n, fused2_blocks(g, QSM_F64, 1));
    return n;
}

int cu_count()
{
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) != hipSuccess
        || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess
        || ncu < 1)
        ncu = 256;
    return ncu;
}

// z-chunk depth of the two-iteration pass: as for the one-iteration pass,
// at most one block per CU (each block is resident for its whole march).
int default_kz2(int dtype, Geom g)
{
    int tx, ty;
    fused2_tile(dtype, tx, ty);
    const int inner = std::max(1, g.nz - 2);
    const int nxy = ((g.nx + tx - 1) / tx) * ((g.ny + ty - 1) / ty);
    const int nch = std::max(1, cu_count() / nxy);
    return std::max(1, (inner + nch - 1) / nch);
}

// Default fused tile shape (the autotuner's table may pick others): 64x16
// tiles (64x8 for double TGV, whose registers allow fewer waves per CU),
// and z-chunks sized so that the grid is at most one block per CU -- the
// march keeps every block resident, and a partial second wave of blocks
// costs a whole chunk.  On MI355X this reproduces the tuned z-chunks of
// 96^3 to 256x256x176 (e.g. FP32 64x16x44, FP64 TGV 64x8x87 at
// 256x256x176).
FusedShape default_shape(int kind, int dtype, Geom g)
{
    FusedShape sh{64, (dtype == QSM_F64 && kind == QSM_TGV) ? 8 : 16, 16};
    const int ncu = cu_count();
    const int inner = std::max(1, g.nz - 2);
    const int nxy = ((g.nx + sh.tx - 1) / sh.tx) * ((g.ny + sh.ty - 1) / sh.ty);
    const int nch = std::max(1, ncu / nxy);  // at most one block per CU
    sh.kz = std::max(1, (inner + nch - 1) / nch);
    return sh;
}

}  // namespace

struct qsm_pd {
    // schedule 0: two-iteration passes, one-iteration passes where a
    // convergence check or the call's count needs them; 1: the reference's
    // two sweeps; 2: one-iteration passes only
    int kind, dtype, schedule;
    size_t sz[3], N;
    double h[3], lam, alpha, tau, sigma;
    hipStream_t stream;
    char *mem;
    Layout L;
    int npart, hist_cap;
    uint32_t n;       // iterations enqueued since the last load
    int dual_ahead;   // fused schedule: dual of iteration n+1 applied
    int loaded;
    FusedShape shape;
    int kz2;          // z-chunk depth of the two-iteration pass
    uint32_t cur;     // fused schedules: bank holding the latest iterates
    // kernel launches enqueued since creation: two-iteration passes,
    // one-iteration passes, two-sweep iterations, convergence checks
    uint64_t launches[4];
    // work list of the two-iteration pass (built at load when it beats the
    // uniform grid): nitems items of at most items_kz planes; 0 = uniform
    int items_on, nitems, items_kz;

    void *bank(int b) const { return mem + size_t(b) * L.bank_bytes; }
    void *f() const { return mem + L.f_off; }
    uint8_t *mask() const { return reinterpret_cast<uint8_t *>(mem + L.mask_off); }
    Ctl *ctl() const { return reinterpret_cast<Ctl *>(mem + L.ctl_off); }
    double *partials() const { return reinterpret_cast<double *>(mem + L.part_off); }
    double *hist() const { return reinterpret_cast<double *>(mem + L.hist_off); }
    F2Item *items() const { return reinterpret_cast<F2Item *>(mem + L.items_off); }
    Geom geom() const { return Geom{int(sz[0]), int(sz[1]), int(sz[2])}; }
};

namespace {

// Work list of the two-iteration pass for a masked volume.  A tile's
// owned voxels outside its active plane range [lo, hi) (no mask bit) are
// never stored by the pass, so marching only [lo, hi) gives the same
// iterates, and any split of it into chunks (each recomputing its own halo,
// like the uniform grid's chunks) too.  The ranges are cut on one grid of
// z-chunks shared by all tiles, chunk j = planes [1 - o + j kz, 1 - o +
// (j + 1) kz): neighbouring tiles' items then march the same planes at the
// same time and their ring reads meet in L2 (ranges split per tile instead
// start at different planes: 50 % more HBM reads, a third of the gain).
// Items are ordered by (chunk, tile row, tile column), so consecutive items
// -- one XCD's blocks -- are patches of neighbours.  An item of len planes
// costs about len + 6 plane steps (the four-stage pipeline's fill) and one
// item per CU runs at a time; the grid (kz >= 8 or the interior, offset o,
// offsets sampled at kz/16 steps) taken is the one with the fewest rounds,
// then the shortest longest item.  The list is built only where it skips
// at least a tenth of the tiles' plane steps, and kept when its rounds x
// (len + 6) beat the uniform grid's by 10 % (items_on == 2: whenever a
// voxel is active).  Synchronous (the item count sizes the launches).
int build_items(qsm_pd *s)
{
    s->nitems = 0;
    s->items_kz = 0;
    const Geom g = s->geom();
    if (!s->items_on || g.nx < 3 || g.ny < 3 || g.nz < 3 || !fused_size_supported(s->kind, s->dtype, s->N))
        return QSM_OK;
    int tx, ty;
    fused2_tile(s->dtype, tx, ty);
    const int ntx = (g.nx + tx - 1) / tx, nty = (g.ny + ty - 1) / ty, nt = ntx * nty;
    int *dr = reinterpret_cast<int *>(s->items());  // 2 nt ints <= npart items
    QSM_HIP(launch_tile_zrange(s->mask(), g, s->dtype, dr, s->stream));
    std::vector<int> r(2 * size_t(nt));
    QSM_HIP(hipMemcpyAsync(r.data(), dr, r.size() * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    QSM_HIP(hipStreamSynchronize(s->stream));
    // A list only pays where it skips planes: with (nearly) every plane of
    // every tile active the uniform grid's tuned chunking is kept.
    long act_planes = 0;
    for (int t = 0; t < nt; ++t)
        act_planes += std::max(0, r[2 * t + 1] - r[2 * t]);
    if (s->items_on == 1 && act_planes * 10 >= long(nt) * (g.nz - 2) * 9)
        return QSM_OK;
    // rounds of one block per CU: a second resident block (TV's smaller LDS
    // allows three) shares the CU's issue slots rather than adding to them
    const long slots = long(cu_count());
    const long nb = fused2_blocks(g, s->dtype, s->kz2);
    const long cost_u = (nb + slots - 1) / slots * (s->kz2 + 6);
    long act = 0;
    for (int t = 0; t < nt; ++t)
        act += r[2 * t + 1] > r[2 * t];
    if (act == 0)
        return QSM_OK;  // no active voxel: the uniform grid
    const long rmin = (act + slots - 1) / slots;  // every active tile needs an item
    long best = -1;
    int bkz = 0, bo = 0;
    for (int kz = std::min(8, g.nz - 2); kz <= g.nz - 2; ++kz)
        for (int o = 0; o < kz; o += std::max(1, kz / 16)) {  // 16 offsets at most
            long cnt = 0;
            int len = 0;
            for (int t = 0; t < nt; ++t) {
                const int lo = r[2 * t], hi = r[2 * t + 1];
                if (hi <= lo) continue;
                const int j0 = (lo - 1 + o) / kz, j1 = (hi - 2 + o) / kz;
                cnt += j1 - j0 + 1;
                for (int j = j0; j <= j1; ++j)
                    len = std::max(len, std::min(hi, 1 - o + (j + 1) * kz)
                                            - std::max(lo, 1 - o + j * kz));
            }
            const long rounds = (cnt + slots - 1) / slots;
            if (rounds > rmin) continue;
            const long cost = rounds * (len + 6);
            if (best < 0 || cost < best) { best = cost; bkz = kz; bo = o; }
        }
    if (bkz == 0 || (s->items_on == 1 && best * 10 >= cost_u * 9))
        return QSM_OK;
    std::vector<F2Item> it;
    for (int j = 0; 1 - bo + j * bkz < g.nz - 1; ++j) {
        const int c0 = 1 - bo + j * bkz, c1 = c0 + bkz;
        for (int by = 0; by < nty; ++by)
            for (int bx = 0; bx < ntx; ++bx) {
                const int t = bx + ntx * by;
                const int k0 = std::max(c0, r[2 * t]), k1 = std::min(c1, r[2 * t + 1]);
                if (k0 < k1)
                    it.push_back(F2Item{bx, by, k0, k1});
            }
    }
    if (it.size() > size_t(s->npart))
        return QSM_OK;
    QSM_HIP(hipMemcpyAsync(s->items(), it.data(), it.size() * sizeof(F2Item),
                           hipMemcpyHostToDevice, s->stream));
    QSM_HIP(hipStreamSynchronize(s->stream));
    s->nitems = int(it.size());
    s->items_kz = bkz;
    return QSM_OK;
}

template <typename T>
Bank<T> make_bank(void *base, int kind, size_t N)
{
    T *a = static_cast<T *>(base);
    Bank<T> b{};
    if (kind == QSM_TGV) {
        b.chi = a; b.eta = a + N;
        b.w1 = a + 2 * N; b.w2 = a + 3 * N; b.w3 = a + 4 * N;
        b.nu = a + 5 * N;
        b.p1 = a + 6 * N; b.p2 = a + 7 * N; b.p3 = a + 8 * N;
        b.q1 = a + 9 * N; b.q2 = a + 10 * N; b.q3 = a + 11 * N;
        b.q4 = a + 12 * N; b.q5 = a + 13 * N; b.q6 = a + 14 * N;
    } else {
        b.chi = a; b.eta = a + N;
        b.nu = a + 2 * N;
        b.p1 = a + 3 * N; b.p2 = a + 4 * N; b.p3 = a + 5 * N;
    }
    return b;
}

// The two-sweep schedule keeps its extrapolated arrays in bank 1.
template <typename T>
Bars<T> bars_of(const Bank<T> &b1)
{
    return Bars<T>{b1.chi, b1.eta, b1.w1, b1.w2, b1.w3};
}

template <typename T>
int iterate_impl(qsm_pd *s, uint32_t niter, double tol)
{
    const Geom g = s->geom();
    const DualCoef<T> dc = make_dual_coef<T>(
        s->h, s->sigma, s->kind == QSM_TGV ? s->alpha * s->lam : s->lam, s->lam);
    const PrimalCoef<T> pc = make_primal_coef<T>(s->h, s->tau);
    const T *f = static_cast<const T *>(s->f());
    const Bank<T> b0 = make_bank<T>(s->bank(0), s->kind, s->N);
    const Bank<T> b1 = make_bank<T>(s->bank(1), s->kind, s->N);
    const int *stop = &s->ctl()->stop;
    const bool tgv = s->kind == QSM_TGV;
    QSM_ARG(s->schedule != 0 || fused_blocks(g, s->shape) <= s->npart,
            "fused tile shape needs %d partial-sum slots, %d allocated",
            fused_blocks(g, s->shape), s->npart);

    QSM_ARG(s->schedule != 0 || fused2_blocks(g, s->dtype, s->kz2) <= s->npart,
            "two-iteration pass needs %d partial-sum slots, %d allocated",
            fused2_blocks(g, s->dtype, s->kz2), s->npart);

    for (uint32_t t = 0; t < niter; ++t) {
        const uint32_t n1 = s->n + 1;
        const bool check = (n1 % 10) == 0;  // pd_tgv_mex.c:298,309
        double *part = check ? s->partials() : nullptr;
        int nblk = 0;
        if (s->schedule != 1 && !s->dual_ahead) {
            // Dual of iteration 1 from the zero primal iterates, which bank 1
            // holds until the first fused pass writes it.
            const hipError_t e = tgv
                ? launch_dual_sweep<T, true>(b0, bars_of(b1), f, s->mask(), 2, g, dc, stop, s->stream)
                : launch_dual_sweep<T, false>(b0, bars_of(b1), f, s->mask(), 2, g, dc, stop, s->stream);
            QSM_HIP(e);
            s->dual_ahead = 1;
        }
        if (s->schedule == 0 && t + 1 < niter && !check) {
            // iterations n1 and n1+1 in one pass; a check falls at most on
            // n1+1 (consecutive iterations), whose chi the pass stores
            const uint32_t n2 = n1 + 1;
            const bool check2 = (n2 % 10) == 0;
            double *part2 = check2 ? s->partials() : nullptr;
            const T *in = static_cast<const T *>(s->bank(s->cur));
            T *out = static_cast<T *>(s->bank(s->cur ^ 1u));
            const F2Item *items = s->nitems ? s->items() : nullptr;
            const hipError_t e = tgv
                ? launch_fused2<T, true>(in, out, s->N, f, s->mask(), g, s->kz2, dc, pc, stop, part2,
                                         items, s->nitems, s->stream)
                : launch_fused2<T, false>(in, out, s->N, f, s->mask(), g, s->kz2, dc, pc, stop, part2,
                                          items, s->nitems, s->stream);
            QSM_HIP(e);
            s->cur ^= 1u;
            ++s->launches[0];
            if (check2) {
                const int nb2 = s->nitems ? s->nitems : fused2_blocks(g, s->dtype, s->kz2);
                QSM_HIP(launch_conv_finalize(s->partials(), nb2, s->ctl(), s->hist(), s->hist_cap,
                                             tol, n2, s->cur, s->stream));
                ++s->launches[3];
            }
            s->n = n2;
            ++t;
            continue;
        }
        if (s->schedule != 1) {
            const T *in = static_cast<const T *>(s->bank(s->cur));
            T *out = static_cast<T *>(s->bank(s->cur ^ 1u));
            const hipError_t e = tgv
                ? launch_fused<T, true>(in, out, s->N, f, s->mask(), g, s->shape, dc, pc, stop, part, s->stream)
                : launch_fused<T, false>(in, out, s->N, f, s->mask(), g, s->shape, dc, pc, stop, part, s->stream);
            QSM_HIP(e);
            s->cur ^= 1u;
            ++s->launches[1];
            nblk = fused_blocks(g, s->shape);
        } else {
            hipError_t e = tgv
                ? launch_dual_sweep<T, true>(b0, bars_of(b1), f, s->mask(), 2, g, dc, stop, s->stream)
                : launch_dual_sweep<T, false>(b0, bars_of(b1), f, s->mask(), 2, g, dc, stop, s->stream);
            QSM_HIP(e);
            e = tgv
                ? launch_primal_sweep<T, true>(b0, bars_of(b1), s->mask(), 1, g, pc, stop, part, s->stream)
                : launch_primal_sweep<T, false>(b0, bars_of(b1), s->mask(), 1, g, pc, stop, part, s->stream);
            QSM_HIP(e);
            ++s->launches[2];
            nblk = primal_sweep_blocks(g);
        }
        if (check) {
            // An empty interior gives nblk == 0: all sums zero and 0/0 norms,
            // as the reference computes them.
            QSM_HIP(launch_conv_finalize(s->partials(), nblk, s->ctl(), s->hist(),
                                         s->hist_cap, tol, n1, s->cur, s->stream));
            ++s->launches[3];
        }
        s->n = n1;
    }
    return QSM_OK;
}

template <typename T>
int get_chi_impl(qsm_pd *s, void *chi, int on_device)
{
    const Bank<T> b0 = make_bank<T>(s->bank(0), s->kind, s->N);
    const Bank<T> b1 = make_bank<T>(s->bank(1), s->kind, s->N);
    if (on_device) {
        if (s->schedule != 1) {
            QSM_HIP(launch_select_copy<T>(static_cast<T *>(chi), b0.chi, b1.chi,
                                          s->ctl(), s->cur, s->N, s->stream));
        } else {
            QSM_HIP(hipMemcpyAsync(chi, b0.chi, s->N * sizeof(T),
                                   hipMemcpyDeviceToDevice, s->stream));
        }
        return QSM_OK;
    }
    const T *src = b0.chi;
    if (s->schedule != 1) {
        Ctl c;
        QSM_HIP(hipMemcpyAsync(&c, s->ctl(), sizeof c, hipMemcpyDeviceToHost, s->stream));
        QSM_HIP(hipStreamSynchronize(s->stream));
        const uint32_t b = c.stop ? c.stop_bank : s->cur;
        src = (b & 1u) ? b1.chi : b0.chi;
    }
    QSM_HIP(hipMemcpyAsync(chi, src, s->N * sizeof(T), hipMemcpyDeviceToHost,
                           s->stream));
    QSM_HIP(hipStreamSynchronize(s->stream));
    return QSM_OK;
}

int create_impl(qsm_pd **out, int kind, int dtype, const size_t sz[3],
                const double h[3], double lam, double alpha, void *stream,
                int hist_cap)
{
    QSM_ARG(out != nullptr, "null handle pointer");
    *out = nullptr;
    QSM_ARG(kind == QSM_TV || kind == QSM_TGV, "kind must be QSM_TV or QSM_TGV");
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "dtype must be QSM_F64 or QSM_F32");
    int rc = check_grid(sz, h, 1);
    if (rc) return rc;
    if ((rc = ensure_device())) return rc;

    qsm_pd *s = new (std::nothrow) qsm_pd();
    QSM_ARG(s != nullptr, "host allocation failed");
    s->kind = kind;
    s->dtype = dtype;
    s->schedule = 0;
    for (int a = 0; a < 3; ++a) { s->sz[a] = sz[a]; s->h[a] = h[a]; }
    s->N = sz[0] * sz[1] * sz[2];
    s->lam = lam;
    s->alpha = alpha;
    s->stream = as_stream(stream);
    s->shape = default_shape(kind, dtype, Geom{int(sz[0]), int(sz[1]), int(sz[2])});
    s->kz2 = default_kz2(dtype, Geom{int(sz[0]), int(sz[1]), int(sz[2])});
    s->cur = 0;
    // Banks above 4 GiB exceed the fused pass's 32-bit offsets: use the
    // two-sweep schedule (same results) for such volumes.
    s->schedule = fused_size_supported(kind, dtype, s->N) ? 0 : 1;
    qsm_step_sizes(h, &s->tau, &s->sigma);
    const Geom g = s->geom();
    s->npart = max_partials(g);
    s->hist_cap = hist_cap;
    s->items_on = 1;
    s->L = layout(kind, dtype, sz, s->npart, hist_cap);
    const hipError_t e = hipMalloc(reinterpret_cast<void **>(&s->mem), s->L.total);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "hipMalloc(solver state)", __FILE__, __LINE__);
    }
    *out = s;
    return QSM_OK;
}

void print_line(qsm_print_fn print, void *ctx, const char *line)
{
    if (print)
        print(line, ctx);
    else
        fputs(line, stdout);
}

// The reference's progress table (pd_tgv_mex.c:290-315), reproduced after
// the solve from the device's norm history.
void print_progress(qsm_pd *s, uint32_t maxiter, uint32_t verbose,
                    qsm_print_fn print, void *ctx)
{
    if (!verbose)
        return;
    uint32_t iters = 0;
    int stopped = 0;
    double nr[2];
    qsm_pd_status(s, &iters, &stopped, nr);
    std::vector<double> hist(size_t(s->hist_cap) * 2);
    const int nh = qsm_pd_norm_history(s, hist.data(), s->hist_cap);
    const uint32_t vb = verbose > 4 ? verbose : 4;
    const uint32_t pmod = maxiter / (vb < maxiter ? vb : maxiter);
    char line[128];
    print_line(print, ctx, "Iter\t\t||delta chi||\t||delta eta||\n");
    double n1 = 1.0, n2 = 1.0;
    for (uint32_t i = 0; i < iters; ++i) {
        const uint32_t it = i + 1;
        const bool flag = stopped && it == iters;
        if (it % 10 == 0) {
            const int c = int(it / 10) - 1;
            if (c < nh) { n1 = hist[2 * c]; n2 = hist[2 * c + 1]; }
        }
        if (i == 0 || i == maxiter - 1 || (it % pmod) == 0 || flag) {
            snprintf(line, sizeof line, "%5d/%d\t%.3e\t%.3e\n", it, maxiter, n1, n2);
            print_line(print, ctx, line);
        }
    }
}

int pd_solve(int kind, int dtype, const void *d2f, const uint8_t *mask,
             const size_t sz[3], const double h[3], double lam, double alpha,
             double tol, uint32_t maxiter, uint32_t verbose, void *chi,
             qsm_print_fn print, void *print_ctx)
{
    QSM_ARG(d2f && mask && chi, "null input or output pointer");
    QSM_ARG(maxiter > 0, "maxiter must be > 0");
    qsm_pd *s = nullptr;
    int rc = create_impl(&s, kind, dtype, sz, h, lam, alpha, nullptr,
                         int(maxiter / 10) + 1);
    if (rc) return rc;
    if (!(rc = qsm_pd_load(s, d2f, mask, 0))
        && !(rc = qsm_pd_iterate(s, maxiter, tol))
        && !(rc = qsm_pd_get_chi(s, chi, 0)))
        print_progress(s, maxiter, verbose, print, print_ctx);
    qsm_pd_destroy(s);
    return rc;
}

// ---- the allocation-free solve ------------------------------------------
//
// Every piece of one solve's device state (the two banks, f, the packed
// mask, the control block, the partial sums and one norm-history slot) is
// laid out in `scratch`, a device array of QSM_PD_STATIC_BYTES that each
// GPU the library runs on holds until the process ends.  Its address is
// looked up once per device (hipGetSymbolAddress with that device
// current) and kept in scratch_ptr; host code uses only that address.

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_ptr[MAX_DEVICES] = {};

int scratch_of_current_device(void **out)
{
    int dev = 0;
    QSM_HIP(hipGetDevice(&dev));
    QSM_ARG(dev >= 0 && dev < MAX_DEVICES, "device %d beyond the %d the library tracks",
            dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        QSM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        QSM_ARG(p != nullptr, "no device address for the static work space");
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return QSM_OK;
}

template <typename T>
int solve_static_impl(qsm_pd *s, const T *d2f, const uint8_t *mask, double tol,
                      uint32_t maxiter, T *chi)
{
    const hipStream_t st = s->stream;
    // qsm_pd_load's work with kernels only (no memset or memcpy nodes when
    // a caller captures the call): zero banks, packed mask straight from
    // the caller's mask, f = sigma * d2f, control block reset
    QSM_HIP(launch_zero_bytes(s->bank(0), 2 * s->L.bank_bytes, st));
    QSM_HIP(launch_mask_init(s->mask(), mask, s->geom(), 0, st));
    QSM_HIP(launch_scale<T>(static_cast<T *>(s->f()), d2f, T(s->sigma), s->N, st));
    QSM_HIP(launch_ctl_reset(s->ctl(), 1, st));
    s->n = 0;
    s->dual_ahead = 0;
    s->cur = 0;
    s->loaded = 1;
    int rc = iterate_impl<T>(s, maxiter, tol);
    if (rc) return rc;
    const Bank<T> b0 = make_bank<T>(s->bank(0), s->kind, s->N);
    const Bank<T> b1 = make_bank<T>(s->bank(1), s->kind, s->N);
    QSM_HIP(launch_select_copy<T>(chi, b0.chi, b1.chi, s->ctl(), s->cur, s->N, st));
    return QSM_OK;
}

int pd_solve_static(int kind, int dtype, const void *d2f, const uint8_t *mask,
                    const size_t sz[3], const double h[3], double lam, double alpha,
                    double tol, uint32_t maxiter, void *chi, void *stream)
{
    QSM_ARG(d2f && mask && chi, "null input or output pointer");
    QSM_ARG(kind == QSM_TV || kind == QSM_TGV, "kind must be QSM_TV or QSM_TGV");
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "dtype must be QSM_F64 or QSM_F32");
    QSM_ARG(maxiter > 0, "maxiter must be > 0");
    int rc = check_grid(sz, h, 1);
    if (rc) return rc;
    if ((rc = ensure_device())) return rc;

    qsm_pd s{};
    s.kind = kind;
    s.dtype = dtype;
    for (int a = 0; a < 3; ++a) { s.sz[a] = sz[a]; s.h[a] = h[a]; }
    s.N = sz[0] * sz[1] * sz[2];
    s.lam = lam;
    s.alpha = alpha;
    s.stream = as_stream(stream);
    s.shape = default_shape(kind, dtype, s.geom());
    s.kz2 = default_kz2(dtype, s.geom());
    s.cur = 0;
    s.schedule = 0;
    qsm_step_sizes(h, &s.tau, &s.sigma);
    s.npart = max_partials(s.geom());
    s.hist_cap = 1;
    s.L = layout(kind, dtype, sz, s.npart, s.hist_cap, false);
    QSM_ARG(s.L.total <= size_t(QSM_PD_STATIC_BYTES),
            "volume %zux%zux%zu needs %zu bytes of solver state, more than the %zu-byte "
            "static work space", sz[0], sz[1], sz[2], s.L.total,
            size_t(QSM_PD_STATIC_BYTES));
    void *base = nullptr;
    if ((rc = scratch_of_current_device(&base))) return rc;
    s.mem = static_cast<char *>(base);
    return dtype == QSM_F64
        ? solve_static_impl<double>(&s, static_cast<const double *>(d2f), mask, tol, maxiter,
                                    static_cast<double *>(chi))
        : solve_static_impl<float>(&s, static_cast<const float *>(d2f), mask, tol, maxiter,
                                   static_cast<float *>(chi));
}

}  // namespace

extern "C" {

int qsm_pd_solve_static(int kind, int dtype, const void *d2f, const uint8_t *mask,
                        const size_t sz[3], const double h[3], double lam, double alpha,
                        double tol, uint32_t maxiter, void *chi, void *stream)
{
    return pd_solve_static(kind, dtype, d2f, mask, sz, h, lam, alpha, tol, maxiter, chi,
                           stream);
}

void qsm_step_sizes(const double h[3], double *tau, double *sigma)
{
    // pd_tgv_mex.c:272-278
    const double gnrm = 4.0 * (1.0 / (h[0] * h[0]) + 1.0 / (h[1] * h[1])
                               + 1.0 / (h[2] * h[2]));
    const double wnrm = 6.0 * (1.0 / (h[0] * h[0]) + 1.0 / (h[1] * h[1])
                               + 2.0 / (h[2] * h[2])) / 3.0;
    const double K2 = gnrm * gnrm + wnrm * wnrm;
    const double t = 1 / std::sqrt(K2);
    *tau = t;
    *sigma = 1 / (t * K2);
}

size_t qsm_pd_footprint(int kind, int dtype, const size_t sz[3])
{
    if (!sz) return 0;
    const Geom g{int(sz[0]), int(sz[1]), int(sz[2])};
    const int npart = max_partials(g);
    return layout(kind, dtype, sz, npart, 1024).total;
}

int qsm_pd_create(qsm_pd **out, int kind, int dtype, const size_t sz[3],
                  const double h[3], double lam, double alpha, void *stream)
{
    return create_impl(out, kind, dtype, sz, h, lam, alpha, stream, 1024);
}

int qsm_pd_load(qsm_pd *s, const void *d2f, const uint8_t *mask, int on_device)
{
    QSM_ARG(s && d2f && mask, "null handle or input pointer");
    const size_t es = dtype_size(s->dtype);
    const hipMemcpyKind kd = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    uint8_t *raw = static_cast<uint8_t *>(s->bank(1));  // staging, cleared below
    QSM_HIP(hipMemcpyAsync(raw, mask, s->N, kd, s->stream));
    QSM_HIP(launch_mask_init(s->mask(), raw, s->geom(), 0, s->stream));
    QSM_HIP(hipMemsetAsync(s->bank(0), 0, 2 * s->L.bank_bytes, s->stream));
    QSM_HIP(hipMemcpyAsync(s->f(), d2f, s->N * es, kd, s->stream));
    if (s->dtype == QSM_F64)
        QSM_HIP(launch_scale<double>(static_cast<double *>(s->f()),
                                     static_cast<const double *>(s->f()),
                                     s->sigma, s->N, s->stream));
    else
        QSM_HIP(launch_scale<float>(static_cast<float *>(s->f()),
                                    static_cast<const float *>(s->f()),
                                    float(s->sigma), s->N, s->stream));
    QSM_HIP(launch_ctl_reset(s->ctl(), 1, s->stream));
    if (!on_device)
        QSM_HIP(hipStreamSynchronize(s->stream));  // host buffers may go away
    const int rc = build_items(s);
    if (rc) return rc;
    s->n = 0;
    s->dual_ahead = 0;
    s->cur = 0;
    s->loaded = 1;
    return QSM_OK;
}

int qsm_pd_iterate(qsm_pd *s, uint32_t niter, double tol)
{
    QSM_ARG(s && s->loaded, "handle not loaded");
    if (s->dtype == QSM_F64)
        return iterate_impl<double>(s, niter, tol);
    return iterate_impl<float>(s, niter, tol);
}

int qsm_pd_status(qsm_pd *s, uint32_t *iters, int *stopped, double nr[2])
{
    QSM_ARG(s, "null handle");
    Ctl c;
    QSM_HIP(hipMemcpyAsync(&c, s->ctl(), sizeof c, hipMemcpyDeviceToHost, s->stream));
    QSM_HIP(hipStreamSynchronize(s->stream));
    if (iters) *iters = c.stop ? c.stop_iter : s->n;
    if (stopped) *stopped = c.stop;
    if (nr) { nr[0] = c.nr[0]; nr[1] = c.nr[1]; }
    return QSM_OK;
}

int qsm_pd_norm_history(qsm_pd *s, double *nr, int cap)
{
    if (!s || !nr || cap <= 0) return 0;
    Ctl c;
    if (hipMemcpyAsync(&c, s->ctl(), sizeof c, hipMemcpyDeviceToHost, s->stream) != hipSuccess
        || hipStreamSynchronize(s->stream) != hipSuccess)
        return 0;
    int n = int(c.checks);
    if (n > s->hist_cap) n = s->hist_cap;
    if (n > cap) n = cap;
    if (n > 0
        && hipMemcpy(nr, s->hist(), size_t(n) * 2 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return 0;
    return n;
}

int qsm_pd_get_chi(qsm_pd *s, void *chi, int on_device)
{
    QSM_ARG(s && chi, "null handle or output pointer");
    if (s->dtype == QSM_F64)
        return get_chi_impl<double>(s, chi, on_device);
    return get_chi_impl<float>(s, chi, on_device);
}

void qsm_pd_destroy(qsm_pd *s)
{
    if (!s) return;
    if (s->mem) {
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(s->mem);
    }
    delete s;
}

int qsm_pd_set_schedule(qsm_pd *s, int schedule)
{
    QSM_ARG(s, "null handle");
    QSM_ARG(schedule >= 0 && schedule <= 2, "schedule must be 0, 1 or 2");
    QSM_ARG(schedule == 1 || fused_size_supported(s->kind, s->dtype, s->N),
            "volume too large for the fused schedules (bank >= 4 GiB)");
    // the fused schedules share the ping-pong banks and may alternate; the
    // in-place two-sweep schedule only starts from a fresh load
    QSM_ARG((s->n == 0 && !s->dual_ahead) || (schedule != 1 && s->schedule != 1),
            "the two-sweep schedule can only be switched to or from before the first "
            "iteration");
    s->schedule = schedule;
    return QSM_OK;
}

int qsm_pd_set_tiles(qsm_pd *s, int tx, int ty, int kz)
{
    QSM_ARG(s, "null handle");
    QSM_ARG(fused_shape_supported(tx, ty), "unsupported tile shape %dx%d", tx, ty);
    QSM_ARG(kz >= 1, "kz must be >= 1");
    const FusedShape sh{tx, ty, kz};
    QSM_ARG(fused_blocks(s->geom(), sh) <= s->npart,
            "tile shape needs %d partial-sum slots, %d allocated",
            fused_blocks(s->geom(), sh), s->npart);
    s->shape = sh;
    return QSM_OK;
}

int qsm_pd_set_kz2(qsm_pd *s, int kz)
{
    QSM_ARG(s, "null handle");
    QSM_ARG(kz >= 1, "kz must be >= 1");
    QSM_ARG(fused2_blocks(s->geom(), s->dtype, kz) <= s->npart,
            "chunk depth needs %d partial-sum slots, %d allocated",
            fused2_blocks(s->geom(), s->dtype, kz), s->npart);
    s->kz2 = kz;
    return QSM_OK;
}

int qsm_pd_set_active_items(qsm_pd *s, int on)
{
    QSM_ARG(s, "null handle");
    QSM_ARG(on >= 0 && on <= 2, "on must be 0, 1 or 2");
    s->items_on = on;
    if (!s->items_on) {
        s->nitems = 0;
        s->items_kz = 0;
        return QSM_OK;
    }
    return s->loaded ? build_items(s) : QSM_OK;
}

int qsm_pd_active_items(qsm_pd *s, int *nitems, int *kz)
{
    QSM_ARG(s, "null handle");
    if (nitems) *nitems = s->nitems;
    if (kz) *kz = s->items_kz;
    return QSM_OK;
}

int qsm_pd_launch_counts(qsm_pd *s, uint64_t counts[4])
{
    QSM_ARG(s && counts, "null handle or output");
    for (int i = 0; i < 4; ++i)
        counts[i] = s->launches[i];
    return QSM_OK;
}

int qsm_pd_get_tiles(qsm_pd *s, int *tx, int *ty, int *kz)
{
    QSM_ARG(s, "null handle");
    if (tx) *tx = s->shape.tx;
    if (ty) *ty = s->shape.ty;
    if (kz) *kz = s->shape.kz;
    return QSM_OK;
}

int qsm_pd_tgv(int dtype, const void *d2f, const uint8_t *mask,
               const size_t sz[3], const double h[3], double lam,
               double alpha, double tol, uint32_t maxiter, uint32_t verbose,
               void *chi, qsm_print_fn print, void *print_ctx)
{
    return pd_solve(QSM_TGV, dtype, d2f, mask, sz, h, lam, alpha, tol, maxiter,
                    verbose, chi, print, print_ctx);
}

// sstgv.m:54-89 on host arrays, echo by echo, with every step after the
// mask's bounding box on the device: one upload of the echo and the mask,
// the crop (box +/-2, sstgv.m:57-66), d2u = lapw(u) on the crop, the solve
// of pd_tgv_mex(-d2u, ...) (alpha > eps) or pd_tv_mex(-d2u, ...) (sstgv.m:
// 76-87) with the same progress output, the uncrop into zeros, one
// download.
int qsm_sstgv(int dtype, const void *phas, const uint8_t *mask, const size_t sz[3],
              size_t nechoes, const double vsz[3], double lam, double alpha, double tol,
              uint32_t maxit, uint32_t verbose, void *x, qsm_print_fn print, void *print_ctx)
{
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    QSM_ARG(phas && mask && x && nechoes >= 1, "null pointer or no echo");
    QSM_ARG(maxit > 0, "maxit must be > 0");
    int rc = check_grid(sz, vsz, 1);
    if (rc) return rc;
    const size_t es = dtype_size(dtype), N = sz[0] * sz[1] * sz[2];
    std::vector<char> any[3];
    axis_any(mask, sz, any);
    size_t c0[3], cs[3];
    for (int a = 0; a < 3; ++a) {
        size_t lo = 0, hi = sz[a];
        while (lo < sz[a] && !any[a][lo]) ++lo;
        while (hi > lo && !any[a][hi - 1]) --hi;
        if (lo == hi) {  // empty mask: chi = 0
            std::memset(x, 0, N * nechoes * es);
            return QSM_OK;
        }
        c0[a] = lo >= 2 ? lo - 2 : 0;
        cs[a] = std::min(sz[a], hi + 2) - c0[a];
    }
    if ((rc = check_grid(cs, vsz, 1))) return rc;
    if ((rc = ensure_device())) return rc;
    const size_t n = cs[0] * cs[1] * cs[2];
    std::vector<int> ix;
    for (int a = 0; a < 3; ++a)
        for (size_t i = 0; i < cs[a]; ++i) ix.push_back(int(c0[a] + i));
    DevBuf dP, dM, dix, du, dm, dd, dchi;
    if ((rc = dP.alloc(N * es)) || (rc = dM.alloc(N)) ||
        (rc = dix.alloc(ix.size() * sizeof(int))) || (rc = du.alloc(n * es)) ||
        (rc = dm.alloc(n)) || (rc = dd.alloc(n * es)) || (rc = dchi.alloc(n * es)))
        return rc;
    QSM_HIP(hipMemcpy(dM.p, mask, N, hipMemcpyHostToDevice));
    QSM_HIP(hipMemcpy(dix.p, ix.data(), ix.size() * sizeof(int), hipMemcpyHostToDevice));
    const int *dixp = static_cast<const int *>(dix.p);
    const CropMap cm{dixp, dixp + cs[0], dixp + cs[0] + cs[1], int(cs[0]), int(cs[1]),
                     int(cs[2]), sz[0], sz[1], N, n};
    crop_gather_kernel<uint8_t, true><<<crop_blocks(cm, 1), 256>>>(
        static_cast<uint8_t *>(dm.p), static_cast<const uint8_t *>(dM.p), cm, 1);
    QSM_HIP(hipGetLastError());
    const bool tgv = alpha > std::numeric_limits<double>::epsilon();
    for (size_t e = 0; e < nechoes; ++e) {
        QSM_HIP(hipMemcpy(dP.p, static_cast<const char *>(phas) + e * N * es, N * es,
                          hipMemcpyHostToDevice));
        // d2f = -lapw(u) on the crop (lapw writes the interior; faces 0)
        QSM_HIP(hipMemsetAsync(dd.p, 0, n * es, nullptr));
        if (dtype == QSM_F32)
            crop_gather_kernel<float><<<crop_blocks(cm, 1), 256>>>(
                static_cast<float *>(du.p), static_cast<const float *>(dP.p), cm, 1);
        else
            crop_gather_kernel<double><<<crop_blocks(cm, 1), 256>>>(
                static_cast<double *>(du.p), static_cast<const double *>(dP.p), cm, 1);
        QSM_HIP(hipGetLastError());
        if ((rc = qsm_lapw_dev(dtype, dd.p, du.p, cs, vsz, nullptr))) return rc;
        if (dtype == QSM_F32)
            QSM_HIP(launch_scale<float>(static_cast<float *>(dd.p),
                                        static_cast<const float *>(dd.p), -1.0f, n, nullptr));
        else
            QSM_HIP(launch_scale<double>(static_cast<double *>(dd.p),
                                         static_cast<const double *>(dd.p), -1.0, n, nullptr));
        // the solve of pd_solve, from and to device arrays
        qsm_pd *h = nullptr;
        if ((rc = create_impl(&h, tgv ? QSM_TGV : QSM_TV, dtype, cs, vsz, lam,
                              tgv ? alpha : 0.0, nullptr, int(maxit / 10) + 1)))
            return rc;
        if (!(rc = qsm_pd_load(h, dd.p, static_cast<const uint8_t *>(dm.p), 1)) &&
            !(rc = qsm_pd_iterate(h, maxit, tol)) && !(rc = qsm_pd_get_chi(h, dchi.p, 1)))
            print_progress(h, maxit, verbose, print, print_ctx);
        qsm_pd_destroy(h);
        if (rc) return rc;
        // uncrop into zeros
        QSM_HIP(hipMemsetAsync(dP.p, 0, N * es, nullptr));
        if (dtype == QSM_F32)
            crop_scatter_kernel<float><<<crop_blocks(cm, 1), 256>>>(
                static_cast<float *>(dP.p), static_cast<const float *>(dchi.p), cm, 1);
        else
            crop_scatter_kernel<double><<<crop_blocks(cm, 1), 256>>>(
                static_cast<double *>(dP.p), static_cast<const double *>(dchi.p), cm, 1);
        QSM_HIP(hipGetLastError());
        QSM_HIP(hipMemcpy(static_cast<char *>(x) + e * N * es, dP.p, N * es,
                          hipMemcpyDeviceToHost));
    }
    return QSM_OK;
}

int qsm_pd_tv(int dtype, const void *d2f, const uint8_t *mask,
              const size_t sz[3], const double h[3], double lam, double tol,
              uint32_t maxiter, uint32_t verbose, void *chi,
              qsm_print_fn print, void *print_ctx)
{
    return pd_solve(QSM_TV, dtype, d2f, mask, sz, h, lam, 0.0, tol, maxiter,
                    verbose, chi, print, print_ctx);
}

// ---- single sweeps on caller device arrays ------------------------------

int qsm_tgv_dual_dev(int dtype, void *nu, void *p, void *q, const void *chi_,
                     const void *eta_, const void *w_, const void *f,
                     const uint8_t *mi, const size_t sz[3], const double h[3],
                     double alpha0, double alpha1, double sigma, void *stream)
{
    int rc = check_grid(sz, h, 1);
    if (rc) return rc;
    const size_t N = sz[0] * sz[1] * sz[2];
    const Geom g{int(sz[0]), int(sz[1]), int(sz[2])};
    if (dtype == QSM_F64) {
        using T = double;
        T *P = static_cast<T *>(p), *Q = static_cast<T *>(q);
        const T *W = static_cast<const T *>(w_);
        Bank<T> b{};
        b.nu = static_cast<T *>(nu);
        b.p1 = P; b.p2 = P + N; b.p3 = P + 2 * N;
        b.q1 = Q; b.q2 = Q + N; b.q3 = Q + 2 * N; b.q4 = Q + 3 * N; b.q5 = Q + 4 * N; b.q6 = Q + 5 * N;
        Bars<T> br{const_cast<T *>(static_cast<const T *>(chi_)),
                   const_cast<T *>(static_cast<const T *>(eta_)),
                   const_cast<T *>(W), const_cast<T *>(W + N), const_cast<T *>(W + 2 * N)};
        QSM_HIP(launch_dual_sweep<T, true>(b, br, static_cast<const T *>(f), mi, 0xff, g,
                                           make_dual_coef<T>(h, sigma, alpha0, alpha1),
                                           nullptr, as_stream(stream)));
    } else if (dtype == QSM_F32) {
        using T = float;
        T *P = static_cast<T *>(p), *Q = static_cast<T *>(q);
        const T *W = static_cast<const T *>(w_);
        Bank<T> b{};
        b.nu = static_cast<T *>(nu);
        b.p1 = P; b.p2 = P + N; b.p3 = P + 2 * N;
        b.q1 = Q; b.q2 = Q + N; b.q3 = Q + 2 * N; b.q4 = Q + 3 * N; b.q5 = Q + 4 * N; b.q6 = Q + 5 * N;
        Bars<T> br{const_cast<T *>(static_cast<const T *>(chi_)),
                   const_cast<T *>(static_cast<const T *>(eta_)),
                   const_cast<T *>(W), const_cast<T *>(W + N), const_cast<T *>(W + 2 * N)};
        QSM_HIP(launch_dual_sweep<T, true>(b, br, static_cast<const T *>(f), mi, 0xff, g,
                                           make_dual_coef<T>(h, sigma, alpha0, alpha1),
                                           nullptr, as_stream(stream)));
    } else {
        QSM_ARG(false, "bad dtype");
    }
    return QSM_OK;
}

}  // extern "C"

namespace {

template <typename T>
int tgv_primal_dev_t(void *chi, void *eta, void *w, void *chi_, void *eta_,
                     void *w_, const void *nu, const void *p, const void *q,
                     const uint8_t *m, const size_t sz[3], const double h[3],
                     double tau, void *stream, bool tgv)
{
    const size_t N = sz[0] * sz[1] * sz[2];
    const Geom g{int(sz[0]), int(sz[1]), int(sz[2])};
    T *P = const_cast<T *>(static_cast<const T *>(p));
    Bank<T> b{};
    b.chi = static_cast<T *>(chi);
    b.eta = static_cast<T *>(eta);
    b.nu = const_cast<T *>(static_cast<const T *>(nu));
    b.p1 = P; b.p2 = P + N; b.p3 = P + 2 * N;
    Bars<T> br{static_cast<T *>(chi_), static_cast<T *>(eta_), nullptr, nullptr, nullptr};
    const PrimalCoef<T> c = make_primal_coef<T>(h, tau);
    if (tgv) {
        T *W = static_cast<T *>(w), *B = static_cast<T *>(w_);
        T *Q = const_cast<T *>(static_cast<const T *>(q));
        b.w1 = W; b.w2 = W + N; b.w3 = W + 2 * N;
        b.q1 = Q; b.q2 = Q + N; b.q3 = Q + 2 * N; b.q4 = Q + 3 * N; b.q5 = Q + 4 * N; b.q6 = Q + 5 * N;
        br.b1 = B; br.b2 = B + N; br.b3 = B + 2 * N;
        QSM_HIP(launch_primal_sweep<T, true>(b, br, m, 0xff, g, c, nullptr, nullptr, as_stream(stream)));
    } else {
        QSM_HIP(launch_primal_sweep<T, false>(b, br, m, 0xff, g, c, nullptr, nullptr, as_stream(stream)));
    }
    return QSM_OK;
}

template <typename T>
int tv_dual_dev_t(void *nu, void *p, const void *chi_, const void *eta_,
                  const void *f, const uint8_t *mi, const size_t sz[3],
                  const double h[3], double lam, double sigma, void *stream)
{
    const size_t N = sz[0] * sz[1] * sz[2];
    const Geom g{int(sz[0]), int(sz[1]), int(sz[2])};
    T *P = static_cast<T *>(p);
    Bank<T> b{};
    b.nu = static_cast<T *>(nu);
    b.p1 = P; b.p2 = P + N; b.p3 = P + 2 * N;
    Bars<T> br{const_cast<T *>(static_cast<const T *>(chi_)),
               const_cast<T *>(static_cast<const T *>(eta_)), nullptr, nullptr, nullptr};
    QSM_HIP(launch_dual_sweep<T, false>(b, br, static_cast<const T *>(f), mi, 0xff, g,
                                        make_dual_coef<T>(h, sigma, lam, lam), nullptr,
                                        as_stream(stream)));
    return QSM_OK;
}

}  // namespace

extern "C" {

int qsm_tgv_primal_dev(int dtype, void *chi, void *eta, void *w, void *chi_,
                       void *eta_, void *w_, const void *nu, const void *p,
                       const void *q, const uint8_t *m, const size_t sz[3],
                       const double h[3], double tau, void *stream)
{
    int rc = check_grid(sz, h, 1);
    if (rc) return rc;
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    return dtype == QSM_F64
        ? tgv_primal_dev_t<double>(chi, eta, w, chi_, eta_, w_, nu, p, q, m, sz, h, tau, stream, true)
        : tgv_primal_dev_t<float>(chi, eta, w, chi_, eta_, w_, nu, p, q, m, sz, h, tau, stream, true);
}

int qsm_tv_primal_dev(int dtype, void *chi, void *eta, void *chi_, void *eta_,
                      const void *nu, const void *p, const uint8_t *m,
                      const size_t sz[3], const double h[3], double tau,
                      void *stream)
{
    int rc = check_grid(sz, h, 1);
    if (rc) return rc;
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    return dtype == QSM_F64
        ? tgv_primal_dev_t<double>(chi, eta, nullptr, chi_, eta_, nullptr, nu, p, nullptr, m, sz, h, tau, stream, false)
        : tgv_primal_dev_t<float>(chi, eta, nullptr, chi_, eta_, nullptr, nu, p, nullptr, m, sz, h, tau, stream, false);
}

int qsm_tv_dual_dev(int dtype, void *nu, void *p, const void *chi_,
                    const void *eta_, const void *f, const uint8_t *mi,
                    const size_t sz[3], const double h[3], double lam,
                    double sigma, void *stream)
{
    int rc = check_grid(sz, h, 1);
    if (rc) return rc;
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    return dtype == QSM_F64
        ? tv_dual_dev_t<double>(nu, p, chi_, eta_, f, mi, sz, h, lam, sigma, stream)
        : tv_dual_dev_t<float>(nu, p, chi_, eta_, f, mi, sz, h, lam, sigma, stream);
}

int qsm_init_mask_dev(uint8_t *mi, const uint8_t *m, const size_t sz[3],
                      void *stream)
{
    const double h1[3] = {1.0, 1.0, 1.0};
    int rc = check_grid(sz, h1, 1);
    if (rc) return rc;
    QSM_ARG(mi && m, "null mask pointer");
    QSM_HIP(launch_mask_init(mi, m, Geom{int(sz[0]), int(sz[1]), int(sz[2])}, 1,
                             as_stream(stream)));
    return QSM_OK;
}

int qsm_convergence_dev(int dtype, const void *chi, const void *chi0,
                        const void *eta, const void *eta0, size_t N,
                        double tol, double out[2], int *flag, void *stream)
{
    QSM_ARG(chi && chi0 && eta && eta0 && out, "null pointer");
    QSM_ARG(dtype == QSM_F64 || dtype == QSM_F32, "bad dtype");
    const hipStream_t st = as_stream(stream);
    const int nblk = 512;
    DevBuf buf;
    int rc = buf.alloc(256 + size_t(nblk) * 4 * sizeof(double));
    if (rc) return rc;
    Ctl *ctl = static_cast<Ctl *>(buf.p);
    double *part = reinterpret_cast<double *>(static_cast<char *>(buf.p) + 256);
    QSM_HIP(launch_ctl_reset(ctl, 1, st));
    if (dtype == QSM_F64)
        QSM_HIP(launch_conv_sums<double>(static_cast<const double *>(chi),
                                         static_cast<const double *>(chi0),
                                         static_cast<const double *>(eta),
                                         static_cast<const double *>(eta0), N, part, nblk, st));
    else
        QSM_HIP(launch_conv_sums<float>(static_cast<const float *>(chi),
                                        static_cast<const float *>(chi0),
                                        static_cast<const float *>(eta),
                                        static_cast<const float *>(eta0), N, part, nblk, st));
    QSM_HIP(launch_conv_finalize(part, nblk, ctl, nullptr, 0, tol, 0, 0, st));
    Ctl c;
    QSM_HIP(hipMemcpyAsync(&c, ctl, sizeof c, hipMemcpyDeviceToHost, st));
    QSM_HIP(hipStreamSynchronize(st));
    out[0] = c.nr[0];
    out[1] = c.nr[1];
    if (flag) *flag = c.stop;
    return QSM_OK;
}

}  // extern "C"
