// fused2_dev.h -- the two-iteration primal-dual march (temporal blocking)
// for gfx950, shared by the single-volume kernel (fused2.hip) and the
// batched persistent kernel (batch.hip).
//
// Pass (n, n+1) reads the state of bank A (primal iterates of n-1, dual of n)
// once and writes the state two iterations later to bank B (primal of n+1,
// dual of n+2): per iteration a voxel's state crosses HBM half as often as in
// the one-iteration pass (fused.hip).  The arithmetic of every update is that
// of pd_math.h, in the reference's order, so the iterates are bitwise those
// of two one-iteration passes.
//
// Block = a BX x BY region of (x, y) columns, one thread each, marching up a
// z-chunk.  Four stages run per plane step, each one plane behind the
// previous (a stage at plane k needs its predecessor at k+1):
//
//   stage 1  primal n      at plane p     on every column of the region
//   stage 2  dual n+1      at plane p-1   on columns at depth >= 1
//   stage 3  primal n+1    at plane p-2   on columns at depth >= 2
//   stage 4  dual n+2      at plane p-3   on columns at depth >= 3 (the tile)
//
// where depth is the distance to the region's edge: each stage needs its
// predecessor's results on the one-cell ring around it, so the owned tile is
// the region less three rings, TX = BX - 6 by TY = BY - 6.  The x/y
// neighbours of a stage's inputs come from LDS (the predecessor's plane,
// double-buffered: one barrier per step), the z neighbours from the
// column's own registers.  Stage 1 reads its inputs (and their x/y
// neighbours) from bank A; only the tile's stage 3 and stage 4 results are
// stored.  Ring columns belong to other blocks' tiles: every launch reads A
// and writes B, so no block reads what another writes.
//
// A march in mode 1 runs stages 1 and 2 only and stores their results on
// the tile: one iteration (primal n, dual n+1), for a volume whose next
// iteration is a convergence check that must see the state after it, or
// whose budget has one iteration left.
#pragma once

#include "fused_dev.h"

// Region (tile plus three rings) of the two-iteration pass per precision:
// 38x22 columns (a 32x16 tile, 836 columns on 896 threads) in FP32, 22x22
// (16x16, 484 columns on 512 threads) in FP64, whose registers allow half
// the waves.  Both keep a tile row at 128 bytes, one cache line.
#ifndef QSM_F2_F32_BX
#define QSM_F2_F32_BX 38
#define QSM_F2_F32_BY 22
#endif
#ifndef QSM_F2_F64_BX
#define QSM_F2_F64_BX 22
#define QSM_F2_F64_BY 22
#endif

namespace qsm {
namespace fdev {

// LDS of one block: two plane slots of the stage 1 extrapolations (E1:
// chi_, eta_, w_[3]), of the stage 2 dual values the primal reads at x/y
// neighbours (D: nu, p1, p2, q1..q5) and of the stage 3 extrapolations (E3).
template <typename T, bool TGV, int NC>
struct Lds2 {
    static constexpr int NE = TGV ? 5 : 2;
    static constexpr int ND = TGV ? 8 : 3;
    T e1[2][NE][NC];
    T d[2][ND][NC];
    T e3[2][NE][NC];
};

// The primal values of one column at one plane that stage 3 needs from
// stage 1 two steps later (the old iterates of its update).
template <typename T>
struct Old {
    T chi, eta, w1, w2, w3;
};

// Threads of a block: the region's columns rounded up to whole waves (the
// extra threads hold no column and only take part in the barriers).
template <int BX, int BY>
constexpr int threads2() { return (BX * BY + 63) / 64 * 64; }

// Thread -> column of the region.  The owned TX x TY tile comes first, row
// by row, so its stores fill whole waves and whole cache lines (TX = 32
// floats or 16 doubles: 128 bytes); then the rings from the inside out
// (depth 2, 1, 0; each a rectangle's perimeter: top row, bottom row, left
// column, right column), so that the waves of the outer rings have no work
// in the later stages and skip them as a whole.  dep = -1: padding thread.
template <int BX, int BY>
__device__ __forceinline__ void region_cell(int t, int &tx, int &ty, int &dep)
{
    constexpr int TX = BX - 6, TY = BY - 6;
    if (t < TX * TY) {
        tx = 3 + t % TX;
        ty = 3 + t / TX;
        dep = 3;
        return;
    }
    int r = t - TX * TY;
    for (int d = 2; d >= 0; --d) {
        const int w = BX - 2 * d, h = BY - 2 * d;
        const int per = 2 * w + 2 * (h - 2);
        if (r < per) {
            dep = d;
            if (r < w) { tx = d + r; ty = d; }
            else if (r < 2 * w) { tx = d + r - w; ty = BY - 1 - d; }
            else if (r < 2 * w + h - 2) { tx = d; ty = d + 1 + (r - 2 * w); }
            else { tx = BX - 1 - d; ty = d + 1 + (r - 2 * w - (h - 2)); }
            return;
        }
        r -= per;
    }
    tx = 0;
    ty = 0;
    dep = -1;
}

// One tile (bx, by), owned planes [k0, k1), of a BX x BY region: MODE 2 advances the
// tile two iterations (primal n and n+1, dual n+1 and n+2), MODE 1 one.
// Convergence terms of the last primal update of the owned voxels
// accumulate into acc (CHECK).  Every thread of the block calls it.
template <typename T, bool TGV, int BX, int BY, bool CHECK, int MODE>
__device__ __forceinline__ void fused2_tile(Lds2<T, TGV, threads2<BX, BY>()> &L, const Ctx<T> &c,
                                            const DualCoef<T> &dc, const PrimalCoef<T> &pc,
                                            int nx, int ny, int nz, int bx, int by, int k0,
                                            int k1, double acc[4])
{
    using I = Ix<TGV>;
    constexpr uint32_t S = sizeof(T);
    constexpr int TX = BX - 6, TY = BY - 6;
    const int t = threadIdx.x;
    int tx, ty, dep;
    region_cell<BX, BY>(t, tx, ty, dep);
    // LDS position of the thread's column (padding threads: a private slot
    // past the region)
    const int lo = dep >= 0 ? ty * BX + tx : t;
    const int x = bx * TX - 3 + tx, y = by * TY - 3 + ty;
    const bool valid = dep >= 0 && x >= 0 && x < nx && y >= 0 && y < ny;
    const bool own = valid && dep >= 3;
    // two iterations: stage 1 from k0 - 3 (stage 4's first plane k0 needs
    // stage 3 at k0 - 1, stage 2 at k0 - 2, stage 1 at k0 - 3) to k1 + 2;
    // one iteration: stages 1-2 only, stage 2 on the owned planes
    constexpr bool two = MODE == 2;
    const int pf = two ? k0 - 3 : k0 - 1;
    const int pl = two ? k1 + 2 : k1;
    const int lo2 = two ? k0 - 2 : k0;  // lowest plane stage 2 computes

    ColQ<T> q;
    q.l0 = valid ? x + nx * y : -1;
    col_init<T, TGV>(c, pf, true, q);

    // registers carried from step to step (plane of each in brackets, p the
    // step's stage 1 plane)
    T e1m_c = T(0), e1m_e = T(0);  // E1 chi_, eta_ [p-2]
    DualVox<T> dv1{};               // dual n [p-1] (stage 2's old values)
    Old<T> oa{}, ob{};              // primal n [p-1], [p-2] (stage 3's old values)
    DualVox<T> da{}, db{};          // dual n+1 [p-2], [p-3]
    T e3m_c = T(0), e3m_e = T(0);  // E3 chi_, eta_ [p-4]
    T fa = T(0), fb = T(0), fc = T(0);  // f [p-1], [p-2], [p-3]
    uint32_t mq = 0;                // mask bytes [p-1], [p-2], [p-3] in bits 0-7, 8-15, 16-23

    for (int p = pf; p <= pl; ++p) {
        const int s = p & 1, ps = s ^ 1;

        // ---- stage 1: primal n at plane p (every column) -----------------
        const uint8_t m1 = q.mc;
        T fcur = T(0);
        if (p >= lo2 && p <= (two ? k1 + 1 : k1 - 1) && (m1 & 2))
            fcur = bld<T>(c.f, uint32_t(q.l0 + c.sk * p) * S, 0);
        DualVox<T> dv;
        bool pact;
        POut<T> o;
        pcol<T, TGV>(c, pc, p, true, q, dv, pact, o);
        if (!two && own && pact && p >= k0 && p < k1) {
            // one iteration: the primal of the owned planes is final
            const uint32_t vb = uint32_t(q.l0 + c.sk * p) * S;
            bst(o.chi, c.out, vb, c.so(I::CHI));
            bst(o.eta, c.out, vb, c.so(I::ETA));
            if constexpr (TGV) {
                bst(o.w1, c.out, vb, c.so(I::W1));
                bst(o.w2, c.out, vb, c.so(I::W2));
                bst(o.w3, c.out, vb, c.so(I::W3));
            }
            if constexpr (CHECK)
                conv_terms(o.chi, o.chi_o, o.eta, o.eta_o, acc);
        }
        L.e1[s][0][lo] = o.cb;
        L.e1[s][1][lo] = o.eb;
        if constexpr (TGV) {
            L.e1[s][2][lo] = o.b1;
            L.e1[s][3][lo] = o.b2;
            L.e1[s][4][lo] = o.b3;
        }

        // ---- stage 2: dual n+1 at plane p-1 (depth >= 1) -----------------
        const uint8_t m2 = uint8_t(mq);
        DualVox<T> d2 = dv1;
        if (p - 1 >= lo2 && dep >= 1 && (m2 & 2)) {
            const Star<T> cs{L.e1[ps][0][lo], L.e1[ps][0][lo - 1], L.e1[ps][0][lo + 1],
                             L.e1[ps][0][lo - BX], L.e1[ps][0][lo + BX], e1m_c, o.cb};
            const Star<T> es{L.e1[ps][1][lo], L.e1[ps][1][lo - 1], L.e1[ps][1][lo + 1],
                             L.e1[ps][1][lo - BX], L.e1[ps][1][lo + BX], e1m_e, o.eb};
            d2.nu = dual_nu(dc, d2.nu, es, cs, fa);
            if constexpr (TGV) {
                tgv_dual_pq(dc, d2, cs,
                            L.e1[ps][2][lo], L.e1[ps][2][lo + 1], L.e1[ps][2][lo + BX], o.b1,
                            L.e1[ps][3][lo], L.e1[ps][3][lo + 1], L.e1[ps][3][lo + BX], o.b2,
                            L.e1[ps][4][lo], L.e1[ps][4][lo + 1], L.e1[ps][4][lo + BX], o.b3);
            } else {
                tv_dual_p(dc, d2, cs);
            }
        }
        L.d[s][0][lo] = d2.nu;
        L.d[s][1][lo] = d2.p1;
        L.d[s][2][lo] = d2.p2;
        if constexpr (TGV) {
            L.d[s][3][lo] = d2.q1;
            L.d[s][4][lo] = d2.q2;
            L.d[s][5][lo] = d2.q3;
            L.d[s][6][lo] = d2.q4;
            L.d[s][7][lo] = d2.q5;
        }
        if constexpr (!two) {
            // one iteration: the dual of the owned planes is final
            if (own && p - 1 >= k0 && p - 1 < k1 && (m2 & 2)) {
                const uint32_t vb = uint32_t(q.l0 + c.sk * (p - 1)) * S;
                if constexpr (TGV) {
                    bst(d2.q1, c.out, vb, c.so(I::Q1));
                    bst(d2.q2, c.out, vb, c.so(I::Q2));
                    bst(d2.q3, c.out, vb, c.so(I::Q3));
                    bst(d2.q4, c.out, vb, c.so(I::Q4));
                    bst(d2.q5, c.out, vb, c.so(I::Q5));
                    bst(d2.q6, c.out, vb, c.so(I::Q6));
                }
                bst(d2.nu, c.out, vb, c.so(I::NU));
                bst(d2.p1, c.out, vb, c.so(I::P1));
                bst(d2.p2, c.out, vb, c.so(I::P2));
                bst(d2.p3, c.out, vb, c.so(I::P3));
            }
            e1m_c = L.e1[ps][0][lo];
            e1m_e = L.e1[ps][1][lo];
            dv1 = dv;
            fa = fcur;
            mq = (mq << 8) | m1;
            __syncthreads();
            continue;
        }

        // ---- stage 3: primal n+1 at plane p-2 (depth >= 2) ---------------
        const int k3 = p - 2;
        const uint8_t m3 = uint8_t(mq >> 8);
        T c3 = T(0), e3 = T(0), b31 = T(0), b32 = T(0), b33 = T(0);
        if (k3 >= k0 - 1 && dep >= 2 && (m3 & 1)) {
            PrimalIn<T> pi;
            pi.nu = Star<T>{da.nu, L.d[ps][0][lo - 1], L.d[ps][0][lo + 1], L.d[ps][0][lo - BX],
                            L.d[ps][0][lo + BX], db.nu, d2.nu};
            pi.p1c = da.p1; pi.p1xm = L.d[ps][1][lo - 1];
            pi.p2c = da.p2; pi.p2ym = L.d[ps][2][lo - BX];
            pi.p3c = da.p3; pi.p3zm = db.p3;
            if constexpr (TGV) {
                pi.q1c = da.q1; pi.q1xm = L.d[ps][3][lo - 1];
                pi.q2c = da.q2; pi.q2xm = L.d[ps][4][lo - 1]; pi.q2ym = L.d[ps][4][lo - BX];
                pi.q3c = da.q3; pi.q3xm = L.d[ps][5][lo - 1]; pi.q3zm = db.q3;
                pi.q4c = da.q4; pi.q4ym = L.d[ps][6][lo - BX];
                pi.q5c = da.q5; pi.q5ym = L.d[ps][7][lo - BX]; pi.q5zm = db.q5;
                pi.q6c = da.q6; pi.q6zm = db.q6;
            }
            T en, cn;
            primal_eta_chi(pc, pi, ob.eta, ob.chi, en, cn);
            c3 = extrap(cn, ob.chi);
            e3 = extrap(en, ob.eta);
            T w1 = ob.w1, w2 = ob.w2, w3 = ob.w3;
            if constexpr (TGV) {
                primal_w(pc, pi, w1, w2, w3);
                b31 = extrap(w1, ob.w1);
                b32 = extrap(w2, ob.w2);
                b33 = extrap(w3, ob.w3);
            }
            if (own && k3 >= k0 && k3 < k1) {
                const uint32_t vb = uint32_t(q.l0 + c.sk * k3) * S;
                bst(cn, c.out, vb, c.so(I::CHI));
                bst(en, c.out, vb, c.so(I::ETA));
                if constexpr (TGV) {
                    bst(w1, c.out, vb, c.so(I::W1));
                    bst(w2, c.out, vb, c.so(I::W2));
                    bst(w3, c.out, vb, c.so(I::W3));
                }
                if constexpr (CHECK)
                    conv_terms(cn, ob.chi, en, ob.eta, acc);  // iteration n+1's check
            }
        }
        L.e3[s][0][lo] = c3;
        L.e3[s][1][lo] = e3;
        if constexpr (TGV) {
            L.e3[s][2][lo] = b31;
            L.e3[s][3][lo] = b32;
            L.e3[s][4][lo] = b33;
        }

        // ---- stage 4: dual n+2 at plane p-3 (the tile) --------------------
        const int k4 = p - 3;
        const uint8_t m4 = uint8_t(mq >> 16);
        if (k4 >= k0 && own && (m4 & 2)) {
            DualVox<T> d4 = db;
            const Star<T> cs{L.e3[ps][0][lo], L.e3[ps][0][lo - 1], L.e3[ps][0][lo + 1],
                             L.e3[ps][0][lo - BX], L.e3[ps][0][lo + BX], e3m_c, c3};
            const Star<T> es{L.e3[ps][1][lo], L.e3[ps][1][lo - 1], L.e3[ps][1][lo + 1],
                             L.e3[ps][1][lo - BX], L.e3[ps][1][lo + BX], e3m_e, e3};
            d4.nu = dual_nu(dc, d4.nu, es, cs, fc);
            const uint32_t vb = uint32_t(q.l0 + c.sk * k4) * S;
            if constexpr (TGV) {
                tgv_dual_pq(dc, d4, cs,
                            L.e3[ps][2][lo], L.e3[ps][2][lo + 1], L.e3[ps][2][lo + BX], b31,
                            L.e3[ps][3][lo], L.e3[ps][3][lo + 1], L.e3[ps][3][lo + BX], b32,
                            L.e3[ps][4][lo], L.e3[ps][4][lo + 1], L.e3[ps][4][lo + BX], b33);
                bst(d4.q1, c.out, vb, c.so(I::Q1));
                bst(d4.q2, c.out, vb, c.so(I::Q2));
                bst(d4.q3, c.out, vb, c.so(I::Q3));
                bst(d4.q4, c.out, vb, c.so(I::Q4));
                bst(d4.q5, c.out, vb, c.so(I::Q5));
                bst(d4.q6, c.out, vb, c.so(I::Q6));
            } else {
                tv_dual_p(dc, d4, cs);
            }
            bst(d4.nu, c.out, vb, c.so(I::NU));
            bst(d4.p1, c.out, vb, c.so(I::P1));
            bst(d4.p2, c.out, vb, c.so(I::P2));
            bst(d4.p3, c.out, vb, c.so(I::P3));
        }

        // ---- advance the column's queues -----------------------------------
        e1m_c = L.e1[ps][0][lo];
        e1m_e = L.e1[ps][1][lo];
        e3m_c = L.e3[ps][0][lo];
        e3m_e = L.e3[ps][1][lo];
        dv1 = dv;
        ob = oa;
        oa = Old<T>{o.chi, o.eta, o.w1, o.w2, o.w3};
        // pcol leaves the old iterates where it did not update: a voxel it
        // skipped keeps zero, which stage 3 never reads (same mask bit)
        db = da;
        da = d2;
        fc = fb;
        fb = fa;
        fa = fcur;
        mq = (mq << 8) | m1;
        __syncthreads();
    }
}

}  // namespace fdev
}  // namespace qsm
