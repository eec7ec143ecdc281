// fused_pair.h -- the single-pass primal-dual march with two voxels per
// lane (single precision).
//
// Each lane owns the x-adjacent voxel pair (x0, x0 + 1), x0 even, and moves
// the pair's state with 8-byte buffer instructions: a lane then issues the
// same number of memory instructions per plane as a double-precision lane,
// for twice the voxels.  Requires an even nx, so that every pair is 8-byte
// aligned (voxel index l = x + nx*(y + ny*z) is even for even x).
//
// The arithmetic of every voxel is the single-voxel march's (pd_math.h, the
// same operands in the same order), so results are bitwise identical to
// fused_tile; only which lane loads what changes:
//   * y and z neighbours are pairs (same x positions);
//   * the x neighbours of a pair are the other element of the pair and, on
//     the outside, voxel x0 - 1 (for element a) and x0 + 2 (for element b),
//     loaded singly;
//   * the ring around the tile is made of pairs too: the -x ring pair is
//     (i0 - 2, i0 - 1), of which element b is the ring cell; the +x ring
//     pair (i0 + W, i0 + W + 1), of which element a is (the other element
//     is computed and ignored);
//   * LDS rows hold pairs, so the dual reads y neighbours as pairs and the
//     outer x neighbours as single floats.
#pragma once

#include "fused_dev.h"

namespace qsm {
namespace fdev {

struct alignas(8) F2 {
    float a, b;
};

// Whole-vector casts only: with this compiler, __builtin_bit_cast of a
// single element of an ext_vector_type value (v.y) yields element 0.
__device__ __forceinline__ F2 bld2(rsrc_t r, uint32_t vo, uint32_t so)
{
    return __builtin_bit_cast(F2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0));
}

__device__ __forceinline__ void bst2(F2 v, rsrc_t r, uint32_t vo, uint32_t so)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)vo, (int)so,
                                          QSM_STORE_AUX);
}

__device__ __forceinline__ uint16_t bld_u16(rsrc_t r, uint32_t vo)
{
    return __builtin_amdgcn_raw_buffer_load_b16(r, (int)vo, 0, 0);
}

// Stores the elements of v selected by ea / eb (one 8-byte store when both).
__device__ __forceinline__ void bst_sel(F2 v, bool ea, bool eb, rsrc_t r, uint32_t vo,
                                        uint32_t so)
{
    if (ea && eb) {
        bst2(v, r, vo, so);
    } else {
        if (ea) bst(v.a, r, vo, so);
        if (eb) bst(v.b, r, vo + 4u, so);
    }
}

// A pair column's queues (as ColQ, per element).
struct ColQ2 {
    int l0;            // index of the pair's first voxel, or -1 outside the grid
    uint8_t ma, mb;    // mask bytes at the current plane kp
    F2 nzm, nzc;       // nu at kp-1, kp
    F2 p3m, q3m, q5m, q6m;
};

struct POut2 {
    POut<float> a, b;
};

// Primal update of one element from the pair's loaded values: x neighbours
// passed explicitly (xm/xp of nu; the x-1 values of p1, q1, q2, q3).
template <bool TGV>
__device__ __forceinline__ void pelem(const PrimalCoef<float> &pc, bool wq, float nuc,
                                      float nxm, float nxp, float nym, float nyp, float nzm,
                                      float nzp, float p1c, float p1xm, float p2c, float p2ym,
                                      float p3c, float p3zm, float q1c, float q1xm, float q2c,
                                      float q2xm, float q2ym, float q3c, float q3xm,
                                      float q3zm, float q4c, float q4ym, float q5c,
                                      float q5ym, float q5zm, float q6c, float q6zm,
                                      float eta_o, float chi_o, float w1o, float w2o,
                                      float w3o, POut<float> &o)
{
    PrimalIn<float> pi;
    pi.nu = Star<float>{nuc, nxm, nxp, nym, nyp, nzm, nzp};
    pi.p1c = p1c; pi.p1xm = p1xm;
    pi.p2c = p2c; pi.p2ym = p2ym;
    pi.p3c = p3c; pi.p3zm = p3zm;
    o.eta_o = eta_o;
    o.chi_o = chi_o;
    primal_eta_chi(pc, pi, o.eta_o, o.chi_o, o.eta, o.chi);
    o.eb = extrap(o.eta, o.eta_o);
    o.cb = extrap(o.chi, o.chi_o);
    if constexpr (TGV) {
        if (wq) {
            pi.q1c = q1c; pi.q1xm = q1xm;
            pi.q2c = q2c; pi.q2xm = q2xm; pi.q2ym = q2ym;
            pi.q3c = q3c; pi.q3xm = q3xm; pi.q3zm = q3zm;
            pi.q4c = q4c; pi.q4ym = q4ym;
            pi.q5c = q5c; pi.q5ym = q5ym; pi.q5zm = q5zm;
            pi.q6c = q6c; pi.q6zm = q6zm;
            o.w1 = w1o; o.w2 = w2o; o.w3 = w3o;
            primal_w(pc, pi, o.w1, o.w2, o.w3);
            o.b1 = extrap(o.w1, w1o);
            o.b2 = extrap(o.w2, w2o);
            o.b3 = extrap(o.w3, w3o);
        }
    }
}

// Primal update of pair column q at plane kp (as pcol, both elements);
// dva/dvb get the elements' dual-side values, pa/pb whether each element
// was primal-active.
template <bool TGV>
__device__ __forceinline__ void pcol2(const Ctx<float> &c, const PrimalCoef<float> &pc,
                                      int kp, bool wq, ColQ2 &q, DualVox<float> &dva,
                                      DualVox<float> &dvb, bool &pa, bool &pb, POut2 &o)
{
    using I = Ix<TGV>;
    const uint8_t ma = q.ma, mb = q.mb;
    uint8_t na = 0, nb = 0;
    F2 np{0.f, 0.f};
    dva = DualVox<float>{};
    dvb = DualVox<float>{};
    o = POut2{};
    pa = pb = false;
    if (q.l0 >= 0) {
        const int l = q.l0 + c.sk * kp;
        const uint32_t vb = uint32_t(l) * 4u;
        if (kp + 1 < c.nz) {
            const uint16_t mn = bld_u16(c.mk, uint32_t(l + c.sk));
            na = uint8_t(mn & 0xffu);
            nb = uint8_t(mn >> 8);
            // nu is zero wherever the dual is inactive: no need to wait
            // for the mask
            np = bld2(c.in, vb + c.skb, c.so(I::NU));
        }
        F2 p1{}, p2{}, p3{}, q1{}, q2{}, q3{}, q4{}, q5{}, q6{};
        if ((ma | mb) & 3) {
            p1 = bld2(c.in, vb, c.so(I::P1));
            p2 = bld2(c.in, vb, c.so(I::P2));
            p3 = bld2(c.in, vb, c.so(I::P3));
            if constexpr (TGV) {
                if (wq) {
                    q1 = bld2(c.in, vb, c.so(I::Q1));
                    q2 = bld2(c.in, vb, c.so(I::Q2));
                    q3 = bld2(c.in, vb, c.so(I::Q3));
                    q4 = bld2(c.in, vb, c.so(I::Q4));
                    q5 = bld2(c.in, vb, c.so(I::Q5));
                    q6 = bld2(c.in, vb, c.so(I::Q6));
                }
            }
        }
        if (ma & 3)
            dva = DualVox<float>{q.nzc.a, p1.a, p2.a, p3.a, q1.a, q2.a, q3.a, q4.a, q5.a, q6.a};
        if (mb & 3)
            dvb = DualVox<float>{q.nzc.b, p1.b, p2.b, p3.b, q1.b, q2.b, q3.b, q4.b, q5.b, q6.b};
        pa = (ma & 1) != 0;
        pb = (mb & 1) != 0;
        if (pa || pb) {
            // an active element is interior; the outer x loads of an
            // inactive element are discarded (out-of-range offsets read 0)
            const uint32_t vym = vb - c.syb, vyp = vb + c.syb;
            const F2 nym = bld2(c.in, vym, c.so(I::NU));
            const F2 nyp = bld2(c.in, vyp, c.so(I::NU));
            const float nxm = bld<float>(c.in, vb - 4u, c.so(I::NU));
            const float nxp = bld<float>(c.in, vb + 8u, c.so(I::NU));
            const float p1xm = bld<float>(c.in, vb - 4u, c.so(I::P1));
            const F2 p2ym = bld2(c.in, vym, c.so(I::P2));
            float q1xm = 0.f, q2xm = 0.f, q3xm = 0.f;
            F2 q2ym{}, q4ym{}, q5ym{}, w1{}, w2{}, w3{};
            if constexpr (TGV) {
                if (wq) {
                    q1xm = bld<float>(c.in, vb - 4u, c.so(I::Q1));
                    q2xm = bld<float>(c.in, vb - 4u, c.so(I::Q2));
                    q3xm = bld<float>(c.in, vb - 4u, c.so(I::Q3));
                    q2ym = bld2(c.in, vym, c.so(I::Q2));
                    q4ym = bld2(c.in, vym, c.so(I::Q4));
                    q5ym = bld2(c.in, vym, c.so(I::Q5));
                    w1 = bld2(c.in, vb, c.so(I::W1));
                    w2 = bld2(c.in, vb, c.so(I::W2));
                    w3 = bld2(c.in, vb, c.so(I::W3));
                }
            }
            const F2 eo = bld2(c.in, vb, c.so(I::ETA));
            const F2 co = bld2(c.in, vb, c.so(I::CHI));
            if (pa)
                pelem<TGV>(pc, wq, q.nzc.a, nxm, q.nzc.b, nym.a, nyp.a, q.nzm.a, np.a,
                           p1.a, p1xm, p2.a, p2ym.a, p3.a, q.p3m.a,
                           q1.a, q1xm, q2.a, q2xm, q2ym.a, q3.a, q3xm, q.q3m.a,
                           q4.a, q4ym.a, q5.a, q5ym.a, q.q5m.a, q6.a, q.q6m.a,
                           eo.a, co.a, w1.a, w2.a, w3.a, o.a);
            if (pb)
                pelem<TGV>(pc, wq, q.nzc.b, q.nzc.a, nxp, nym.b, nyp.b, q.nzm.b, np.b,
                           p1.b, p1.a, p2.b, p2ym.b, p3.b, q.p3m.b,
                           q1.b, q1.a, q2.b, q2.a, q2ym.b, q3.b, q3.a, q.q3m.b,
                           q4.b, q4ym.b, q5.b, q5ym.b, q.q5m.b, q6.b, q.q6m.b,
                           eo.b, co.b, w1.b, w2.b, w3.b, o.b);
        }
    }
    q.ma = na;
    q.mb = nb;
    q.nzm = q.nzc;
    q.nzc = np;
    q.p3m = F2{dva.p3, dvb.p3};
    if constexpr (TGV) {
        q.q3m = F2{dva.q3, dvb.q3};
        q.q5m = F2{dva.q5, dvb.q5};
        q.q6m = F2{dva.q6, dvb.q6};
    }
}

// Loads a pair column's queues for a march whose first primal plane is kp0.
template <bool TGV>
__device__ __forceinline__ void col_init2(const Ctx<float> &c, int kp0, bool wq, ColQ2 &q)
{
    using I = Ix<TGV>;
    q.ma = q.mb = 0;
    q.nzm = q.nzc = q.p3m = q.q3m = q.q5m = q.q6m = F2{0.f, 0.f};
    if (q.l0 < 0)
        return;
    const int l = q.l0 + c.sk * kp0;
    const uint32_t vb = uint32_t(l) * 4u;
    const uint16_t m = bld_u16(c.mk, uint32_t(l));
    q.ma = uint8_t(m & 0xffu);
    q.mb = uint8_t(m >> 8);
    if ((q.ma | q.mb) & 2) {
        const F2 v = bld2(c.in, vb, c.so(I::NU));
        q.nzc = F2{(q.ma & 2) ? v.a : 0.f, (q.mb & 2) ? v.b : 0.f};
    }
    if (kp0 >= 1) {
        const uint16_t mm = bld_u16(c.mk, uint32_t(l - c.sk));
        const uint8_t ba = uint8_t(mm & 0xffu), bb = uint8_t(mm >> 8);
        const uint32_t vm = vb - c.skb;
        if ((ba | bb) & 2) {
            const F2 v = bld2(c.in, vm, c.so(I::NU));
            q.nzm = F2{(ba & 2) ? v.a : 0.f, (bb & 2) ? v.b : 0.f};
        }
        if ((ba | bb) & 3) {
            const F2 v = bld2(c.in, vm, c.so(I::P3));
            q.p3m = F2{(ba & 3) ? v.a : 0.f, (bb & 3) ? v.b : 0.f};
            if constexpr (TGV) {
                if (wq) {
                    const F2 a3 = bld2(c.in, vm, c.so(I::Q3));
                    const F2 a5 = bld2(c.in, vm, c.so(I::Q5));
                    const F2 a6 = bld2(c.in, vm, c.so(I::Q6));
                    q.q3m = F2{(ba & 3) ? a3.a : 0.f, (bb & 3) ? a3.b : 0.f};
                    q.q5m = F2{(ba & 3) ? a5.a : 0.f, (bb & 3) ? a5.b : 0.f};
                    q.q6m = F2{(ba & 3) ? a6.a : 0.f, (bb & 3) ? a6.b : 0.f};
                }
            }
        }
    }
}

// LDS of one block: two plane slots of the extrapolations on the tile plus
// its ring, rows of TX pairs plus one pair on each side.
template <bool TGV, int TX, int TY>
struct PairLds {
    static constexpr int W = 2 * TX + 4;   // floats per row (even: pairs stay 8-byte aligned)
    alignas(16) float a[2][TGV ? 5 : 2][W * (TY + 2)];
};

__device__ __forceinline__ void lds_put2(float *row, int p, float va, float vb)
{
    *reinterpret_cast<F2 *>(row + 2 * p) = F2{va, vb};
}

__device__ __forceinline__ F2 lds_get2(const float *row, int p)
{
    return *reinterpret_cast<const F2 *>(row + 2 * p);
}

// One tile of TX pairs x TY rows of z-chunk bz: primal of iteration n, dual
// of n+1 (as fused_tile).
template <bool TGV, int TX, int TY, bool CHECK>
__device__ __forceinline__ void fused_tile_pair(PairLds<TGV, TX, TY> &L, const Ctx<float> &c,
                                                const DualCoef<float> &dc,
                                                const PrimalCoef<float> &pc, int nx, int ny,
                                                int nz, int bx, int by, int bz, int kz,
                                                double acc[4])
{
    using I = Ix<TGV>;
    constexpr int NT = TX * TY;
    constexpr int NRING = 2 * TX + 2 * TY;
    static_assert(NRING <= NT, "one ring pair per thread at most");
    constexpr int W = PairLds<TGV, TX, TY>::W;
    auto &lds = L.a;
    const int t = threadIdx.x;
    const int tx = t % TX, ty = t / TX;
    const int i0 = bx * 2 * TX, j0 = by * TY;
    const int k0 = 1 + bz * kz;
    const int k1 = min(k0 + kz, nz - 1);  // owned planes [k0, k1)

    const int oi = i0 + 2 * tx, oj = j0 + ty;
    ColQ2 qo;
    qo.l0 = (oi < nx && oj < ny) ? oi + nx * oj : -1;
    ColQ2 qr;
    int rx = 0, ry = 0;
    qr.l0 = -1;
    if (t < NRING) {
        ring_pos<TX, TY>(t, rx, ry);    // rx in pairs: -1 .. TX
        const int ri = i0 + 2 * rx, rj = j0 + ry;
        if (ri >= 0 && ri < nx && rj >= 0 && rj < ny)
            qr.l0 = ri + nx * rj;
    }
    const int po = tx + 1;                 // own pair in the LDS row
    const int lro = (ty + 1) * W;          // own row offset
    const int pr = rx + 1;
    const int lrr = (ry + 1) * W;
    const bool rplus = rx == TX || ry == TY;

    col_init2<TGV>(c, k0 - 1, true, qo);
    if (t < NRING)
        col_init2<TGV>(c, k0 - 1, rplus, qr);

    F2 cbm{0.f, 0.f}, cbc{0.f, 0.f}, ebm{0.f, 0.f}, ebc{0.f, 0.f};
    F2 wc1{0.f, 0.f}, wc2{0.f, 0.f}, wc3{0.f, 0.f};
    DualVox<float> dpa{}, dpb{};
    uint8_t mpa = 0, mpb = 0;
    F2 fprev{0.f, 0.f};

    for (int kp = k0 - 1; kp <= k1; ++kp) {
        const int slot = kp & 1;
        const uint8_t ma = qo.ma, mb = qo.mb;
        F2 fcur{0.f, 0.f};
        if (kp >= k0 && kp < k1 && ((ma | mb) & 2))
            fcur = bld2(c.f, uint32_t(qo.l0 + c.sk * kp) * 4u, 0);
        DualVox<float> dva, dvb;
        bool pa, pb;
        POut2 o;
        pcol2<TGV>(c, pc, kp, true, qo, dva, dvb, pa, pb, o);
        lds_put2(&lds[slot][0][lro], po, o.a.cb, o.b.cb);
        lds_put2(&lds[slot][1][lro], po, o.a.eb, o.b.eb);
        if constexpr (TGV) {
            lds_put2(&lds[slot][2][lro], po, o.a.b1, o.b.b1);
            lds_put2(&lds[slot][3][lro], po, o.a.b2, o.b.b2);
            lds_put2(&lds[slot][4][lro], po, o.a.b3, o.b.b3);
        }
        if ((pa || pb) && kp >= k0 && kp < k1) {
            const uint32_t vb = uint32_t(qo.l0 + c.sk * kp) * 4u;
            bst_sel(F2{o.a.chi, o.b.chi}, pa, pb, c.out, vb, c.so(I::CHI));
            bst_sel(F2{o.a.eta, o.b.eta}, pa, pb, c.out, vb, c.so(I::ETA));
            if constexpr (TGV) {
                bst_sel(F2{o.a.w1, o.b.w1}, pa, pb, c.out, vb, c.so(I::W1));
                bst_sel(F2{o.a.w2, o.b.w2}, pa, pb, c.out, vb, c.so(I::W2));
                bst_sel(F2{o.a.w3, o.b.w3}, pa, pb, c.out, vb, c.so(I::W3));
            }
            if constexpr (CHECK) {
                if (pa) conv_terms(o.a.chi, o.a.chi_o, o.a.eta, o.a.eta_o, acc);
                if (pb) conv_terms(o.b.chi, o.b.chi_o, o.b.eta, o.b.eta_o, acc);
            }
        }
        if (t < NRING) {
            DualVox<float> ra, rb;
            bool rpa, rpb;
            POut2 ro;
            pcol2<TGV>(c, pc, kp, rplus, qr, ra, rb, rpa, rpb, ro);
            lds_put2(&lds[slot][0][lrr], pr, ro.a.cb, ro.b.cb);
            lds_put2(&lds[slot][1][lrr], pr, ro.a.eb, ro.b.eb);
            if constexpr (TGV) {
                lds_put2(&lds[slot][2][lrr], pr, ro.a.b1, ro.b.b1);
                lds_put2(&lds[slot][3][lrr], pr, ro.a.b2, ro.b.b2);
                lds_put2(&lds[slot][4][lrr], pr, ro.a.b3, ro.b.b3);
            }
        }

        // ---- dual at plane kd = kp-1 from the other slot
        const int kd = kp - 1;
        const bool da = (mpa & 2) != 0, db = (mpb & 2) != 0;
        if (kd >= k0 && kd < k1 && (da || db)) {
            const int ps = slot ^ 1;
            const uint32_t vb = uint32_t(qo.l0 + c.sk * kd) * 4u;
            const float *C = lds[ps][0], *E = lds[ps][1];
            // x neighbours outside the pair: element b of pair po-1 and
            // element a of pair po+1
            const float cxm = C[lro + 2 * po - 1], cxp = C[lro + 2 * po + 2];
            const float exm = E[lro + 2 * po - 1], exp_ = E[lro + 2 * po + 2];
            const F2 cym = lds_get2(C + lro - W, po), cyp = lds_get2(C + lro + W, po);
            const F2 eym = lds_get2(E + lro - W, po), eyp = lds_get2(E + lro + W, po);
            DualVox<float> ua = dpa, ub = dpb;
            const Star<float> csa{cbc.a, cxm, cbc.b, cym.a, cyp.a, cbm.a, o.a.cb};
            const Star<float> csb{cbc.b, cbc.a, cxp, cym.b, cyp.b, cbm.b, o.b.cb};
            const Star<float> esa{ebc.a, exm, ebc.b, eym.a, eyp.a, ebm.a, o.a.eb};
            const Star<float> esb{ebc.b, ebc.a, exp_, eym.b, eyp.b, ebm.b, o.b.eb};
            if (da) ua.nu = dual_nu(dc, ua.nu, esa, csa, fprev.a);
            if (db) ub.nu = dual_nu(dc, ub.nu, esb, csb, fprev.b);
            if constexpr (TGV) {
                const float *B1 = lds[ps][2], *B2 = lds[ps][3], *B3 = lds[ps][4];
                const F2 v1y = lds_get2(B1 + lro + W, po);
                const F2 v2y = lds_get2(B2 + lro + W, po);
                const F2 v3y = lds_get2(B3 + lro + W, po);
                const float v1x = B1[lro + 2 * po + 2];
                const float v2x = B2[lro + 2 * po + 2];
                const float v3x = B3[lro + 2 * po + 2];
                if (da)
                    tgv_dual_pq(dc, ua, csa, wc1.a, wc1.b, v1y.a, o.a.b1,
                                wc2.a, wc2.b, v2y.a, o.a.b2, wc3.a, wc3.b, v3y.a, o.a.b3);
                if (db)
                    tgv_dual_pq(dc, ub, csb, wc1.b, v1x, v1y.b, o.b.b1,
                                wc2.b, v2x, v2y.b, o.b.b2, wc3.b, v3x, v3y.b, o.b.b3);
                bst_sel(F2{ua.q1, ub.q1}, da, db, c.out, vb, c.so(I::Q1));
                bst_sel(F2{ua.q2, ub.q2}, da, db, c.out, vb, c.so(I::Q2));
                bst_sel(F2{ua.q3, ub.q3}, da, db, c.out, vb, c.so(I::Q3));
                bst_sel(F2{ua.q4, ub.q4}, da, db, c.out, vb, c.so(I::Q4));
                bst_sel(F2{ua.q5, ub.q5}, da, db, c.out, vb, c.so(I::Q5));
                bst_sel(F2{ua.q6, ub.q6}, da, db, c.out, vb, c.so(I::Q6));
            } else {
                if (da) tv_dual_p(dc, ua, csa);
                if (db) tv_dual_p(dc, ub, csb);
            }
            bst_sel(F2{ua.nu, ub.nu}, da, db, c.out, vb, c.so(I::NU));
            bst_sel(F2{ua.p1, ub.p1}, da, db, c.out, vb, c.so(I::P1));
            bst_sel(F2{ua.p2, ub.p2}, da, db, c.out, vb, c.so(I::P2));
            bst_sel(F2{ua.p3, ub.p3}, da, db, c.out, vb, c.so(I::P3));
        }
        cbm = cbc; cbc = F2{o.a.cb, o.b.cb};
        ebm = ebc; ebc = F2{o.a.eb, o.b.eb};
        wc1 = F2{o.a.b1, o.b.b1};
        wc2 = F2{o.a.b2, o.b.b2};
        wc3 = F2{o.a.b3, o.b.b3};
        dpa = dva;
        dpb = dvb;
        mpa = ma;
        mpb = mb;
        fprev = fcur;
        __syncthreads();
    }
}

}  // namespace fdev
}  // namespace qsm
