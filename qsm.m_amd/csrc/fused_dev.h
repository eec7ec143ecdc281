// fused_dev.h -- device building blocks of the single-pass primal-dual
// kernel: buffer-resource accessors, the per-column primal step, the ring
// layout and the per-tile march (fused_tile).  Shared by the single-volume
// kernel (fused.hip) and the batched persistent kernel (batch.hip).
#pragma once

#include "pd_kernels.h"

namespace qsm {
namespace fdev {


using rsrc_t = __amdgpu_buffer_rsrc_t;

// Raw buffer (stride 0, byte offsets), 32-bit data format: the CDNA
// encoding of the resource's fourth dword.
constexpr int RSRC_DW3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0,
                                             (int)bytes, RSRC_DW3);
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Loads one value; AUX is the instruction's cache-policy operand (0: the
// default, 2: nontemporal).
template <typename T, int AUX = 0>
__device__ __forceinline__ T bld(rsrc_t r, uint32_t vo, uint32_t so)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, AUX));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, AUX));
}

// Cache policy of the state stores (the instruction's aux operand; on
// gfx950 bit 1 is the nontemporal hint).  The out bank is not read again
// in the same pass, so its stores stream past the caches that hold the
// in bank's halo (measured 4-6 % faster at 256x256x176, both precisions).
#ifndef QSM_STORE_AUX
#define QSM_STORE_AUX 2
#endif

__device__ __forceinline__ void bst(float v, rsrc_t r, uint32_t vo, uint32_t so)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vo, (int)so,
                                          QSM_STORE_AUX);
}

__device__ __forceinline__ void bst(double v, rsrc_t r, uint32_t vo, uint32_t so)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)vo, (int)so,
                                          QSM_STORE_AUX);
}

__device__ __forceinline__ uint8_t bld_u8(rsrc_t r, uint32_t vo)
{
    return __builtin_amdgcn_raw_buffer_load_b8(r, (int)vo, 0, 0);
}

// Array indices inside a bank (the handle's layout).
template <bool TGV>
struct Ix;
template <>
struct Ix<true> {
    enum { CHI = 0, ETA, W1, W2, W3, NU, P1, P2, P3, Q1, Q2, Q3, Q4, Q5, Q6, NARR };
};
template <>
struct Ix<false> {
    enum { CHI = 0, ETA, NU, P1, P2, P3, NARR, W1 = 0, W2 = 0, W3 = 0,
           Q1 = 0, Q2 = 0, Q3 = 0, Q4 = 0, Q5 = 0, Q6 = 0 };
};

template <typename T>
struct Ctx {
    rsrc_t in, out, f, mk;
    uint32_t nb;         // N * sizeof(T): byte distance between arrays
    uint32_t syb, skb;   // y / z strides in bytes
    int sk, nz;          // z stride in voxels, planes
    __device__ __forceinline__ uint32_t so(int a) const { return uint32_t(a) * nb; }
};

// A column's own state at one plane (what pcol reads at the centre voxel).
template <typename T>
struct Cen {
    T p1, p2, p3, q1, q2, q3, q4, q5, q6, chi, eta, w1, w2, w3;
};

// The x/y neighbours the primal reads around a primal-active voxel.
template <typename T>
struct Nbr {
    T nxm, nxp, nym, nyp;             // nu at -x, +x, -y, +y
    T p1xm, p2ym;
    T q1xm, q2xm, q2ym, q3xm, q4ym, q5ym;
};

template <typename T, bool TGV>
__device__ __forceinline__ void load_nbr(const Ctx<T> &c, uint32_t vb, uint8_t m, bool wq,
                                         Nbr<T> &x)
{
    using I = Ix<TGV>;
    constexpr uint32_t S = sizeof(T);
    x = Nbr<T>{};
    if (m & 1) {
        // a primal-active voxel is interior: every neighbour offset is in range
        const uint32_t vxm = vb - S, vym = vb - c.syb, vyp = vb + c.syb;
        x.nxm = bld<T>(c.in, vxm, c.so(I::NU));
        x.nxp = bld<T>(c.in, vb + S, c.so(I::NU));
        x.nym = bld<T>(c.in, vym, c.so(I::NU));
        x.nyp = bld<T>(c.in, vyp, c.so(I::NU));
        x.p1xm = bld<T>(c.in, vxm, c.so(I::P1));
        x.p2ym = bld<T>(c.in, vym, c.so(I::P2));
        if constexpr (TGV) {
            if (wq) {
                x.q1xm = bld<T>(c.in, vxm, c.so(I::Q1));
                x.q2xm = bld<T>(c.in, vxm, c.so(I::Q2));
                x.q2ym = bld<T>(c.in, vym, c.so(I::Q2));
                x.q3xm = bld<T>(c.in, vxm, c.so(I::Q3));
                x.q4ym = bld<T>(c.in, vym, c.so(I::Q4));
                x.q5ym = bld<T>(c.in, vym, c.so(I::Q5));
            }
        }
    }
}

template <typename T>
struct ColQ {
    int l0;             // i + nx*j, or -1 when the column is outside the grid
    uint8_t mc;         // mask byte at the current plane kp
    T nzm, nzc;         // nu at kp-1, kp
    T p3m, q3m, q5m, q6m;  // z-backward neighbours (plane kp-1)
};

// Own state at byte offset vb for mask byte m: the dual state where either
// mask bit is set (q only with wq), the primal state where bit 0 is set.
template <typename T, bool TGV>
__device__ __forceinline__ void load_cen(const Ctx<T> &c, uint32_t vb, uint8_t m, bool wq,
                                         Cen<T> &x)
{
    using I = Ix<TGV>;
    x = Cen<T>{};
    if (m & 3) {
        x.p1 = bld<T>(c.in, vb, c.so(I::P1));
        x.p2 = bld<T>(c.in, vb, c.so(I::P2));
        x.p3 = bld<T>(c.in, vb, c.so(I::P3));
        if constexpr (TGV) {
            if (wq) {
                x.q1 = bld<T>(c.in, vb, c.so(I::Q1));
                x.q2 = bld<T>(c.in, vb, c.so(I::Q2));
                x.q3 = bld<T>(c.in, vb, c.so(I::Q3));
                x.q4 = bld<T>(c.in, vb, c.so(I::Q4));
                x.q5 = bld<T>(c.in, vb, c.so(I::Q5));
                x.q6 = bld<T>(c.in, vb, c.so(I::Q6));
            }
        }
    }
    if (m & 1) {
        x.eta = bld<T>(c.in, vb, c.so(I::ETA));
        x.chi = bld<T>(c.in, vb, c.so(I::CHI));
        if constexpr (TGV) {
            if (wq) {
                x.w1 = bld<T>(c.in, vb, c.so(I::W1));
                x.w2 = bld<T>(c.in, vb, c.so(I::W2));
                x.w3 = bld<T>(c.in, vb, c.so(I::W3));
            }
        }
    }
}

template <typename T>
struct POut {
    T chi, eta, w1, w2, w3;   // new iterates
    T cb, eb, b1, b2, b3;     // extrapolations
    T chi_o, eta_o;           // old iterates (convergence terms)
};

// Primal update of column q at plane kp; returns the voxel's dual-side
// values (dv; zero where both mask bits are clear, which is where the dual
// state never leaves zero) and advances the column's queues to kp+1.
// wq = false (TGV ring cells on the -x/-y sides) skips w and q: the dual
// reads w_ only on the + sides.
template <typename T, bool TGV>
__device__ __forceinline__ void pcol(const Ctx<T> &c, const PrimalCoef<T> &pc,
                                     int kp, bool wq, ColQ<T> &q,
                                     DualVox<T> &dv, bool &pact, POut<T> &o)
{
    using I = Ix<TGV>;
    constexpr uint32_t S = sizeof(T);
    const uint8_t mc = q.mc;
    uint8_t mn = 0;
    T np = T(0);
    Cen<T> cur{};
    Nbr<T> nb{};
    dv = DualVox<T>{};
    o = POut<T>{};
    pact = false;
    if (q.l0 >= 0) {
        const int l = q.l0 + c.sk * kp;
        const uint32_t vb = uint32_t(l) * S;
        if (kp + 1 >= 0 && kp + 1 < c.nz) {
            mn = bld_u8(c.mk, uint32_t(l + c.sk));
            // nu is zero wherever the dual is inactive (never written
            // there), so the load need not wait for the mask: one
            // dependent memory round trip less per plane step
            np = bld<T>(c.in, vb + c.skb, c.so(I::NU));
        }
        load_cen<T, TGV>(c, vb, mc, wq, cur);
        if (mc & 3) {
            dv.nu = q.nzc;
            dv.p1 = cur.p1;
            dv.p2 = cur.p2;
            dv.p3 = cur.p3;
            if constexpr (TGV) {
                if (wq) {
                    dv.q1 = cur.q1;
                    dv.q2 = cur.q2;
                    dv.q3 = cur.q3;
                    dv.q4 = cur.q4;
                    dv.q5 = cur.q5;
                    dv.q6 = cur.q6;
                }
            }
        }
        pact = (mc & 1) != 0;
        if (pact) {
            load_nbr<T, TGV>(c, vb, mc, wq, nb);
            PrimalIn<T> pi;
            pi.nu = Star<T>{q.nzc, nb.nxm, nb.nxp, nb.nym, nb.nyp, q.nzm, np};
            pi.p1c = dv.p1; pi.p1xm = nb.p1xm;
            pi.p2c = dv.p2; pi.p2ym = nb.p2ym;
            pi.p3c = dv.p3; pi.p3zm = q.p3m;
            if constexpr (TGV) {
                if (wq) {
                    pi.q1c = dv.q1; pi.q1xm = nb.q1xm;
                    pi.q2c = dv.q2; pi.q2xm = nb.q2xm;
                    pi.q2ym = nb.q2ym;
                    pi.q3c = dv.q3; pi.q3xm = nb.q3xm;
                    pi.q3zm = q.q3m;
                    pi.q4c = dv.q4; pi.q4ym = nb.q4ym;
                    pi.q5c = dv.q5; pi.q5ym = nb.q5ym;
                    pi.q5zm = q.q5m;
                    pi.q6c = dv.q6; pi.q6zm = q.q6m;
                }
            }
            o.eta_o = cur.eta;
            o.chi_o = cur.chi;
            primal_eta_chi(pc, pi, o.eta_o, o.chi_o, o.eta, o.chi);
            o.eb = extrap(o.eta, o.eta_o);
            o.cb = extrap(o.chi, o.chi_o);
            if constexpr (TGV) {
                if (wq) {
                    const T a1 = cur.w1;
                    const T a2 = cur.w2;
                    const T a3 = cur.w3;
                    o.w1 = a1; o.w2 = a2; o.w3 = a3;
                    primal_w(pc, pi, o.w1, o.w2, o.w3);
                    o.b1 = extrap(o.w1, a1);
                    o.b2 = extrap(o.w2, a2);
                    o.b3 = extrap(o.w3, a3);
                }
            }
        }
    }
    q.mc = mn;
    q.nzm = q.nzc;
    q.nzc = np;
    q.p3m = dv.p3;
    if constexpr (TGV) {
        q.q3m = dv.q3;
        q.q5m = dv.q5;
        q.q6m = dv.q6;
    }
}

// Loads a column's queues for a march whose first primal plane is kp0.
template <typename T, bool TGV>
__device__ __forceinline__ void col_init(const Ctx<T> &c, int kp0, bool wq, ColQ<T> &q)
{
    using I = Ix<TGV>;
    constexpr uint32_t S = sizeof(T);
    q.mc = 0;
    q.nzm = q.nzc = q.p3m = q.q3m = q.q5m = q.q6m = T(0);
    // planes outside the volume hold no active voxel (a march may start
    // below plane 0; pcol then loads plane 0 as its next plane)
    if (q.l0 < 0 || kp0 < 0 || kp0 >= c.nz)
        return;
    const int l = q.l0 + c.sk * kp0;
    const uint32_t vb = uint32_t(l) * S;
    q.mc = bld_u8(c.mk, uint32_t(l));
    if (q.mc & 2)
        q.nzc = bld<T>(c.in, vb, c.so(I::NU));
    if (kp0 >= 1) {
        const uint8_t mb = bld_u8(c.mk, uint32_t(l - c.sk));
        const uint32_t vm = vb - c.skb;
        if (mb & 2)
            q.nzm = bld<T>(c.in, vm, c.so(I::NU));
        if (mb & 3) {
            q.p3m = bld<T>(c.in, vm, c.so(I::P3));
            if constexpr (TGV) {
                if (wq) {
                    q.q3m = bld<T>(c.in, vm, c.so(I::Q3));
                    q.q5m = bld<T>(c.in, vm, c.so(I::Q5));
                    q.q6m = bld<T>(c.in, vm, c.so(I::Q6));
                }
            }
        }
    }
}

// Ring cells: the -y and +y rows, then the -x and +x columns; no corners
// (the stencils have no diagonal terms).
template <int TX, int TY>
__device__ __forceinline__ void ring_pos(int r, int &rx, int &ry)
{
    if (r < TX) { rx = r; ry = -1; }
    else if (r < 2 * TX) { rx = r - TX; ry = TY; }
    else if (r < 2 * TX + TY) { rx = -1; ry = r - 2 * TX; }
    else { rx = TX; ry = r - 2 * TX - TY; }
}

// Sums s[0..3] over the block in a fixed order and writes them from thread 0
// (NT, the block size, a multiple of the 64-lane wave).
template <int NT>
__device__ __forceinline__ void block_sum4_f(double s[4], double *dst)
{
    static_assert(NT % 64 == 0, "whole waves only");
    __shared__ double red[NT / 64][4];
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 4; ++q)
            s[q] += __shfl_down(s[q], off, 64);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0)
        for (int q = 0; q < 4; ++q)
            red[tid >> 6][q] = s[q];
    __syncthreads();
    if (tid == 0) {
        double t[4] = {0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < NT / 64; ++w)
            for (int q = 0; q < 4; ++q)
                t[q] += red[w][q];
        for (int q = 0; q < 4; ++q)
            dst[q] = t[q];
    }
}

// LDS of one block: two plane slots of the extrapolations (chi_, eta_ and,
// for TGV, w_[3]) on the tile plus its one-cell ring.
template <typename T, bool TGV, int TX, int TY>
struct FusedLds {
    T a[2][TGV ? 5 : 2][(TX + 2) * (TY + 2)];
};

// One tile (bx, by) of z-chunk bz: primal of iteration n then dual of n+1.
// Convergence terms of the owned primal updates accumulate into acc.
template <typename T, bool TGV, int TX, int TY, bool CHECK>
__device__ __forceinline__ void fused_tile(FusedLds<T, TGV, TX, TY> &L, const Ctx<T> &c,
                                           const DualCoef<T> &dc, const PrimalCoef<T> &pc,
                                           int nx, int ny, int nz, int bx, int by, int bz,
                                           int kz, double acc[4])
{
    using I = Ix<TGV>;
    constexpr uint32_t S = sizeof(T);
    constexpr int NT = TX * TY;
    constexpr int NRING = 2 * TX + 2 * TY;
    static_assert(NRING <= NT, "one ring cell per thread at most");
    constexpr int LW = TX + 2;
    auto &lds = L.a;
    const int t = threadIdx.x;
    const int tx = t % TX, ty = t / TX;
    const int i0 = bx * TX, j0 = by * TY;
    const int k0 = 1 + bz * kz;
    const int k1 = min(k0 + kz, nz - 1);  // owned planes [k0, k1)

    const int oi = i0 + tx, oj = j0 + ty;
    ColQ<T> qo;
    qo.l0 = (oi < nx && oj < ny) ? oi + nx * oj : -1;
    ColQ<T> qr;
    int rx = 0, ry = 0;
    qr.l0 = -1;
    if (t < NRING) {
        ring_pos<TX, TY>(t, rx, ry);
        const int ri = i0 + rx, rj = j0 + ry;
        if (ri >= 0 && ri < nx && rj >= 0 && rj < ny)
            qr.l0 = ri + nx * rj;
    }
    const int lo = (ty + 1) * LW + (tx + 1);
    const int lr = (ry + 1) * LW + (rx + 1);
    const bool rplus = rx == TX || ry == TY;  // +x / +y side of the ring

    col_init<T, TGV>(c, k0 - 1, true, qo);
    if (t < NRING)
        col_init<T, TGV>(c, k0 - 1, rplus, qr);

    // own-column extrapolation queues and the dual-side values one plane back
    T cbm = T(0), cbc = T(0), ebm = T(0), ebc = T(0);
    T wc1 = T(0), wc2 = T(0), wc3 = T(0);
    DualVox<T> dprev{};
    uint8_t mprev = 0;
    T fprev = T(0);  // f at the dual plane, loaded one step ahead

    for (int kp = k0 - 1; kp <= k1; ++kp) {
        const int slot = kp & 1;
        // ---- primal at plane kp: own column and (some threads) a ring cell
        const uint8_t mcur = qo.mc;
        // f of plane kp, consumed by the dual of the next step
        T fcur = T(0);
        if (kp >= k0 && kp < k1 && (mcur & 2))
            fcur = bld<T>(c.f, uint32_t(qo.l0 + c.sk * kp) * S, 0);
        DualVox<T> dv;
        bool pact;
        POut<T> o;
        pcol<T, TGV>(c, pc, kp, true, qo, dv, pact, o);
        lds[slot][0][lo] = o.cb;
        lds[slot][1][lo] = o.eb;
        if constexpr (TGV) {
            lds[slot][2][lo] = o.b1;
            lds[slot][3][lo] = o.b2;
            lds[slot][4][lo] = o.b3;
        }
        if (pact && kp >= k0 && kp < k1) {
            const uint32_t vb = uint32_t(qo.l0 + c.sk * kp) * S;
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
        if (t < NRING) {
            DualVox<T> rdv;
            bool ract;
            POut<T> ro;
            pcol<T, TGV>(c, pc, kp, rplus, qr, rdv, ract, ro);
            lds[slot][0][lr] = ro.cb;
            lds[slot][1][lr] = ro.eb;
            if constexpr (TGV) {
                lds[slot][2][lr] = ro.b1;
                lds[slot][3][lr] = ro.b2;
                lds[slot][4][lr] = ro.b3;
            }
        }

        // ---- dual at plane kd = kp-1 (neighbour extrapolations were
        // written to the other LDS slot in the previous step)
        const int kd = kp - 1;
        if (kd >= k0 && kd < k1 && (mprev & 2)) {
            const int ps = slot ^ 1;
            const uint32_t vb = uint32_t(qo.l0 + c.sk * kd) * S;
            const Star<T> cs{cbc, lds[ps][0][lo - 1], lds[ps][0][lo + 1],
                             lds[ps][0][lo - LW], lds[ps][0][lo + LW], cbm, o.cb};
            const Star<T> es{ebc, lds[ps][1][lo - 1], lds[ps][1][lo + 1],
                             lds[ps][1][lo - LW], lds[ps][1][lo + LW], ebm, o.eb};
            DualVox<T> d = dprev;
            d.nu = dual_nu(dc, d.nu, es, cs, fprev);
            if constexpr (TGV) {
                tgv_dual_pq(dc, d, cs,
                            wc1, lds[ps][2][lo + 1], lds[ps][2][lo + LW], o.b1,
                            wc2, lds[ps][3][lo + 1], lds[ps][3][lo + LW], o.b2,
                            wc3, lds[ps][4][lo + 1], lds[ps][4][lo + LW], o.b3);
                bst(d.q1, c.out, vb, c.so(I::Q1));
                bst(d.q2, c.out, vb, c.so(I::Q2));
                bst(d.q3, c.out, vb, c.so(I::Q3));
                bst(d.q4, c.out, vb, c.so(I::Q4));
                bst(d.q5, c.out, vb, c.so(I::Q5));
                bst(d.q6, c.out, vb, c.so(I::Q6));
            } else {
                tv_dual_p(dc, d, cs);
            }
            bst(d.nu, c.out, vb, c.so(I::NU));
            bst(d.p1, c.out, vb, c.so(I::P1));
            bst(d.p2, c.out, vb, c.so(I::P2));
            bst(d.p3, c.out, vb, c.so(I::P3));
        }
        cbm = cbc; cbc = o.cb;
        ebm = ebc; ebc = o.eb;
        wc1 = o.b1; wc2 = o.b2; wc3 = o.b3;
        dprev = dv;
        mprev = mcur;
        fprev = fcur;
        __syncthreads();
    }

}

}  // namespace fdev
}  // namespace qsm
