// pd_math.h -- per-voxel primal and dual updates of the QSM.m primal-dual
// TV/TGV solver, shared by every kernel schedule (two sweeps, fused single
// pass, batched).  The arithmetic and its association order are those of
// src/inversion/mex/pd_tgv_mex.c and pd_tv_mex.c, so that with FMA
// contraction disabled (-ffp-contract=off) the results are bitwise those of
// the reference's CPU loops.  C evaluates a+b+c as (a+b)+c; the parentheses
// below make that explicit.
#pragma once

#include "common.h"

namespace qsm {

// Coefficients of a dual sweep.  Double path: pd_tgv_mex.c:529-550; single
// path: :367-389 (each computed in double and cast, except lhh and whh which
// the single path forms in float from the cast values).
template <typename T>
struct DualCoef {
    T hx, hy, hz, hx2, hy2, hz2;
    T lhx, lhy, lhz, lhh, whx, why, whz, whh;
    T a0, a1, a0sq, a1sq, sg;
};

// Coefficients of a primal sweep: pd_tgv_mex.c:858-872 / :692-707.
template <typename T>
struct PrimalCoef {
    T hx, hy, hz;
    T lhx, lhy, lhz, lhh, whx, why, whz, whh;
    T tr, tau1;
};

// alpha0/alpha1 are the TGV weights; TV passes (lam, lam) (pd_tv_mex.c:
// 459-464 uses lam for both the projection radius and its square).
template <typename T>
inline DualCoef<T> make_dual_coef(const double h[3], double sigma,
                                  double alpha0, double alpha1)
{
    DualCoef<T> c;
    c.hx = T(sigma / h[0]);
    c.hy = T(sigma / h[1]);
    c.hz = T(sigma / h[2]);
    c.hx2 = T(0.5 * sigma / h[0]);
    c.hy2 = T(0.5 * sigma / h[1]);
    c.hz2 = T(0.5 * sigma / h[2]);
    c.lhx = T(sigma / (h[0] * h[0]));
    c.lhy = T(sigma / (h[1] * h[1]));
    c.lhz = T(sigma / (h[2] * h[2]));
    c.lhh = T(-2.0) * (c.lhx + c.lhy + c.lhz);
    c.whx = T(-sigma / (3.0 * h[0] * h[0]));
    c.why = T(-sigma / (3.0 * h[1] * h[1]));
    c.whz = T(sigma * 2.0 / (3.0 * h[2] * h[2]));
    c.whh = T(-2.0) * (c.whx + c.why + c.whz);
    c.a0 = T(alpha0);
    c.a1 = T(alpha1);
    c.a0sq = T(alpha0 * alpha0);
    c.a1sq = T(alpha1 * alpha1);
    c.sg = T(sigma);
    return c;
}

template <typename T>
inline PrimalCoef<T> make_primal_coef(const double h[3], double tau)
{
    PrimalCoef<T> c;
    c.hx = T(tau / h[0]);
    c.hy = T(tau / h[1]);
    c.hz = T(tau / h[2]);
    c.lhx = T(tau / (h[0] * h[0]));
    c.lhy = T(tau / (h[1] * h[1]));
    c.lhz = T(tau / (h[2] * h[2]));
    c.lhh = T(-2.0) * (c.lhx + c.lhy + c.lhz);
    c.whx = T(-tau / (3.0 * h[0] * h[0]));
    c.why = T(-tau / (3.0 * h[1] * h[1]));
    c.whz = T(tau * 2.0 / (3.0 * h[2] * h[2]));
    c.whh = T(-2.0) * (c.whx + c.why + c.whz);
    c.tr = T(tau);
    c.tau1 = T(1.0 / (1.0 + tau));
    return c;
}

// 7-point neighbourhood: centre and the -/+ neighbours along x, y, z.
template <typename T>
struct Star {
    T c, xm, xp, ym, yp, zm, zp;
};

// Dual-side data of one voxel: nu, p[3], q[6].
template <typename T>
struct DualVox {
    T nu, p1, p2, p3, q1, q2, q3, q4, q5, q6;
};

// nu <- nu + (laplace(eta_) - wave(chi_) + f): pd_tgv_mex.c:575-588 (the
// f term arrives pre-scaled by sigma, :280-283).
template <typename T>
__device__ __forceinline__ T dual_nu(const DualCoef<T> &c, T nu,
                                     const Star<T> &eb, const Star<T> &cb,
                                     T f)
{
    const T lap = ((c.lhh * eb.c + c.lhx * (eb.xm + eb.xp))
                   + c.lhy * (eb.ym + eb.yp))
                  + c.lhz * (eb.zm + eb.zp);
    const T wav = ((c.whh * cb.c + c.whx * (cb.xm + cb.xp))
                   + c.why * (cb.ym + cb.yp))
                  + c.whz * (cb.zm + cb.zp);
    return nu + ((lap - wav) + f);
}

// TGV: p <- P_{|.|<=alpha1}(p + sigma*(grad chi_ - w_)) and
//      q <- P_{|.|<=alpha0}(q + sigma*symgrad w_): pd_tgv_mex.c:590-639.
// v{1,2,3}{c,xp,yp,zp}: w_ components at the voxel and its + neighbours.
template <typename T>
__device__ __forceinline__ void tgv_dual_pq(
    const DualCoef<T> &c, DualVox<T> &d, const Star<T> &cb,
    T v1c, T v1x, T v1y, T v1z, T v2c, T v2x, T v2y, T v2z,
    T v3c, T v3x, T v3y, T v3z)
{
    T a = (d.p1 + c.hx * (cb.xp - cb.c)) - c.sg * v1c;
    T b = (d.p2 + c.hy * (cb.yp - cb.c)) - c.sg * v2c;
    T e = (d.p3 + c.hz * (cb.zp - cb.c)) - c.sg * v3c;
    T x = (a * a + b * b) + e * e;
    if (x > c.a1sq) {
        x = c.a1 / rsqrt_cr(x);
        a = a * x; b = b * x; e = e * x;
    }
    d.p1 = a; d.p2 = b; d.p3 = e;

    T e1 = d.q1 + c.hx * (v1x - v1c);
    T e2 = (d.q2 + c.hx2 * (v2x - v2c)) + c.hy2 * (v1y - v1c);
    T e3 = (d.q3 + c.hx2 * (v3x - v3c)) + c.hz2 * (v1z - v1c);
    T e4 = d.q4 + c.hy * (v2y - v2c);
    T e5 = (d.q5 + c.hy2 * (v3y - v3c)) + c.hz2 * (v2z - v2c);
    T e6 = d.q6 + c.hz * (v3z - v3c);
    x = ((e1 * e1 + e4 * e4) + e6 * e6)
        + T(2.0) * ((e2 * e2 + e3 * e3) + e5 * e5);
    if (x > c.a0sq) {
        x = c.a0 / rsqrt_cr(x);
        e1 = e1 * x; e2 = e2 * x; e3 = e3 * x;
        e4 = e4 * x; e5 = e5 * x; e6 = e6 * x;
    }
    d.q1 = e1; d.q2 = e2; d.q3 = e3; d.q4 = e4; d.q5 = e5; d.q6 = e6;
}

// TV: p <- P_{|.|<=lam}(p + sigma*grad chi_): pd_tv_mex.c:493-506.
template <typename T>
__device__ __forceinline__ void tv_dual_p(const DualCoef<T> &c, DualVox<T> &d,
                                          const Star<T> &cb)
{
    T a = d.p1 + c.hx * (cb.xp - cb.c);
    T b = d.p2 + c.hy * (cb.yp - cb.c);
    T e = d.p3 + c.hz * (cb.zp - cb.c);
    T x = (a * a + b * b) + e * e;
    if (x > c.a1sq) {
        x = c.a1 / rsqrt_cr(x);
        a = a * x; b = b * x; e = e * x;
    }
    d.p1 = a; d.p2 = b; d.p3 = e;
}

// Primal inputs of one voxel: nu's 7-point star, and p, q at the voxel and
// at the backward neighbours their divergences use.
template <typename T>
struct PrimalIn {
    Star<T> nu;
    T p1c, p1xm, p2c, p2ym, p3c, p3zm;
    T q1c, q1xm, q2c, q2xm, q2ym, q3c, q3xm, q3zm;
    T q4c, q4ym, q5c, q5ym, q5zm, q6c, q6zm;
};

// eta, chi updates: pd_tgv_mex.c:914-931 (identical in pd_tv_mex.c:696-713).
// Returns the new values through en/cn; the extrapolations
// eta_ = eta + 1.0*(eta - eta_old) (pd_tgv_mex.c:965-966) are formed by the
// caller from the old and new values.
template <typename T>
__device__ __forceinline__ void primal_eta_chi(const PrimalCoef<T> &c,
                                               const PrimalIn<T> &in,
                                               T eta_old, T chi_old,
                                               T &eta_new, T &chi_new)
{
    const T x = in.nu.xm + in.nu.xp;
    const T y = in.nu.ym + in.nu.yp;
    const T z = in.nu.zm + in.nu.zp;
    eta_new = c.tau1 * ((((eta_old - c.lhh * in.nu.c) - c.lhx * x)
                         - c.lhy * y) - c.lhz * z);
    const T dvp = (c.hx * (in.p1c - in.p1xm) + c.hy * (in.p2c - in.p2ym))
                  + c.hz * (in.p3c - in.p3zm);
    const T wnu = ((c.whh * in.nu.c + c.whx * x) + c.why * y) + c.whz * z;
    chi_new = chi_old + (dvp + wnu);
}

// w update: pd_tgv_mex.c:934-959.  Each component is scaled by that
// component's h (hx for w1 etc.), as the reference does.
template <typename T>
__device__ __forceinline__ void primal_w(const PrimalCoef<T> &c,
                                         const PrimalIn<T> &in,
                                         T &w1, T &w2, T &w3)
{
    w1 = w1 + (c.tr * in.p1c
               + c.hx * (((((in.q1c - in.q1xm) + in.q2c) - in.q2ym) + in.q3c)
                         - in.q3zm));
    w2 = w2 + (c.tr * in.p2c
               + c.hy * (((((in.q2c - in.q2xm) + in.q4c) - in.q4ym) + in.q5c)
                         - in.q5zm));
    w3 = w3 + (c.tr * in.p3c
               + c.hz * (((((in.q3c - in.q3xm) + in.q5c) - in.q5ym) + in.q6c)
                         - in.q6zm));
}

// theta-extrapolation with THETA = 1.0: pd_tgv_mex.c:964-970.
template <typename T>
__device__ __forceinline__ T extrap(T nw, T old)
{
    return nw + T(1.0) * (nw - old);
}

// Convergence summands of one voxel (pd_tgv_mex.c:997-1002 double,
// :1031-1036 single).  s[0] = chi^2, s[1] = eta^2, s[2] = dchi^2,
// s[3] = deta^2, accumulated in double.
__device__ __forceinline__ void conv_terms(double chi, double chi0,
                                           double eta, double eta0,
                                           double s[4])
{
    s[0] += chi * chi;
    s[1] += eta * eta;
    s[2] += (chi - chi0) * (chi - chi0);
    s[3] += (eta - eta0) * (eta - eta0);
}

__device__ __forceinline__ void conv_terms(float chi, float chi0,
                                           float eta, float eta0,
                                           double s[4])
{
    s[0] += double(chi) * double(chi);
    s[1] += double(eta) * double(eta);
    s[2] += double(chi - chi0) * (double(chi) - chi0);
    s[3] += double(eta - eta0) * (double(eta) - eta0);
}

}  // namespace qsm
