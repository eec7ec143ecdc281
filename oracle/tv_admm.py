"""CPU restatement of src/inversion/tv.m (total-variation dipole inversion by
ADMM) and src/inversion/rts.m (rapid two-step inversion: LSMR in k-space,
then the same ADMM), numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and by bench.py's CPU
baseline leg), never by the product package.  It follows tv.m line by line
with MATLAB's precision rules: arrays keep the class of f (float32 or
float64); a double scalar or a double array combined with a single array is
converted to single first.  The gradients are the oracle's C restatement of
gradfp_mex / gradfp_adj_mex (bitwise the reference's).  The FFTs are
numpy's (pocketfft, computed in double and rounded to the class of f):
MATLAB's FFT is closed, so parity with it is pinned by numpy and by the
closed forms in tests/test_oracle_tv.py, not by an import.  Sums are numpy's
(pairwise) in the class of f.
"""
from __future__ import annotations

import numpy as np

from . import oracle as orc

DEFAULTS = dict(mu=7.5e3, tv=1, rtol=1e-1, atol=1e-4, maxit=20, verbose=0, rho=9.0,
                alpha=1.5, rtau=0.7, stau=1e2, gamma=3.0)  # tv.m:38-86


def _cplx(T):
    return np.complex64 if T == np.float32 else np.complex128


def fft3(x, T):
    """fft3.m:5 -- fftn, in the class of x."""
    return np.fft.fftn(x).astype(_cplx(T))


def ifft3(x, T):
    """ifft3.m:5 -- ifftn (1/M included), in the class of x."""
    return np.fft.ifftn(x).astype(_cplx(T))


def dipole_kernel_k(sz, vsz, bdir, T):
    """dipoleKernel.m:56-66 (k-space branch) with dipoleGrid (:74-87):
    grid -low:h:high per axis at h = 1/(n vsz), then ifftshift, D(1) = 0,
    computed in double and cast to T."""
    sz = tuple(int(n) for n in sz)
    vsz = np.broadcast_to(np.asarray(vsz, dtype=np.float64), (3,))
    b = np.asarray(bdir, dtype=np.float64)
    if np.linalg.norm(b) != 1:
        b = b / np.linalg.norm(b)
    axes = []
    for n, h in zip(sz, vsz):
        dk = 1.0 / (n * h)
        low = n // 2
        axes.append((np.arange(n) - low) * dk)
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    kz = b[0] * X + b[1] * Y + b[2] * Z
    k2 = X * X + Y * Y + Z * Z
    with np.errstate(divide="ignore", invalid="ignore"):
        D = 1.0 / 3.0 - (kz * kz / k2)
    D = np.fft.ifftshift(D)
    D[0, 0, 0] = 0.0
    return np.asfortranarray(D.astype(T))


def laplace_kernel(vsz):
    """laplaceKernel.m:14-24: the 7-point stencil, 3x3x3, double."""
    vsz = np.broadcast_to(np.asarray(vsz, dtype=np.float64), (3,))
    ihx2, ihy2, ihz2 = 1.0 / vsz[0] ** 2, 1.0 / vsz[1] ** 2, 1.0 / vsz[2] ** 2
    ihh2 = -2 * (ihx2 + ihy2 + ihz2)
    L = np.zeros((3, 3, 3))
    L[1, 1, 0] = ihz2
    L[1, 1, 2] = ihz2
    L[0, 1, 1] = ihx2
    L[2, 1, 1] = ihx2
    L[1, 0, 1] = ihy2
    L[1, 2, 1] = ihy2
    L[1, 1, 1] = ihh2
    return L


def laplace_symbol(sz, vsz):
    """tv.m:163-165: real(fft3(circshift(padarray(-laplaceKernel(vsz),
    sz - 3, 'post'), -[1, 1, 1]))), double, with L(0) = 0 exactly.

    The stencil sums to zero, so L(0) is 0; the transform leaves a rounding
    residue there when the voxel is not cubic (1/h^2 terms that do not
    cancel exactly).  In the reference iA(0) = 1/(K(0) + rho L(0)) is then
    a huge finite number instead of the Inf that tv.m:168 sets to 0, and
    the mean of u becomes that number times the rounding residue of b(0)
    (the sum of a divergence): noise, not a value to reproduce.  The
    restatement takes the exact L(0) = 0, i.e. the mean of u is 0, as it is
    in the reference whenever the residue vanishes (cubic voxels)."""
    L = -laplace_kernel(vsz)
    P = np.zeros(tuple(int(n) for n in sz))
    P[:3, :3, :3] = L
    P = np.roll(P, (-1, -1, -1), axis=(0, 1, 2))
    S = np.real(np.fft.fftn(P))
    S[0, 0, 0] = 0.0
    return S


def grad_(u, h):
    """tv.m:313-319: gradfp_mex."""
    return orc.gradfp(u, h)


def grad_adj_(x, y, z, h):
    """tv.m:336-339: gradfp_adj_mex."""
    return orc.gradfp_adj(x, y, z, h)


def inv_a(K, L, rho):
    """tv.m:167-168 / rts.m:158-159: iA = 1 ./ (K + rho*L), non-finite -> 0,
    in the class of K (rho*L is double, converted to it)."""
    T = K.dtype.type
    with np.errstate(divide="ignore", invalid="ignore"):
        iA = T(1) / (K + (rho * L).astype(T))
    iA[~np.isfinite(iA)] = 0
    return iA


def u_step(F, iA, v, lm, rho, vsz, T):
    """tv.m:196-203: the u-subproblem."""
    r1 = T(rho) * v[0] - lm[0]
    r2 = T(rho) * v[1] - lm[1]
    r3 = T(rho) * v[2] - lm[2]
    r1 = grad_adj_(r1, r2, r3, vsz)
    b = F + fft3(r1, T)
    return np.asfortranarray(np.real(ifft3(iA * b, T)).astype(T))


def shrink(r1, r2, r3, irho, mtv, T):
    """tv.m:226-242: anisotropic (mtv == 1) or isotropic soft threshold."""
    irho = T(irho)
    if mtv == 1:
        return tuple(np.maximum(np.abs(r) - irho, T(0)) * np.sign(r) for r in (r1, r2, r3))
    s = np.sqrt(r1 * r1 + r2 * r2 + r3 * r3)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = (s - irho) * (s > irho).astype(T) / s
    s[~np.isfinite(s)] = 0
    return r1 * s, r2 * s, r3 * s


def _sum(a):
    return float(np.sum(a))


def _norm2(a):
    return float(np.linalg.norm(np.ravel(a)))


def tv_(f, mask, D, vsz, opts):
    """tv.m:107-310 for one echo.  Returns (x, trace) with trace rows
    (||r||, eps_pri, ||s||, eps_dual, rho after the update)."""
    T = f.dtype.type
    f = np.asfortranarray(f)
    mu = opts["mu"]
    F = (T(mu) * np.conj(D)) * fft3(f, T)       # tv.m:159
    K = T(mu) * np.abs(D) ** 2                  # tv.m:161
    return admm_(f, F, K, mask, vsz, opts)


def admm_(u0, F, K, mask, vsz, opts):
    """The ADMM loop shared by tv.m:141-308 and rts.m:145-308: u starts at
    u0, F is the data term's spectrum, K its k-space weight (iA = 1/(K +
    rho L))."""
    T = u0.dtype.type
    N = u0.size
    rtol, atol = opts["rtol"], opts["atol"]
    atolp = np.sqrt(3 * N) * atol
    atold = np.sqrt(N) * atol
    rho, alpha = opts["rho"], opts["alpha"]
    rtau, stau, gamma = opts["rtau"], opts["stau"], opts["gamma"]
    mtv, maxit = opts["tv"], opts["maxit"]
    irho = 1 / rho
    f = u0
    L = laplace_symbol(u0.shape, vsz)
    iA = inv_a(K, L, rho)

    u = f
    lm = [np.zeros(f.shape, T, order="F") for _ in range(3)]
    v = list(grad_(u, vsz))
    nr = _sum(np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]))
    flag = 0
    trace = []
    a, oma = T(alpha), T(1 - alpha)
    for ii in range(1, maxit + 1):
        u = u_step(F, iA, v, lm, rho, vsz, T)
        v_ = [c.copy() for c in v]
        dx, dy, dz = grad_(u, vsz)
        dx1 = a * dx + oma * v[0]
        dy1 = a * dy + oma * v[1]
        dz1 = a * dz + oma * v[2]
        r1 = dx1 + T(irho) * lm[0]
        r2 = dy1 + T(irho) * lm[1]
        r3 = dz1 + T(irho) * lm[2]
        v = list(shrink(r1, r2, r3, irho, mtv, T))
        lm[0] = lm[0] + T(rho) * (dx1 - v[0])
        lm[1] = lm[1] + T(rho) * (dy1 - v[1])
        lm[2] = lm[2] + T(rho) * (dz1 - v[2])
        nr_ = nr
        e1, e2, e3 = dx - v[0], dy - v[1], dz - v[2]
        s = T(rho) * grad_adj_(v[0] - v_[0], v[1] - v_[1], v[2] - v_[2], vsz)
        nr = _sum(np.sqrt(e1 * e1 + e2 * e2 + e3 * e3))
        ns = _norm2(s)
        na = _sum(np.sqrt(dx * dx + dy * dy + dz * dz))
        nb = _sum(np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]))
        ep = atolp + rtol * max(na, nb)
        ed = atold + rtol * _norm2(T(rho) * grad_adj_(lm[0], lm[1], lm[2], vsz))
        if nr < ep and ns < ed:
            trace.append((nr, ep, ns, ed, rho))
            break
        if nr > stau * ns or (nr > rtau * nr_ and flag < 3):
            rho = gamma * rho
            irho = 1 / rho
            iA = inv_a(K, L, rho)
            if flag > 5:
                flag = 0
        elif nr < ns / stau:
            rho = rho / gamma ** 2
            irho = 1 / rho
            iA = inv_a(K, L, rho)
            flag = flag + 1
        trace.append((nr, ep, ns, ed, rho))
    x = np.asfortranarray((np.asarray(mask) != 0).astype(T) * u)
    return x, np.array(trace)


def lsmr(d, b, atol, btol, itnlim, conlim=1e8):
    """third_party/lsmr/lsmr.m (Fong and Saunders' LSMR, lambda = 0, no
    local reorthogonalisation) for the diagonal operator A = A' = diag(d)
    and the right-hand side b, vectors in the class of b, scalars in double.
    Returns (x, itn, istop)."""
    d = np.asarray(d)
    u = b.copy()
    beta = float(np.linalg.norm(u))
    if beta > 0:
        u = u * (1 / beta)
    v = d * u
    alpha = float(np.linalg.norm(v))
    if alpha > 0:
        v = v * (1 / alpha)
    itn, zetabar, alphabar, rho, rhobar, cbar, sbar = 0, alpha * beta, alpha, 1.0, 1.0, 1.0, 0.0
    h, hbar, x = v.copy(), np.zeros_like(v), np.zeros_like(v)
    betadd, betad, rhodold, tautildeold, thetatilde, zeta, dd = beta, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0
    normA2, maxrbar, minrbar, normb = alpha ** 2, 0.0, 1e100, beta
    ctol = 1 / conlim if conlim > 0 else 0.0
    istop = 0
    if alpha * beta == 0:
        return x, 0, 0
    while itn < itnlim:
        itn += 1
        u = d * v - alpha * u
        beta = float(np.linalg.norm(u))
        if beta > 0:
            u = u * (1 / beta)
            v = d * u - beta * v
            alpha = float(np.linalg.norm(v))
            if alpha > 0:
                v = v * (1 / alpha)
        alphahat, chat, shat = alphabar, 1.0, 0.0
        rhoold = rho
        rho = float(np.hypot(alphahat, beta))
        c, s = alphahat / rho, beta / rho
        thetanew = s * alpha
        alphabar = c * alpha
        rhobarold, zetaold = rhobar, zeta
        thetabar, rhotemp = sbar * rho, cbar * rho
        rhobar = float(np.hypot(cbar * rho, thetanew))
        cbar, sbar = cbar * rho / rhobar, thetanew / rhobar
        zeta = cbar * zetabar
        zetabar = -sbar * zetabar
        hbar = h - (thetabar * rho / (rhoold * rhobarold)) * hbar
        x = x + (zeta / (rho * rhobar)) * hbar
        h = v - (thetanew / rho) * h
        betaacute, betacheck = chat * betadd, -shat * betadd
        betahat = c * betaacute
        betadd = -s * betaacute
        thetatildeold = thetatilde
        rhotildeold = float(np.hypot(rhodold, thetabar))
        ctildeold, stildeold = rhodold / rhotildeold, thetabar / rhotildeold
        thetatilde = stildeold * rhobar
        rhodold = ctildeold * rhobar
        betad = -stildeold * betad + ctildeold * betahat
        tautildeold = (zetaold - thetatildeold * tautildeold) / rhotildeold
        taud = (zeta - thetatilde * tautildeold) / rhodold
        dd = dd + betacheck ** 2
        normr = np.sqrt(dd + (betad - taud) ** 2 + betadd ** 2)
        normA2 = normA2 + beta ** 2
        normA = np.sqrt(normA2)
        normA2 = normA2 + alpha ** 2
        maxrbar = max(maxrbar, rhobarold)
        if itn > 1:
            minrbar = min(minrbar, rhobarold)
        condA = max(maxrbar, rhotemp) / min(minrbar, rhotemp)
        normAr = abs(zetabar)
        normx = float(np.linalg.norm(x))
        test1, test2, test3 = normr / normb, normAr / (normA * normr), 1 / condA
        t1 = test1 / (1 + normA * normx / normb)
        rtol = btol + atol * normA * normx / normb
        if itn >= itnlim:
            istop = 7
        if 1 + test3 <= 1:
            istop = 6
        if 1 + test2 <= 1:
            istop = 5
        if 1 + t1 <= 1:
            istop = 4
        if test3 <= ctol:
            istop = 3
        if test2 <= atol:
            istop = 2
        if test1 <= rtol:
            istop = 1
        if istop > 0:
            break
    return x, itn, istop


RTS_DEFAULTS = dict(tolls=4, mu=1e5, delta=0.15, tv=1, rtol=1e-1, atol=1e-2, maxit=20,
                    verbose=0, rho=10.0, alpha=1.5, rtau=0.7, stau=1e2, gamma=3.0)  # rts.m:50-102


def rts_(f, mask, D, vsz, opts):
    """rts.m:121-308 for one echo.  Returns (x, trace)."""
    T = f.dtype.type
    f = np.asfortranarray(f)
    eps = float(np.finfo(T).eps)
    # step 1: well-conditioned (rts.m:128-139)
    x1, _, _ = lsmr(D.reshape(-1, order="F"), fft3(f, T).reshape(-1, order="F"), eps, eps,
                    opts["tolls"])
    x1 = x1.reshape(f.shape, order="F")
    m = (np.asarray(mask) != 0).astype(T)
    x1 = np.asfortranarray(m * np.real(ifft3(x1, T)).astype(T))
    # step 2: ill-conditioned (rts.m:145-159); M = mu*(|D| > delta) is double
    M = opts["mu"] * (np.abs(D) > T(opts["delta"])).astype(np.float64)
    F = (M * fft3(x1, T)).astype(_cplx(T))
    return admm_(x1, F, M, mask, vsz, opts)


def rts(f, mask, vsz, bdir=(0.0, 0.0, 1.0), opts=None):
    """rts.m:1-118: per echo of a 3-D or 4-D f.  Returns (x, traces)."""
    o = dict(RTS_DEFAULTS)
    o.update({k: v for k, v in (opts or {}).items() if v is not None})
    f = np.asarray(f)
    if f.dtype not in (np.float32, np.float64):
        f = f.astype(np.float64)
    T = f.dtype.type
    vsz = np.broadcast_to(np.asarray(vsz, dtype=np.float64), (3,))
    D = dipole_kernel_k(np.shape(mask), vsz, bdir, T)
    if f.ndim == 3:
        return rts_(f, mask, D, vsz, o)
    x = np.zeros(f.shape, T, order="F")
    traces = []
    for t in range(f.shape[3]):
        x[..., t], tr = rts_(np.asfortranarray(f[..., t]), mask, D, vsz, o)
        traces.append(tr)
    return x, traces


def tv(f, mask, vsz, bdir=(0.0, 0.0, 1.0), opts=None):
    """tv.m:1-104: per echo of a 3-D or 4-D f.  Returns (x, traces)."""
    o = dict(DEFAULTS)
    o.update({k: v for k, v in (opts or {}).items() if v is not None})
    f = np.asarray(f)
    if f.dtype not in (np.float32, np.float64):
        f = f.astype(np.float64)
    T = f.dtype.type
    vsz = np.broadcast_to(np.asarray(vsz, dtype=np.float64), (3,))
    D = dipole_kernel_k(np.shape(mask), vsz, bdir, T)
    if f.ndim == 3:
        return tv_(f, mask, D, vsz, o)
    x = np.zeros(f.shape, T, order="F")
    traces = []
    for t in range(f.shape[3]):
        x[..., t], tr = tv_(np.asfortranarray(f[..., t]), mask, D, vsz, o)
        traces.append(tr)
    return x, traces
