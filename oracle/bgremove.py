"""CPU restatement of src/bgremove/sharp.m and vsharp.m (background field
removal by the spherical mean value property), numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and by bench.py's CPU
baseline leg), never by the product package.  It follows the reference
line by line with MATLAB's precision rules (arrays keep the class of f; the
mask is smoothed in single precision, as `fft3(single(mask))` does); the
helpers restate cropIndices.m, padfastfft.m / unpadfastfft.m ('both') and
nextfastfft.m (LTFAT 2.4.0, vendored as third_party/nextfastfft.m: the
smallest 2,3,5,7-smooth integer >= n).  The FFTs are numpy's (pocketfft,
computed in double and rounded to the class of the operands): MATLAB's FFT
is closed, so parity with it is pinned by numpy and by the closed forms in
tests/test_oracle_bgremove.py.
"""
from __future__ import annotations

import numpy as np


def crop_indices(mask):
    """cropIndices.m:7-16: indices (0-based) of the non-empty slices along
    each axis (not necessarily contiguous)."""
    m = np.asarray(mask) != 0
    return (np.nonzero(m.any(axis=(1, 2)))[0], np.nonzero(m.any(axis=(0, 2)))[0],
            np.nonzero(m.any(axis=(0, 1)))[0])


def nextfastfft(n):
    """Smallest 2,3,5,7-smooth integer >= n (n >= 1)."""
    k = int(n)
    while True:
        r = k
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        if r == 1:
            return k
        k += 1


def _pads(sz, padsize):
    """padfastfft.m:37-58 ('both'): leading and trailing pad per axis."""
    pre, post = [], []
    for s, p in zip(sz, padsize):
        pad = 0 if p == 0 else nextfastfft(s + 2 * p) - s
        oddpad, odds = pad & 1, s & 1
        pre.append(pad // 2 + (oddpad & odds))
        post.append(pad // 2 + (oddpad & (1 - odds)))
    return pre, post


def padfastfft(u, padsize):
    """padfastfft.m with direction 'both' and zero padding; u 3-D or 4-D
    (the fourth axis is not padded)."""
    u = np.asarray(u)
    pre, post = _pads(u.shape[:3], padsize)
    width = list(zip(pre, post)) + [(0, 0)] * (u.ndim - 3)
    return np.asfortranarray(np.pad(u, width))


def unpadfastfft(u, sz0, padsize):
    """unpadfastfft.m with direction 'both'."""
    pre, _ = _pads(sz0, padsize)
    sl = tuple(slice(p, p + s) for p, s in zip(pre, sz0))
    return np.asfortranarray(u[sl])


def smv_kernel(sz, r, vsz, T):
    """smvKernel.m: the normalised ball (Y^2 + X^2 + Z^2 <= r^2, that
    order) on the grid -low:vsz:high per axis, centred at floor(sz/2)."""
    vsz = np.broadcast_to(np.asarray(vsz, dtype=np.float64), (3,))
    ax = [(np.arange(n) - n // 2) * h for n, h in zip(sz, vsz)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    h = ((Y * Y + X * X + Z * Z) <= r * r).astype(np.float64)
    h = h / h.sum()
    return h.astype(T)


def _cplx(T):
    return np.complex64 if T == np.float32 else np.complex128


def _smv_symbol(sz, r, vsz, T):
    """real(fft3(ifftshift(smvKernel(sz, r, vsz, T)))) in the class T."""
    h = np.fft.ifftshift(smv_kernel(sz, r, vsz, T))
    return np.real(np.fft.fftn(h).astype(_cplx(T))).astype(T)


def _echoes(f):
    return [f] if f.ndim == 3 else [f[..., t] for t in range(f.shape[3])]


def sharp(f, mask, vsz, r=9, thr=0.05):
    """sharp.m:47-84 -> (fl, eroded mask)."""
    f = np.asarray(f)
    if f.dtype not in (np.float32, np.float64):
        f = f.astype(np.float64)
    T = f.dtype.type
    vsz = np.broadcast_to(np.asarray(vsz, dtype=np.float64), (3,))
    mask = np.asarray(mask) != 0
    fl = np.zeros(f.shape, T, order="F")
    ix, iy, iz = crop_indices(mask)
    pad = [int(np.ceil(r / vsz.min()))] * 3
    fc = f[np.ix_(ix, iy, iz)] if f.ndim == 3 else f[np.ix_(ix, iy, iz, np.arange(f.shape[3]))]
    fp = padfastfft(fc, pad)
    m1 = padfastfft(mask[np.ix_(ix, iy, iz)], pad)
    sz = fp.shape[:3]
    H = _smv_symbol(sz, r, vsz, T)
    # mask1 = real(ifft3(h .* fft3(single(mask1)))) > 0.999
    sm = np.fft.ifftn(H.astype(np.float32) * np.fft.fftn(m1.astype(np.float32)).astype(np.complex64))
    m1 = np.real(sm.astype(np.complex64)) > np.float32(0.999)
    h = T(1) - H
    out = []
    for fe in _echoes(fp):
        g = m1 * np.real(np.fft.ifftn(h * np.fft.fftn(fe).astype(_cplx(T))).astype(_cplx(T)))
        G = np.fft.fftn(g.astype(T)).astype(_cplx(T))
        with np.errstate(divide="ignore", invalid="ignore"):
            G = G / h
        G[np.abs(h) < T(thr)] = 0
        out.append((m1 * np.real(np.fft.ifftn(G).astype(_cplx(T)))).astype(T))
    fo = out[0] if f.ndim == 3 else np.stack(out, axis=3)
    szc = (len(ix), len(iy), len(iz))
    res = unpadfastfft(fo, szc, pad)
    if f.ndim == 3:
        fl[np.ix_(ix, iy, iz)] = res
    else:
        fl[np.ix_(ix, iy, iz, np.arange(f.shape[3]))] = res
    mo = np.zeros(mask.shape, bool, order="F")
    mo[np.ix_(ix, iy, iz)] = unpadfastfft(m1, szc, pad)
    return fl, mo


def vsharp(f, mask, vsz, rs=None, thr=0.05):
    """vsharp.m:47-153 -> (fl, mask).  The reference convolves two radii per
    complex kernel; for a real field that equals one real convolution per
    radius, which this restatement does.  The returned mask is the
    reference's last `mask1`: the eroded mask of the first radius of the
    last pair (the second smallest radius for an even count), or of the
    last radius for an odd count."""
    f = np.asarray(f)
    if f.dtype not in (np.float32, np.float64):
        f = f.astype(np.float64)
    T = f.dtype.type
    vsz = np.broadcast_to(np.asarray(vsz, dtype=np.float64), (3,))
    if rs is None:
        step = 2 * vsz.max()
        rs = np.arange(9, step - 1e-12, -step)  # 9:-2*max(vsz):2*max(vsz)
        rs = rs[rs >= step - 1e-12]
    rs = np.sort(np.asarray(rs, dtype=np.float64).reshape(-1))[::-1]
    mask = np.asarray(mask) != 0
    fl = np.zeros(f.shape, T, order="F")
    ix, iy, iz = crop_indices(mask)
    pad = [int(np.ceil(rs.max() / vsz.min()))] * 3
    fc = f[np.ix_(ix, iy, iz)] if f.ndim == 3 else f[np.ix_(ix, iy, iz, np.arange(f.shape[3]))]
    fp = padfastfft(fc, pad)
    m0 = padfastfft(mask[np.ix_(ix, iy, iz)], pad)
    sz = fp.shape[:3]
    echoes = _echoes(fp)
    F = [np.fft.fftn(fe).astype(_cplx(T)) for fe in echoes]
    M = np.fft.fftn(m0.astype(np.float32)).astype(np.complex64)
    acc = [np.zeros(sz, T, order="F") for _ in echoes]
    prev = np.zeros(sz, bool)
    ih = None
    mask1 = maskn = prev

    def radius(r):
        H = _smv_symbol(sz, r, vsz, T)
        mr = np.real(np.fft.ifftn(H.astype(np.float32) * M).astype(np.complex64)) > np.float32(0.999)
        return T(1) - H, mr

    def accumulate(h, mr, prev):
        for e, Fe in enumerate(F):
            c = np.real(np.fft.ifftn(h * Fe).astype(_cplx(T))).astype(T)
            acc[e] = acc[e] + (mr.astype(T) - prev.astype(T)) * c

    # vsharp.m:87-140: radii in pairs (one complex kernel each in the
    # reference), a last single radius alone
    for ii in range(0, len(rs), 2):
        if ii + 1 < len(rs):
            h1, mask1 = radius(rs[ii])
            h2, mask2 = radius(rs[ii + 1])
            if ii == 0:
                with np.errstate(divide="ignore"):
                    ih = T(1) / h1
                ih[np.abs(h1) < T(thr)] = 0
            accumulate(h1, mask1, prev)
            accumulate(h2, mask2, mask1)
            prev = maskn = mask2
        else:
            h, mask1 = radius(rs[ii])
            if ih is None:
                raise ValueError("vsharp needs at least two radii (the reference sets its "
                                 "inverse kernel from the first pair)")
            accumulate(h, mask1, prev)
            maskn = mask1
    mlast = maskn
    out = []
    for a in acc:
        out.append((mlast * np.real(np.fft.ifftn(ih * np.fft.fftn(a).astype(_cplx(T)))
                                    .astype(_cplx(T)))).astype(T))
    fo = out[0] if f.ndim == 3 else np.stack(out, axis=3)
    szc = (len(ix), len(iy), len(iz))
    res = unpadfastfft(fo, szc, pad)
    if f.ndim == 3:
        fl[np.ix_(ix, iy, iz)] = res
    else:
        fl[np.ix_(ix, iy, iz, np.arange(f.shape[3]))] = res
    mo = np.zeros(mask.shape, bool, order="F")
    mo[np.ix_(ix, iy, iz)] = unpadfastfft(mask1, szc, pad)  # vsharp.m:149-151
    return fl, mo
