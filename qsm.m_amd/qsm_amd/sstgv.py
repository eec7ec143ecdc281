"""sstgv -- single-step TGV/TV QSM (src/inversion/sstgv.m:1-136) on the GPU.

Same arguments, defaults and flow as the reference: validate, crop phase
and mask to the mask's bounding box +/-2 voxels (sstgv.m:57-66), and per
echo compute d2u = lapw(u, vsz) and call pd_tgv_mex(-d2u, ...) when
alpha > eps, pd_tv_mex(-d2u, ...) otherwise (sstgv.m:76-87), then uncrop
into zeros.  lapw and the inversion run on the GPU.
"""
from __future__ import annotations

import numpy as np

from . import fd, pd
from .fastfft import crop_indices

EPS = np.finfo(np.float64).eps


def _validate(phas, mask, vsz, lam, alpha, tol, maxit, verbose):
    phas = np.asarray(phas)
    if phas.dtype not in (np.float32, np.float64):
        raise TypeError("phas must be single or double")
    if phas.ndim not in (3, 4):
        raise ValueError("phas must be 3-D or 4-D")
    if not np.all(np.isfinite(phas)):
        raise ValueError("phas must be finite")
    m = np.asarray(mask)
    if m.shape != phas.shape[:3]:
        raise ValueError("mask must match the first three dimensions of phas")
    if not np.all((m == 0) | (m == 1)):
        raise ValueError("mask must be binary")
    v = np.asarray(vsz, dtype=np.float64).reshape(-1)
    if v.size != 3 or not np.all(np.isfinite(v)) or np.any(v <= 0):
        raise ValueError("vsz must be 3 finite values > 0")
    for name, x in (("lambda", lam),):
        if not (np.isfinite(x) and x > 0):
            raise ValueError(f"{name} must be finite and > 0")
    if not (np.isfinite(alpha) and np.isfinite(tol)):
        raise ValueError("alpha and tol must be finite")
    if int(maxit) != maxit or maxit <= 0:
        raise ValueError("maxit must be a positive integer")
    if not isinstance(verbose, bool) and not (0 <= int(verbose) <= maxit):
        raise ValueError("verbose must be logical or an integer in [0, maxit]")


def sstgv(phas, mask, vsz, lam=1e-4, alpha=3.0, tol=3e-4, maxit=5000, verbose=True):
    """x = sstgv(phas, mask, vsz, [lambda], [alpha], [tol], [maxit], [verbose])

    alpha1 = lambda, alpha0 = alpha*lambda; alpha < eps selects TV.
    """
    _validate(phas, mask, vsz, lam, alpha, tol, maxit, verbose)
    phas = np.asarray(phas)
    mask = np.asarray(mask) != 0
    vsz = np.asarray(vsz, dtype=np.float64).reshape(3)
    x = np.zeros(phas.shape, dtype=phas.dtype, order="F")
    ix, iy, iz = crop_indices(mask)
    if ix.size == 0:
        return x
    sl = tuple(slice(max(0, i[0] - 2), min(n, i[-1] + 3)) for i, n in zip((ix, iy, iz), mask.shape))
    m = np.asfortranarray(mask[sl])
    echoes = 1 if phas.ndim == 3 else phas.shape[3]
    for t in range(echoes):
        u = phas[sl] if phas.ndim == 3 else phas[sl + (t,)]
        d2u = fd.lapw(np.asfortranarray(u), vsz)
        vb = int(verbose)
        if alpha > EPS:
            chi = pd.pd_tgv_mex(-d2u, m, vsz, lam, alpha, tol, maxit, vb)
        else:
            chi = pd.pd_tv_mex(-d2u, m, vsz, lam, tol, maxit, vb)
        if phas.ndim == 3:
            x[sl] = chi
        else:
            x[sl + (t,)] = chi
    return x
