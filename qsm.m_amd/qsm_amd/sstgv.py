"""sstgv -- single-step TGV/TV QSM (src/inversion/sstgv.m:1-136) on the GPU.

Same arguments, defaults and flow as the reference: validate, crop phase
and mask to the mask's bounding box +/-2 voxels (sstgv.m:57-66), and per
echo compute d2u = lapw(u, vsz) and call pd_tgv_mex(-d2u, ...) when
alpha > eps, pd_tv_mex(-d2u, ...) otherwise (sstgv.m:76-87), then uncrop
into zeros.  The validation runs here; everything after it is one library
call (qsm_sstgv), which crops, runs lapw and the inversion and uncrops on
the GPU.
"""
from __future__ import annotations

import numpy as np

from ._lib import check, dtype_code, fortran, h3, lib, mask_u8, ptr, sz3
from .pd import _printer

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
    phas = fortran(np.asarray(phas))
    m = mask_u8(mask)
    vsz = np.asarray(vsz, dtype=np.float64).reshape(3)
    x = np.empty(phas.shape, dtype=phas.dtype, order="F")  # written whole by the call
    echoes = 1 if phas.ndim == 3 else phas.shape[3]
    cb, keep = _printer(int(verbose))
    check(lib().qsm_sstgv(dtype_code(phas.dtype), ptr(phas), ptr(m), sz3(phas.shape[:3]), echoes,
                          h3(vsz), float(lam), float(alpha), float(tol), int(maxit),
                          int(verbose), ptr(x), cb, None))
    del keep
    return x
