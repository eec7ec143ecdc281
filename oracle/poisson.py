"""CPU restatement of the masked Poisson problem of
src/utils/poisson_solver (mgpcg.m, mex/lapmg_mex.c, mgpcg_mex.c) and of
src/unwrap/unwrapLaplacian.m, numpy/scipy.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and by bench.py's CPU
baseline leg), never by the product package.  mgpcg solves -Lap_h v = f on
the mask's voxels with v = 0 off the mask and on a one-voxel border around
the volume, to ||r|| <= tol ||f||; its result is the discrete solution to
that tolerance, whatever the preconditioner.  The restatement assembles the
same system (lapmg_mex.c's stencil with mx_pad_boundary's border) and
solves it exactly (scipy's sparse direct solver); the tests compare the GPU
solution with it to a tolerance derived from the stopping rule.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from . import oracle as orc


def neg_laplacian(mask, vsz):
    """-Lap_h restricted to the mask's voxels (lapmg_mex.c:62-89 with the
    zero border of mgpcg_mex.c), as a sparse matrix on the mask's voxels in
    column-major order, and their linear indices."""
    m = np.asarray(mask) != 0
    hx, hy, hz = (1.0 / (float(h) * float(h)) for h in vsz)
    hh = 2.0 * (hx + hy + hz)
    idx = np.flatnonzero(m.reshape(-1, order="F"))
    pos = -np.ones(m.size, dtype=np.int64)
    pos[idx] = np.arange(idx.size)
    nx, ny, nz = m.shape
    i, j, k = np.unravel_index(idx, m.shape, order="F")
    rows, cols, vals = [np.arange(idx.size)], [np.arange(idx.size)], [np.full(idx.size, hh)]
    for (di, dj, dk), w in (((1, 0, 0), hx), ((-1, 0, 0), hx), ((0, 1, 0), hy), ((0, -1, 0), hy),
                            ((0, 0, 1), hz), ((0, 0, -1), hz)):
        a, b, c = i + di, j + dj, k + dk
        ok = (a >= 0) & (a < nx) & (b >= 0) & (b < ny) & (c >= 0) & (c < nz)
        nb = np.full(idx.size, -1, dtype=np.int64)
        nb[ok] = pos[a[ok] + nx * (b[ok] + ny * c[ok])]
        sel = nb >= 0
        rows.append(np.arange(idx.size)[sel])
        cols.append(nb[sel])
        vals.append(np.full(sel.sum(), -w))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(idx.size, idx.size))
    return A, idx


def solve(f, mask, vsz):
    """The exact discrete solution of mgpcg's system (float64)."""
    A, idx = neg_laplacian(mask, vsz)
    v = np.zeros(np.asarray(f).size)
    if idx.size:
        v[idx] = spl.spsolve(A.tocsc(), np.asarray(f, dtype=np.float64).reshape(-1, order="F")[idx])
    return v.reshape(np.shape(f), order="F")


def unwrap_laplacian(phas, mask, vsz):
    """unwrapLaplacian.m:1-98 with the exact solve: crop to the mask's box
    grown by one voxel, per echo d2u = lapw(u, vsz) (the oracle's C
    restatement of lapw_mex.c) and v solving -Lap v = -mask .* d2u, uncrop.
    Returns (uphas, d2phas) in float64."""
    p = np.asarray(phas)
    m = np.asarray(mask) != 0
    up = np.zeros(p.shape)
    d2 = np.zeros(p.shape)
    nz = np.argwhere(m)
    lo = np.maximum(nz.min(axis=0) - 1, 0)
    hi = np.minimum(nz.max(axis=0) + 1, np.array(m.shape) - 1)
    sl = tuple(slice(a, b + 1) for a, b in zip(lo, hi))
    mc = m[sl]
    for t in range(1 if p.ndim == 3 else p.shape[3]):
        u = np.asfortranarray(p[sl] if p.ndim == 3 else p[sl + (t,)])
        d = orc.lapw(u, vsz)
        v = solve(-(mc * d), mc, vsz)
        if p.ndim == 3:
            up[sl], d2[sl] = v, d
        else:
            up[sl + (t,)], d2[sl + (t,)] = v, d
    return up, d2
