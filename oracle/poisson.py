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


# ---- the reference's algorithm (mgpcg_mex.c), for timing on the CPU ------
#
# A numpy restatement of mgpcg_mex.c's multigrid-preconditioned CG in
# float64: the padded hierarchy (mx_init, the voxel size doubling per
# level), coarsen_grid_mex.c, boundary_mask_mex.c, the red-black
# gauss_seidel_mex.c with the exact voxel sets its loops visit, residual,
# restrict and correct, the coarsest level swept in rounds of 128 + 128 to
# a relative residual of 1e-8 (mx_cycle), mx_mg and mx_pcg.  Each colour of
# a sweep has no dependency inside it, so updating it as one vector
# operation visits the same values as the C loops.


class _Level:
    def __init__(self, shape, h, dtype=np.float64):
        self.shape = shape
        self.h = h
        n = int(np.prod(shape))
        self.v = np.zeros(n, dtype)
        self.f = np.zeros(n, dtype)
        self.r = np.zeros(n, dtype)
        self.G = np.zeros(shape, bool)
        self.B = np.zeros(shape, bool)

    def finish(self):
        nx, ny, nz = self.shape
        i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        inner = (i > 0) & (j > 0) & (k > 0) & (i < nx - 1) & (j < ny - 1) & (k < nz - 1)
        col = (i + j + k) & 1
        row_eq = (j & 1) == (k & 1)
        # gauss_seidel_mex.c:141-187: forward passes s = 0, 1; reverse passes
        # start at the far corner (first colour parity nx + 1); the second
        # pass skips i = 1 (forward) or i = nx - 2 (reverse) on rows j = k mod 2
        fwd = [inner & (col == 0), inner & (col == 1) & ~((i == 1) & row_eq)]
        c0 = (nx + 1) & 1
        rev = [inner & (col == c0), inner & (col == 1 - c0) & ~((i == nx - 2) & row_eq)]
        F = lambda a: a.reshape(-1, order="F")  # noqa: E731
        self.sets = {}
        for name, M in (("G", self.G), ("B", self.B)):
            self.sets[name] = ([np.flatnonzero(F(s & M)) for s in fwd],
                               [np.flatnonzero(F(s & M)) for s in rev])
        self.gidx = np.flatnonzero(F(self.G & inner))
        hx, hy, hz = (1.0 / (h * h) for h in self.h)
        T = self.v.dtype.type
        self.coef = tuple(T(c) for c in (hx, hy, hz, 1.0 / (2.0 * (hx + hy + hz)), -2.0 * (hx + hy + hz)))
        self.st = (1, nx, nx * ny)


def _gs(L, which, iters, reverse):
    hx, hy, hz, hh, _ = L.coef
    sx, sy, sz = L.st
    v, f = L.v, L.f
    sets = L.sets[which][1 if reverse else 0]
    for _ in range(iters):
        for idx in sets:
            v[idx] = hh * (f[idx] + ((hx * (v[idx - sx] + v[idx + sx]) + hy * (v[idx - sy] + v[idx + sy]))
                                     + hz * (v[idx - sz] + v[idx + sz])))


def _residual(L, out):
    hx, hy, hz, _, hh = L.coef
    sx, sy, sz = L.st
    v, idx = L.v, L.gidx
    out[:] = 0.0
    out[idx] = L.f[idx] + (((hh * v[idx] + hx * (v[idx - sx] + v[idx + sx]))
                            + hy * (v[idx - sy] + v[idx + sy])) + hz * (v[idx - sz] + v[idx + sz]))


def _neg_lap(L, p, out):
    hx, hy, hz, _, hh = L.coef
    sx, sy, sz = L.st
    idx = L.gidx
    out[:] = 0.0
    out[idx] = -((((hh * p[idx] + hx * (p[idx - sx] + p[idx + sx]))
                   + hy * (p[idx - sy] + p[idx + sy])) + hz * (p[idx - sz] + p[idx + sz])))


def _restrict(C, L):
    """restrict_mex.c: full weighting onto C.G (0 elsewhere)."""
    nx, ny, nz = L.shape
    r = L.r.reshape(L.shape, order="F")
    n2 = C.shape
    sl = lambda o, n: slice(o, o + 2 * (n - 2), 2)  # noqa: E731  (fine 2 i2 - 2 + o)
    blk = {}
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                blk[a, b, c] = r[sl(1 + a, n2[0]), sl(1 + b, n2[1]), sl(1 + c, n2[2])]
    corners = sum(blk[a, b, c] for a in (-1, 1) for b in (-1, 1) for c in (-1, 1))
    edges = sum(blk[t] for t in blk if sum(x == 0 for x in t) == 1)
    faces = sum(blk[t] for t in blk if sum(x == 0 for x in t) == 2)
    val = 0.015625 * corners + 0.03125 * edges + 0.0625 * faces + 0.125 * blk[0, 0, 0]
    f2 = np.zeros(n2, L.r.dtype)
    f2[1:-1, 1:-1, 1:-1] = val
    f2[~C.G] = 0.0
    C.f[:] = f2.reshape(-1, order="F")


def _correct(L, C):
    """correct_mex.c: trilinear interpolation of C.v into L.v on L.G."""
    n2 = C.shape
    v2 = C.v.reshape(n2, order="F")
    v = L.v.reshape(L.shape, order="F")  # a view (Fortran reshape of a 1-D array)
    G = L.G
    m = [n - 2 for n in n2]
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                acc = 0.0
                for ei in range(di + 1):
                    for ej in range(dj + 1):
                        for ek in range(dk + 1):
                            acc = acc + v2[1 + ei:1 + ei + m[0], 1 + ej:1 + ej + m[1], 1 + ek:1 + ek + m[2]]
                w = v.dtype.type(1.0 / ((1 + di) * (1 + dj) * (1 + dk)))
                fi = slice(1 + di, 1 + di + 2 * m[0], 2)
                fj = slice(1 + dj, 1 + dj + 2 * m[1], 2)
                fk = slice(1 + dk, 1 + dk + 2 * m[2], 2)
                sub = v[fi, fj, fk]
                g = G[fi, fj, fk]
                add = (w * acc)[: sub.shape[0], : sub.shape[1], : sub.shape[2]]
                sub[g] += add[g]
    L.v[:] = v.reshape(-1, order="F")


def _cycle(lv, l, o):
    L = lv[l]
    if l < len(lv) - 1:
        bs = (1 << l) * o["nboundary"]
        _gs(L, "G", o["npre"], False)
        _gs(L, "B", bs, False)
        _residual(L, L.r)
        C = lv[l + 1]
        _restrict(C, L)
        C.v[:] = 0.0
        for _ in range(o["mu"]):
            _cycle(lv, l + 1, o)
            if l + 1 == len(lv) - 1:
                break
        _correct(L, C)
        _gs(L, "B", bs, True)
        _gs(L, "G", o["npost"], True)
    else:
        reltol = (1e-4 if L.f.dtype == np.float32 else 1e-8) * np.linalg.norm(L.f.astype(np.float64))
        for _ in range(1024):
            _gs(L, "G", 128, False)
            _gs(L, "G", 128, True)
            _residual(L, L.r)
            if np.linalg.norm(L.r.astype(np.float64)) < reltol:
                break


def _hierarchy(mask, vsz, nlevels, dtype=np.float64):
    m = np.pad(np.asarray(mask) != 0, 1)
    lv = [_Level(m.shape, np.asarray(vsz, float), dtype)]
    lv[0].G = m
    for l in range(1, nlevels):
        P = lv[-1]
        shape = tuple(((n - 2 + 1) >> 1) + 2 for n in P.shape)
        C = _Level(shape, P.h * 2.0, dtype)
        nx, ny, nz = P.shape
        G = P.G
        g = np.zeros(shape, bool)
        # coarsen_grid_mex.c: fine (2 i2 - 1, ...) and its +1 neighbours
        # inside the interior must all be active
        I = 2 * np.arange(1, shape[0] - 1) - 1
        J = 2 * np.arange(1, shape[1] - 1) - 1
        K = 2 * np.arange(1, shape[2] - 1) - 1
        acc = G[np.ix_(I, J, K)].copy()
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    if di == dj == dk == 0:
                        continue
                    Ii = np.where(I + di <= nx - 2, I + di, I)
                    Jj = np.where(J + dj <= ny - 2, J + dj, J)
                    Kk = np.where(K + dk <= nz - 2, K + dk, K)
                    acc &= G[np.ix_(Ii, Jj, Kk)]
        g[1:-1, 1:-1, 1:-1] = acc
        C.G = g
        lv.append(C)
    for L in lv:
        L.B = _band(L.G)
        L.finish()
    return lv


def _band(G):
    """boundary_mask_mex.c."""
    nx, ny, nz = G.shape
    b = np.zeros_like(G)
    inner = np.zeros_like(G)
    inner[1:-1, 1:-1, 1:-1] = True
    face = inner.copy()
    face[2:-2, 2:-2, 2:-2] = False
    b[face] = G[face]
    # 2x2x2 blocks from odd corners that are not fully active
    I, J, K = (np.arange(1, n - 1, 2) for n in G.shape)
    full = np.ones((I.size, J.size, K.size), bool)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                full &= G[np.ix_(I + di, J + dj, K + dk)]
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                sub = np.ix_(I + di, J + dj, K + dk)
                b[sub] = np.where(full, b[sub], G[sub])
    grown = b.copy()
    grown[1:-1, 1:-1, 1:-1] |= (b[:-2, 1:-1, 1:-1] | b[2:, 1:-1, 1:-1] | b[1:-1, :-2, 1:-1]
                                | b[1:-1, 2:, 1:-1] | b[1:-1, 1:-1, :-2] | b[1:-1, 1:-1, 2:])
    return G & inner & grown


def mgpcg_port(f, mask, vsz, tol=None, maxit=None, mgopts=None):
    """mgpcg.m + mgpcg_mex.c (mx_pcg, mx_mg, mx_cycle) in the class of f
    (float32 or float64; norms and dot products in float64).  Returns (v,
    iterations, residual norms)."""
    f = np.asarray(f)
    T = f.dtype.type if f.dtype in (np.float32, np.float64) else np.float64
    f = f.astype(T)
    o = {"tol": -1.0, "maxit": 1, "mu": 1, "npre": 1, "npost": 1, "nboundary": 2,
         "nlevels": max(1, int(np.floor(np.log2(min(f.shape)))) - 3)}
    o.update({k: v for k, v in (mgopts or {}).items() if v is not None})
    tol = np.sqrt(np.finfo(T).eps) if tol is None else tol
    maxit = f.size if maxit is None else maxit
    lv = _hierarchy(mask, vsz, int(o["nlevels"]), T)
    L0 = lv[0]
    L0.f[:] = np.pad(f, 1).reshape(-1, order="F")
    x = np.zeros_like(L0.f)
    p = np.zeros_like(L0.f)
    r, q = L0.f, L0.v
    nrm = lambda a: float(np.linalg.norm(a.astype(np.float64)))  # noqa: E731
    res = nrm(r)
    reltol = tol * res
    rho, hist = 1.0, []
    for _ in range(int(maxit)):
        if res <= reltol:
            break
        prevrho = rho
        q[:] = 0.0
        if o["tol"] > 0:
            mtol = o["tol"] * nrm(r)
            for _ in range(o["maxit"]):
                _residual(L0, L0.r)
                if nrm(L0.r) < mtol:
                    break
                _cycle(lv, 0, o)
        else:
            for _ in range(o["maxit"]):
                _cycle(lv, 0, o)
        rho = float(r.astype(np.float64) @ q.astype(np.float64))
        beta = rho / prevrho
        p *= T(beta)
        p += q
        _neg_lap(L0, p, q)
        a = rho / float(p.astype(np.float64) @ q.astype(np.float64))
        x += T(a) * p
        r -= T(a) * q
        res = nrm(r)
        hist.append(res)
    v = x.reshape(L0.shape, order="F")[1:-1, 1:-1, 1:-1]
    return np.asfortranarray(v), len(hist), np.array(hist)


def unwrap_laplacian_port(phas, mask, vsz):
    """unwrapLaplacian.m with mgpcg_port in place of the exact solve
    (tol 1e-8, maxit ceil(sqrt(numel)), mgopts maxit = npre = npost =
    nboundary = 2) in the class of phas.  3-D phase; returns (uphas,
    d2phas, iterations)."""
    p = np.asarray(phas)
    if p.dtype not in (np.float32, np.float64):
        p = p.astype(np.float64)
    m = np.asarray(mask) != 0
    up = np.zeros(p.shape, p.dtype)
    d2 = np.zeros(p.shape, p.dtype)
    nz = np.argwhere(m)
    lo = np.maximum(nz.min(axis=0) - 1, 0)
    hi = np.minimum(nz.max(axis=0) + 1, np.array(m.shape) - 1)
    sl = tuple(slice(a, b + 1) for a, b in zip(lo, hi))
    mc = m[sl]
    d = orc.lapw(np.asfortranarray(p[sl]), vsz)
    v, it, _ = mgpcg_port(-(mc * d).astype(p.dtype), mc, vsz, 1e-8, int(np.ceil(np.sqrt(mc.size))),
                          {"maxit": 2, "npre": 2, "npost": 2, "nboundary": 2})
    up[sl], d2[sl] = v, d
    return up, d2, it
