"""Finite-difference operators on the GPU, mirroring the reference MEX
(src/inversion/mex/gradfp*_mex.c, src/utils/fd/mex/grad{f,fm,b,bm}{,_adj}_mex.c,
gradc{,m}_mex.c, lap1_mex.c, src/unwrap/mex/lapw_mex.c): outputs are written IN PLACE into the arrays the
caller passes first, which must be column-major (Fortran-ordered) and of the
input's dtype, exactly as the MEX files overwrite their caller-allocated
prhs[0..2].  The wrappers below also offer allocating forms (gradfp(u, h)
etc.) like src/utils/fd/gradf.m.
"""
from __future__ import annotations

import numpy as np

from ._lib import check, dtype_code, fortran, h3, lib, mask_u8, ptr, sz3


def _out_ok(a, ref):
    if not (isinstance(a, np.ndarray) and a.flags.f_contiguous and a.dtype == ref.dtype
            and a.shape == ref.shape):
        raise ValueError("output arrays must be Fortran-contiguous, same shape/dtype as the input")


def _vec_op(name, outs, u, h, G=None):
    u = fortran(u)
    if u.dtype not in (np.float32, np.float64):
        raise TypeError("u must be single or double")
    for o in outs:
        _out_ok(o, u)
    args = [dtype_code(u.dtype), *(ptr(o) for o in outs), ptr(u)]
    if G is not None:
        g = mask_u8(G)
        args.append(ptr(g))
    check(getattr(lib(), name)(*args, sz3(u.shape), h3(h)))


def _adj_op(name, du, x, y, z, h, G=None):
    x, y, z = (fortran(a, du.dtype) for a in (x, y, z))
    _out_ok(du, x)
    args = [dtype_code(du.dtype), ptr(du), ptr(x), ptr(y), ptr(z)]
    if G is not None:
        g = mask_u8(G)
        args.append(ptr(g))
    check(getattr(lib(), name)(*args, sz3(du.shape), h3(h)))


# ---- MEX-shaped, in place ------------------------------------------------

def gradfp_mex(dx, dy, dz, u, h):
    """gradfp_mex(dx, dy, dz, u, h): periodic forward gradient, in place."""
    _vec_op("qsm_gradfp", (dx, dy, dz), u, h)
    return 1.0


def gradfp_adj_mex(du, x, y, z, h):
    """gradfp_adj_mex(du, x, y, z, h): -(periodic backward divergence)."""
    _adj_op("qsm_gradfp_adj", du, x, y, z, h)
    return 1.0


def gradf_mex(dx, dy, dz, u, h):
    """gradf_mex(dx, dy, dz, u, h): forward gradient, last face backward."""
    _vec_op("qsm_gradf", (dx, dy, dz), u, h)
    return 1.0


def gradfm_mex(dx, dy, dz, u, G, h):
    """gradfm_mex(dx, dy, dz, u, G, h): masked forward/backward gradient."""
    _vec_op("qsm_gradfm", (dx, dy, dz), u, h, G)
    return 1.0


def gradf_adj_mex(du, x, y, z, h):
    """gradf_adj_mex(du, x, y, z, h)."""
    _adj_op("qsm_gradf_adj", du, x, y, z, h)
    return 1.0


def gradfm_adj_mex(du, x, y, z, G, h):
    """gradfm_adj_mex(du, x, y, z, G, h)."""
    _adj_op("qsm_gradfm_adj", du, x, y, z, h, G)
    return 1.0


def gradb_mex(dx, dy, dz, u, h):
    """gradb_mex(dx, dy, dz, u, h): backward gradient, first face forward."""
    _vec_op("qsm_gradb", (dx, dy, dz), u, h)
    return 1.0


def gradbm_mex(dx, dy, dz, u, G, h):
    """gradbm_mex(dx, dy, dz, u, G, h): masked backward gradient."""
    _vec_op("qsm_gradbm", (dx, dy, dz), u, h, G)
    return 1.0


def gradb_adj_mex(du, x, y, z, h):
    """gradb_adj_mex(du, x, y, z, h): -1/h forward differences summed; the
    voxel (nx-1, ny-1, nz-1) keeps its value (the reference never writes
    it)."""
    _adj_op("qsm_gradb_adj", du, x, y, z, h)
    return 1.0


def gradbm_adj_mex(du, x, y, z, G, h):
    """gradbm_adj_mex(du, x, y, z, G, h)."""
    _adj_op("qsm_gradbm_adj", du, x, y, z, h, G)
    return 1.0


def gradc_mex(dx, dy, dz, u, h):
    """gradc_mex(dx, dy, dz, u, h): central gradient, one-sided faces, with
    the reference's plane-0 rewrite and its unwritten last corner."""
    _vec_op("qsm_gradc", (dx, dy, dz), u, h)
    return 1.0


def gradcm_mex(dx, dy, dz, u, G, h):
    """gradcm_mex(dx, dy, dz, u, G, h): masked central gradient."""
    _vec_op("qsm_gradcm", (dx, dy, dz), u, h, G)
    return 1.0


def lap1_mex(d2u, u, G, h):
    """lap1_mex(d2u, u, G, h): masked Laplacian, written in place on G."""
    u = fortran(u)
    if u.dtype not in (np.float32, np.float64):
        raise TypeError("u must be single or double")
    _out_ok(d2u, u)
    g = mask_u8(G)
    if g.shape != u.shape:
        raise ValueError("G must have the shape of u")
    check(lib().qsm_lap1(dtype_code(u.dtype), ptr(d2u), ptr(u), ptr(g), sz3(u.shape), h3(h)))
    return 1.0


def lapw_mex(d2u, u, h):
    """lapw_mex(d2u, u, h): Laplacian of wrapped phase, interior, in place."""
    u = fortran(u)
    _out_ok(d2u, u)
    check(lib().qsm_lapw(dtype_code(u.dtype), ptr(d2u), ptr(u), sz3(u.shape), h3(h)))
    return 1.0


# ---- allocating forms (src/utils/fd/gradf.m, gradfAdj.m, src/unwrap/lapw.m)

def _zeros_like(u):
    return np.zeros(u.shape, dtype=u.dtype, order="F")


def gradfp(u, h=(1.0, 1.0, 1.0)):
    u = fortran(u)
    d = [_zeros_like(u) for _ in range(3)]
    gradfp_mex(*d, u, h)
    return tuple(d)


def gradfp_adj(x, y, z, h=(1.0, 1.0, 1.0)):
    x = fortran(x)
    du = _zeros_like(x)
    gradfp_adj_mex(du, x, y, z, h)
    return du


def gradf(u, mask=None, h=(1.0, 1.0, 1.0)):
    """src/utils/fd/gradf.m:24-36: unmasked MEX when the mask is all true."""
    u = fortran(u)
    d = [_zeros_like(u) for _ in range(3)]
    if mask is None or np.all(np.asarray(mask) != 0):
        gradf_mex(*d, u, h)
    else:
        gradfm_mex(*d, u, mask, h)
    return tuple(d)


def gradf_adj(x, y, z, mask=None, h=(1.0, 1.0, 1.0)):
    """src/utils/fd/gradfAdj.m:29-39."""
    x = fortran(x)
    du = _zeros_like(x)
    if mask is None or np.all(np.asarray(mask) != 0):
        gradf_adj_mex(du, x, y, z, h)
    else:
        gradfm_adj_mex(du, x, y, z, mask, h)
    return du


def gradb(u, mask=None, h=(1.0, 1.0, 1.0)):
    """src/utils/fd/gradb.m:30-46: unmasked MEX when the mask is all true."""
    u = fortran(u)
    d = [_zeros_like(u) for _ in range(3)]
    if mask is None or np.all(np.asarray(mask) != 0):
        gradb_mex(*d, u, h)
    else:
        gradbm_mex(*d, u, mask, h)
    return tuple(d)


def gradb_adj(x, y, z, mask=None, h=(1.0, 1.0, 1.0)):
    """src/utils/fd/gradbAdj.m."""
    x = fortran(x)
    du = _zeros_like(x)
    if mask is None or np.all(np.asarray(mask) != 0):
        gradb_adj_mex(du, x, y, z, h)
    else:
        gradbm_adj_mex(du, x, y, z, mask, h)
    return du


def gradc(u, mask=None, h=(1.0, 1.0, 1.0)):
    """src/utils/fd/gradc.m:30-46."""
    u = fortran(u)
    d = [_zeros_like(u) for _ in range(3)]
    if mask is None or np.all(np.asarray(mask) != 0):
        gradc_mex(*d, u, h)
    else:
        gradcm_mex(*d, u, mask, h)
    return tuple(d)


def lap(u, mask=None, vsz=(1.0, 1.0, 1.0)):
    """src/utils/fd/lap.m:1-26: zeros like u, then lap1_mex on the mask (all
    true by default) -- the MEX runs even for a full mask."""
    u = fortran(u)
    d = _zeros_like(u)
    lap1_mex(d, u, np.ones(u.shape, np.uint8) if mask is None else mask, vsz)
    return d


def lapw(u, h=(1.0, 1.0, 1.0)):
    """src/unwrap/lapw.m:32-37: zeros, then the MEX fills the interior."""
    u = fortran(u)
    d = _zeros_like(u)
    lapw_mex(d, u, h)
    return d
