"""Laplacian phase unwrapping (src/unwrap/unwrapLaplacian.m) and the masked
Poisson solver it uses (src/utils/poisson_solver/mgpcg.m) on the GPU.

Same arguments, defaults and outputs as the reference:
[uphas, d2phas] = unwrap_laplacian(phas, mask, vsz) for a 3-D or 4-D phase
(one call of qsm_unwrap_laplacian), and v = mgpcg(f, mask, vsz, tol, maxit,
mgopts) (qsm_mgpcg).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dtype_code, fortran, h3, lib, mask_u8, ptr, sz3

_sz = C.c_size_t
_p = C.c_void_p
_dp = C.POINTER(C.c_double)


class MgOpts(C.Structure):
    """qsm_mg_opts (include/qsm_hip.h): mgopts of mgpcg.m:11-52."""
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int), ("mu", C.c_int), ("npre", C.c_int),
                ("npost", C.c_int), ("nboundary", C.c_int), ("nlevels", C.c_int)]


_lib.register("qsm_mg_opts_default", None, [C.POINTER(MgOpts)])
_lib.register("qsm_mgpcg", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _dp, C.c_double, C.c_int,
                                     C.POINTER(MgOpts), _p, C.POINTER(C.c_int), _dp])
_lib.register("qsm_unwrap_laplacian", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _sz, _dp, _p, _p])


def make_mgopts(mgopts=None) -> MgOpts:
    o = MgOpts()
    lib().qsm_mg_opts_default(C.byref(o))
    if mgopts:
        get = mgopts.get if isinstance(mgopts, dict) else (lambda k, d=None: getattr(mgopts, k, d))
        for k, _ in MgOpts._fields_:
            val = get(k, None)
            if val is not None and np.size(val) > 0:
                setattr(o, k, float(val) if k == "tol" else int(val))
    return o


def mgpcg(f, mask, vsz, tol=None, maxit=None, mgopts=None, return_info=False):
    """v = mgpcg(f, mask, vsz, [tol], [maxit], [mgopts]) (mgpcg.m:1-58):
    -Lap v = f on the mask, v = 0 off it.  return_info: also (iterations,
    residual norms)."""
    f = np.asarray(f)
    if f.dtype not in (np.float32, np.float64):
        raise TypeError("f must be single or double")
    f = fortran(f, f.dtype)
    m = mask_u8(mask)
    if m.shape != f.shape or f.ndim != 3:
        raise ValueError("f and mask must be 3-D of the same size")
    if tol is None or np.size(tol) == 0:
        tol = float(np.sqrt(np.finfo(f.dtype).eps))
    if maxit is None or np.size(maxit) == 0:
        maxit = f.size
    o = make_mgopts(mgopts)
    v = np.zeros(f.shape, dtype=f.dtype, order="F")
    it = C.c_int(0)
    # the history has maxit entries (numel(f) by default): only when asked
    res = np.zeros(int(maxit), dtype=np.float64) if return_info else None
    check(lib().qsm_mgpcg(dtype_code(f.dtype), ptr(f), ptr(m), (_sz * 3)(*f.shape), h3(vsz),
                          float(tol), int(maxit), C.byref(o), ptr(v), C.byref(it),
                          res.ctypes.data_as(_dp) if res is not None else None))
    if return_info:
        return v, (it.value, res[: it.value])
    return v


def unwrap_laplacian(phas, mask, vsz):
    """[uphas, d2phas] = unwrapLaplacian(phas, mask, vsz) (unwrapLaplacian.m:
    1-76)."""
    p = np.asarray(phas)
    if p.dtype not in (np.float32, np.float64):
        raise TypeError("phas must be single or double")
    if p.ndim not in (3, 4) or not np.all(np.isfinite(p)):
        raise ValueError("phas must be a finite 3-D or 4-D array")
    m = np.asarray(mask)
    if m.shape != p.shape[:3] or not np.all((m == 0) | (m == 1)):
        raise ValueError("mask must be binary and match phas")
    v = np.asarray(vsz, dtype=np.float64).reshape(-1)
    if v.size != 3 or not np.all(np.isfinite(v)) or np.any(v <= 0):
        raise ValueError("vsz must be 3 finite values > 0")
    p = fortran(p, p.dtype)
    m8 = mask_u8(m)
    ne = 1 if p.ndim == 3 else p.shape[3]
    up = np.zeros(p.shape, dtype=p.dtype, order="F")
    d2 = np.zeros(p.shape, dtype=p.dtype, order="F")
    check(lib().qsm_unwrap_laplacian(dtype_code(p.dtype), ptr(p), ptr(m8), (_sz * 3)(*p.shape[:3]),
                                     ne, h3(v), ptr(up), ptr(d2)))
    return up, d2


_lib.register("qsm_mgpcg_work_bytes", _sz, [C.c_int, C.POINTER(_sz), C.c_int])
_lib.register("qsm_mgpcg_dev", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _dp, C.c_double, C.c_int,
                                         C.POINTER(MgOpts), _p, _p, C.POINTER(C.c_int), _dp, _p])
_lib.register("qsm_unwrap_laplacian_dev", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _dp, _p, _p,
                                                    _p, C.POINTER(C.c_int), _p])


def work_bytes(dtype, shape, nlevels=0) -> int:
    code = _lib.F32 if str(dtype).endswith("float32") else _lib.F64
    return int(lib().qsm_mgpcg_work_bytes(code, sz3(shape), int(nlevels)))


def _torch_code(t):
    import torch

    return _lib.F64 if t.dtype == torch.float64 else _lib.F32


def _work(t, shape, nlevels, work):
    import torch

    nb = work_bytes(t.dtype, shape, nlevels)
    if work is None:
        return torch.empty(nb, dtype=torch.uint8, device=t.device)
    if work.numel() * work.element_size() < nb:
        raise ValueError(f"work space too small: {nb} bytes needed")
    return work


def mgpcg_dev(f, mask, vsz, tol, maxit, mgopts, v, work=None, stream=None):
    """Device form on torch tensors (3-D Fortran-order views; mask uint8):
    mgpcg on the given stream.  Returns (iterations, residual norms)."""
    shape = tuple(int(s) for s in f.shape)
    o = make_mgopts(mgopts)
    work = _work(f, shape, o.nlevels, work)
    it = C.c_int(0)
    res = np.zeros(min(int(maxit), 1 << 16), dtype=np.float64)
    sp = stream.cuda_stream if stream is not None else None
    if int(maxit) > res.size:
        res = None     # a long history is not kept
    check(lib().qsm_mgpcg_dev(_torch_code(f), f.data_ptr(), mask.data_ptr(), sz3(shape), h3(vsz),
                              float(tol), int(maxit), C.byref(o), v.data_ptr(), work.data_ptr(),
                              C.byref(it), res.ctypes.data_as(_dp) if res is not None else None,
                              sp))
    return it.value, (res[: it.value] if res is not None else None)


def unwrap_laplacian_dev(u, mask, vsz, uphas, d2u, work=None, stream=None):
    """Device form of one echo of unwrapLaplacian on an already cropped
    volume (torch tensors, 3-D Fortran-order views).  Returns the PCG
    iteration count."""
    shape = tuple(int(s) for s in u.shape)
    work = _work(u, shape, 0, work)
    it = C.c_int(0)
    sp = stream.cuda_stream if stream is not None else None
    check(lib().qsm_unwrap_laplacian_dev(_torch_code(u), u.data_ptr(), mask.data_ptr(), sz3(shape),
                                         h3(vsz), uphas.data_ptr(), d2u.data_ptr(), work.data_ptr(),
                                         C.byref(it), sp))
    return it.value
