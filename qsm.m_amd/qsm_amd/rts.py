"""rts -- rapid two-step dipole inversion (src/inversion/rts.m) on the GPU.

Same arguments, opts fields, defaults and flow as the reference
(rts.m:1-118): x = rts(f, mask, vsz, bdir, opts) for a 3-D or 4-D local
field f, per echo step 1 (tolls iterations of LSMR with the k-space dipole
kernel, x1 = mask .* real(ifft3(.))) and step 2 (the ADMM of tv.m on the
ill-conditioned k-space region, M = mu (|D| > delta), starting from x1),
x = mask .* x.  Each echo is one call of qsm_rts (libqsm_hip.so).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dtype_code, fortran, h3, lib, mask_u8, ptr, sz3
from .tv import _d3

_sz = C.c_size_t
_p = C.c_void_p
_dp = C.POINTER(C.c_double)


class RtsOpts(C.Structure):
    """qsm_rts_opts (include/qsm_hip.h): the opts fields of rts.m:50-102."""
    _fields_ = [("tolls", C.c_int), ("mu", C.c_double), ("delta", C.c_double),
                ("tv", C.c_int), ("rtol", C.c_double), ("atol", C.c_double),
                ("maxit", C.c_int), ("verbose", C.c_int), ("rho", C.c_double),
                ("alpha", C.c_double), ("rtau", C.c_double), ("stau", C.c_double),
                ("gamma", C.c_double)]


_lib.register("qsm_rts_opts_default", None, [C.POINTER(RtsOpts)])
_lib.register("qsm_rts_work_bytes", _sz, [C.c_int, C.POINTER(_sz)])
_lib.register("qsm_rts", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _dp, _dp,
                                   C.POINTER(RtsOpts), _p, _dp, C.POINTER(C.c_int)])
_lib.register("qsm_rts_dev", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _dp, _dp,
                                       C.POINTER(RtsOpts), _p, _p, _dp, C.POINTER(C.c_int), _p])

INT_FIELDS = ("tolls", "tv", "maxit", "verbose")
FIELDS = ("tolls", "mu", "delta", "tv", "rtol", "atol", "maxit", "verbose", "rho", "alpha",
          "rtau", "stau", "gamma")


def make_opts(opts=None) -> RtsOpts:
    """rts.m:50-102: fields missing (or None/empty) take the defaults."""
    o = RtsOpts()
    lib().qsm_rts_opts_default(C.byref(o))
    if opts is None:
        return o
    get = opts.get if isinstance(opts, dict) else (lambda k, d=None: getattr(opts, k, d))
    for k in FIELDS:
        val = get(k, None)
        if val is None or (np.size(val) == 0):
            continue
        setattr(o, k, int(val) if k in INT_FIELDS else float(val))
    return o


def rts(f, mask, vsz, bdir=None, opts=None, return_trace=False):
    """x = rts(f, mask, vsz, [bdir], [opts]) (rts.m:1-118).

    return_trace: also return, per echo, the step-2 rows (||r||, eps_pri,
    ||s||, eps_dual, rho after the update)."""
    f = np.asarray(f)
    if f.ndim not in (3, 4):
        raise ValueError("f must be 3-D or 4-D")
    if f.dtype not in (np.float32, np.float64):
        f = f.astype(np.float64)
    m = np.asarray(mask)
    if m.shape != f.shape[:3]:
        raise ValueError("mask must match the first three dimensions of f")
    if bdir is None or np.size(bdir) == 0:
        bdir = (0.0, 0.0, 1.0)
    o = make_opts(opts)
    m8 = mask_u8(m)
    code = dtype_code(f.dtype)
    echoes = 1 if f.ndim == 3 else f.shape[3]
    x = np.zeros(f.shape, dtype=f.dtype, order="F")
    traces = []
    for t in range(echoes):
        if o.verbose and echoes > 1:
            print(f"Echo: {t + 1}/{echoes}")
        ft = fortran(f if f.ndim == 3 else f[..., t], f.dtype)
        xt = np.zeros(ft.shape, dtype=f.dtype, order="F")
        tr = np.zeros(5 * o.maxit, dtype=np.float64)
        it = C.c_int(0)
        check(lib().qsm_rts(code, ptr(ft), ptr(m8), sz3(ft.shape), h3(vsz), _d3(bdir),
                            C.byref(o), ptr(xt), tr.ctypes.data_as(_dp), C.byref(it)))
        if f.ndim == 3:
            x = xt
        else:
            x[..., t] = xt
        traces.append(tr[: 5 * it.value].reshape(it.value, 5))
    if return_trace:
        return x, (traces[0] if f.ndim == 3 else traces)
    return x


def work_bytes(dtype, shape) -> int:
    code = _lib.F32 if str(dtype).endswith("float32") else _lib.F64
    return int(lib().qsm_rts_work_bytes(code, sz3(shape)))


def rts_dev(f, mask, vsz, bdir, opts, out, work=None, stream=None, trace=None):
    """Device form on torch tensors (3-D Fortran-order views; mask uint8):
    one echo of rts.m on the given stream.  Returns the step-2 iteration
    count."""
    import torch

    shape = tuple(int(s) for s in f.shape)
    code = _lib.F64 if f.dtype == torch.float64 else _lib.F32
    o = make_opts(opts)
    nb = work_bytes(f.dtype, shape)
    if work is None:
        work = torch.empty(nb, dtype=torch.uint8, device=f.device)
    elif work.numel() * work.element_size() < nb:
        raise ValueError(f"work space too small: {nb} bytes needed")
    tr = trace if trace is not None else np.zeros(5 * o.maxit, dtype=np.float64)
    it = C.c_int(0)
    sp = stream.cuda_stream if stream is not None else None
    check(lib().qsm_rts_dev(code, f.data_ptr(), mask.data_ptr(), sz3(shape), h3(vsz), _d3(bdir),
                            C.byref(o), out.data_ptr(), work.data_ptr(),
                            tr.ctypes.data_as(_dp), C.byref(it), sp))
    return it.value
