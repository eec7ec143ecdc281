"""Background field removal on the GPU: SHARP (src/bgremove/sharp.m) and
V-SHARP (src/bgremove/vsharp.m).

Same arguments, defaults and outputs as the reference: [fl, mask] =
sharp(f, mask, vsz, r=9, thr=0.05) and vsharp(f, mask, vsz, rs, thr=0.05)
for a 3-D or 4-D field (echoes share the mask); each is one call of
qsm_sharp / qsm_vsharp (libqsm_hip.so), which crops to the mask
(cropIndices.m), runs the spherical-mean-value convolutions on the device
and uncrops.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dtype_code, fortran, h3, lib, mask_u8, ptr

_sz = C.c_size_t
_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_lib.register("qsm_sharp_work_bytes", _sz, [C.c_int, C.POINTER(_sz), _dp, C.c_double])
_lib.register("qsm_vsharp_work_bytes", _sz, [C.c_int, C.POINTER(_sz), _dp, C.c_double, _sz])
_lib.register("qsm_sharp", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _sz, _dp, C.c_double,
                                     C.c_double, _p, _p])
_lib.register("qsm_vsharp", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _sz, _dp, _dp, _sz,
                                      C.c_double, _p, _p])
_lib.register("qsm_sharp_dev", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _sz, _dp, C.c_double,
                                         C.c_double, _p, _p, _p, _p])
_lib.register("qsm_vsharp_dev", C.c_int, [C.c_int, _p, _p, C.POINTER(_sz), _sz, _dp, _dp, _sz,
                                          C.c_double, _p, _p, _p, _p])


def _prep(f, mask):
    f = np.asarray(f)
    if f.ndim not in (3, 4):
        raise ValueError("f must be 3-D or 4-D")
    if f.dtype not in (np.float32, np.float64):
        f = f.astype(np.float64)
    m = np.asarray(mask)
    if m.shape != f.shape[:3]:
        raise ValueError("mask must match the first three dimensions of f")
    ne = 1 if f.ndim == 3 else f.shape[3]
    return fortran(f, f.dtype), mask_u8(m), ne


def default_radii(vsz):
    """vsharp.m:49: 9:-2*max(vsz):2*max(vsz) (MATLAB's colon: 9 - k*step)."""
    step = 2.0 * float(np.max(np.asarray(vsz, dtype=np.float64)))
    n = int(np.floor((9.0 - step) / step + 1e-10)) + 1
    return np.array([9.0 - k * step for k in range(max(n, 0))])


def sharp(f, mask, vsz, r=9, thr=0.05):
    """[fl, mask] = sharp(f, mask, vsz, [r], [thr]) (sharp.m:1-84)."""
    f, m, ne = _prep(f, mask)
    r = 9 if r is None or np.size(r) == 0 else float(r)
    fl = np.zeros(f.shape, dtype=f.dtype, order="F")
    mo = np.zeros(m.shape, dtype=np.uint8, order="F")
    check(lib().qsm_sharp(dtype_code(f.dtype), ptr(f), ptr(m), (_sz * 3)(*f.shape[:3]), ne,
                          h3(vsz), float(r), float(thr), ptr(fl), ptr(mo)))
    return fl, mo.view(bool)  # 0/1 bytes


def vsharp(f, mask, vsz, rs=None, thr=0.05):
    """[fl, mask] = vsharp(f, mask, vsz, [rs], [thr]) (vsharp.m:1-153)."""
    f, m, ne = _prep(f, mask)
    if rs is None or np.size(rs) == 0:
        rs = default_radii(vsz)
    rs = np.ascontiguousarray(np.asarray(rs, dtype=np.float64).reshape(-1))
    fl = np.zeros(f.shape, dtype=f.dtype, order="F")
    mo = np.zeros(m.shape, dtype=np.uint8, order="F")
    check(lib().qsm_vsharp(dtype_code(f.dtype), ptr(f), ptr(m), (_sz * 3)(*f.shape[:3]), ne,
                           h3(vsz), rs.ctypes.data_as(_dp), rs.size, float(thr), ptr(fl),
                           ptr(mo)))
    return fl, mo.view(bool)  # 0/1 bytes


def _code(dtype):
    return _lib.F32 if str(dtype).endswith("float32") else _lib.F64


def _checked(nb, what, rmax, vsz):
    if nb == 0:
        raise ValueError(f"{what}: radius {rmax!r} and voxel size {tuple(vsz)!r} must be "
                         f"finite and > 0, the crop non-empty")
    return nb


def work_bytes(dtype, szc, vsz, rmax) -> int:
    """Work space of sharp_dev (qsm_sharp_work_bytes)."""
    return _checked(int(lib().qsm_sharp_work_bytes(_code(dtype), (_sz * 3)(*szc), h3(vsz),
                                                   float(rmax))), "sharp", rmax, vsz)


def vsharp_work_bytes(dtype, szc, vsz, rmax, nechoes=1) -> int:
    """Work space of vsharp_dev (qsm_vsharp_work_bytes)."""
    return _checked(int(lib().qsm_vsharp_work_bytes(_code(dtype), (_sz * 3)(*szc), h3(vsz),
                                                    float(rmax), int(nechoes))), "vsharp",
                    rmax, vsz)


def _dev_args(f, mask, work, rmax, vsz, stream, v=False):
    import torch

    shape = tuple(int(s) for s in f.shape[:3])
    ne = 1 if f.dim() == 3 else int(f.shape[3])
    code = _lib.F64 if f.dtype == torch.float64 else _lib.F32
    nb = (vsharp_work_bytes(f.dtype, shape, vsz, rmax, ne) if v
          else work_bytes(f.dtype, shape, vsz, rmax))
    if work is None:
        work = torch.empty(nb, dtype=torch.uint8, device=f.device)
    elif work.numel() * work.element_size() < nb:
        raise ValueError(f"work space too small: {nb} bytes needed")
    sp = stream.cuda_stream if stream is not None else None
    return shape, ne, code, work, sp


def sharp_dev(fc, mc, vsz, r, thr, flc, mask_out, work=None, stream=None):
    """Device form on an already cropped volume (torch tensors, Fortran-order
    views; echoes along a fourth axis): qsm_sharp_dev, asynchronous."""
    shape, ne, code, work, sp = _dev_args(fc, mc, work, r, vsz, stream)
    check(lib().qsm_sharp_dev(code, fc.data_ptr(), mc.data_ptr(), (_sz * 3)(*shape), ne, h3(vsz),
                              float(r), float(thr), flc.data_ptr(), mask_out.data_ptr(),
                              work.data_ptr(), sp))
    return work


def vsharp_dev(fc, mc, vsz, rs, thr, flc, mask_out, work=None, stream=None):
    """Device form of vsharp on an already cropped volume: qsm_vsharp_dev."""
    rs = np.ascontiguousarray(np.asarray(rs, dtype=np.float64).reshape(-1))
    if rs.size == 0 or not np.all(np.isfinite(rs)):
        raise ValueError("vsharp: radii must be finite")
    shape, ne, code, work, sp = _dev_args(fc, mc, work, float(rs.max()), vsz, stream, v=True)
    check(lib().qsm_vsharp_dev(code, fc.data_ptr(), mc.data_ptr(), (_sz * 3)(*shape), ne,
                               h3(vsz), rs.ctypes.data_as(_dp), rs.size, float(thr),
                               flc.data_ptr(), mask_out.data_ptr(), work.data_ptr(), sp))
    return work
