"""ctypes binding of libqsm_hip.so (the C ABI declared in include/qsm_hip.h).

The library is the only compute path: if it is missing or cannot be loaded
every entry point raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("QSM_HIP_LIB", PKG_DIR / "lib" / "libqsm_hip.so"))

QSM_OK, QSM_EARG, QSM_EHIP, QSM_ENOMEM, QSM_ENODEV, QSM_EFULL = range(6)
F64, F32 = 0, 1
TV, TGV = 0, 1

_sz_t = C.c_size_t
_p = C.c_void_p
_u8p = C.c_void_p
_dp = C.POINTER(C.c_double)
PRINT_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_void_p)


class QsmError(RuntimeError):
    """Raised when a libqsm_hip entry point returns a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"qsm error {code}: {msg}")
        self.code = code


class QsmAdmissionError(QsmError):
    """The serving runtime refused a request (HBM budget)."""


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "qsm_version": (C.c_int, []),
    "qsm_last_error": (C.c_char_p, []),
    "qsm_pd_tgv": (C.c_int, [C.c_int, _p, _u8p, C.POINTER(_sz_t), _dp, C.c_double,
                             C.c_double, C.c_double, C.c_uint32, C.c_uint32, _p,
                             PRINT_FN, _p]),
    "qsm_pd_tv": (C.c_int, [C.c_int, _p, _u8p, C.POINTER(_sz_t), _dp, C.c_double,
                            C.c_double, C.c_uint32, C.c_uint32, _p, PRINT_FN, _p]),
    "qsm_sstgv": (C.c_int, [C.c_int, _p, _u8p, C.POINTER(_sz_t), C.c_size_t, _dp,
                            C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_uint32, _p,
                            PRINT_FN, _p]),
    "qsm_pd_create": (C.c_int, [C.POINTER(_p), C.c_int, C.c_int, C.POINTER(_sz_t),
                                _dp, C.c_double, C.c_double, _p]),
    "qsm_pd_footprint": (_sz_t, [C.c_int, C.c_int, C.POINTER(_sz_t)]),
    "qsm_pd_load": (C.c_int, [_p, _p, _u8p, C.c_int]),
    "qsm_pd_iterate": (C.c_int, [_p, C.c_uint32, C.c_double]),
    "qsm_pd_status": (C.c_int, [_p, C.POINTER(C.c_uint32), C.POINTER(C.c_int), _dp]),
    "qsm_pd_norm_history": (C.c_int, [_p, _dp, C.c_int]),
    "qsm_pd_get_chi": (C.c_int, [_p, _p, C.c_int]),
    "qsm_pd_destroy": (None, [_p]),
    "qsm_pd_set_schedule": (C.c_int, [_p, C.c_int]),
    "qsm_pd_set_tiles": (C.c_int, [_p, C.c_int, C.c_int, C.c_int]),
    "qsm_pd_set_kz2": (C.c_int, [_p, C.c_int]),
    "qsm_pd_launch_counts": (C.c_int, [_p, C.POINTER(C.c_uint64)]),
    "qsm_pd_set_active_items": (C.c_int, [_p, C.c_int]),
    "qsm_pd_active_items": (C.c_int, [_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "qsm_pd_get_tiles": (C.c_int, [_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "qsm_tgv_dual_dev": (C.c_int, [C.c_int, _p, _p, _p, _p, _p, _p, _p, _u8p,
                                   C.POINTER(_sz_t), _dp, C.c_double, C.c_double,
                                   C.c_double, _p]),
    "qsm_tgv_primal_dev": (C.c_int, [C.c_int, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                     _u8p, C.POINTER(_sz_t), _dp, C.c_double, _p]),
    "qsm_tv_dual_dev": (C.c_int, [C.c_int, _p, _p, _p, _p, _p, _u8p, C.POINTER(_sz_t),
                                  _dp, C.c_double, C.c_double, _p]),
    "qsm_tv_primal_dev": (C.c_int, [C.c_int, _p, _p, _p, _p, _p, _p, _u8p,
                                    C.POINTER(_sz_t), _dp, C.c_double, _p]),
    "qsm_init_mask_dev": (C.c_int, [_u8p, _u8p, C.POINTER(_sz_t), _p]),
    "qsm_convergence_dev": (C.c_int, [C.c_int, _p, _p, _p, _p, _sz_t, C.c_double,
                                      _dp, C.POINTER(C.c_int), _p]),
    "qsm_step_sizes": (None, [_dp, _dp, _dp]),
    "qsm_pd_solve_static": (C.c_int, [C.c_int, C.c_int, _p, _u8p, C.POINTER(_sz_t), _dp,
                                      C.c_double, C.c_double, C.c_double, C.c_uint32, _p, _p]),
    "qsm_copy_dev": (C.c_int, [_p, _p, _sz_t, _p]),
}
for _n in ("gradfp", "gradf"):
    _SIGS[f"qsm_{_n}"] = (C.c_int, [C.c_int, _p, _p, _p, _p, C.POINTER(_sz_t), _dp])
    _SIGS[f"qsm_{_n}_dev"] = (C.c_int, [C.c_int, _p, _p, _p, _p, C.POINTER(_sz_t), _dp, _p])
for _n in ("gradfp_adj", "gradf_adj"):
    _SIGS[f"qsm_{_n}"] = (C.c_int, [C.c_int, _p, _p, _p, _p, C.POINTER(_sz_t), _dp])
    _SIGS[f"qsm_{_n}_dev"] = (C.c_int, [C.c_int, _p, _p, _p, _p, C.POINTER(_sz_t), _dp, _p])
_SIGS["qsm_gradfm"] = (C.c_int, [C.c_int, _p, _p, _p, _p, _u8p, C.POINTER(_sz_t), _dp])
_SIGS["qsm_gradfm_dev"] = (C.c_int, [C.c_int, _p, _p, _p, _p, _u8p, C.POINTER(_sz_t), _dp, _p])
_SIGS["qsm_gradfm_adj"] = (C.c_int, [C.c_int, _p, _p, _p, _p, _u8p, C.POINTER(_sz_t), _dp])
_SIGS["qsm_gradfm_adj_dev"] = (C.c_int, [C.c_int, _p, _p, _p, _p, _u8p, C.POINTER(_sz_t), _dp, _p])
_SIGS["qsm_lapw"] = (C.c_int, [C.c_int, _p, _p, C.POINTER(_sz_t), _dp])
_SIGS["qsm_lap1"] = (C.c_int, [C.c_int, _p, _p, _p, C.POINTER(_sz_t), _dp])
for _n in ("gradb", "gradc", "gradb_adj"):
    _SIGS[f"qsm_{_n}"] = (C.c_int, [C.c_int, _p, _p, _p, _p, C.POINTER(_sz_t), _dp])
    _SIGS[f"qsm_{_n}_dev"] = (C.c_int, [C.c_int, _p, _p, _p, _p, C.POINTER(_sz_t), _dp, _p])
for _n in ("gradbm", "gradcm", "gradbm_adj"):
    _SIGS[f"qsm_{_n}"] = (C.c_int, [C.c_int, _p, _p, _p, _p, _u8p, C.POINTER(_sz_t), _dp])
    _SIGS[f"qsm_{_n}_dev"] = (C.c_int, [C.c_int, _p, _p, _p, _p, _u8p, C.POINTER(_sz_t), _dp,
                                        _p])
_SIGS["qsm_lapw_dev"] = (C.c_int, [C.c_int, _p, _p, C.POINTER(_sz_t), _dp, _p])
_SIGS["qsm_lap1_dev"] = (C.c_int, [C.c_int, _p, _p, _p, C.POINTER(_sz_t), _dp, _p])


def register(name: str, restype, argtypes) -> None:
    """Adds a signature (used by modules binding further entry points)."""
    _SIGS[name] = (restype, argtypes)
    if _lib is not None:
        _bind(_lib, name)


def _bind(lib, name):
    fn = getattr(lib, name)
    fn.restype, fn.argtypes = _SIGS[name]


def lib():
    """Loads libqsm_hip.so once; raises if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise QsmError(QSM_ENODEV, f"{LIB_PATH} is not built; run qsm.m_amd/build.py")
        L = C.CDLL(str(LIB_PATH))
        for name in _SIGS:
            _bind(L, name)
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != QSM_OK:
        msg = lib().qsm_last_error().decode(errors="replace")
        if rc == QSM_EFULL:
            raise QsmAdmissionError(rc, msg)
        raise QsmError(rc, msg)


def sz3(shape) -> C.Array:
    if len(shape) != 3:
        raise ValueError(f"expected a 3-D volume, got shape {tuple(shape)}")
    return (_sz_t * 3)(*[int(s) for s in shape])


def h3(vsz) -> C.Array:
    v = np.asarray(vsz, dtype=np.float64).reshape(-1)
    if v.size == 1:
        v = np.repeat(v, 3)
    if v.size != 3:
        raise ValueError("voxel size needs 3 elements")
    return (C.c_double * 3)(*v.tolist())


def dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise TypeError(f"unsupported dtype {dt} (single or double)")


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def fortran(a, dtype=None) -> np.ndarray:
    """Column-major (MATLAB) contiguous copy-or-view of a."""
    a = np.asarray(a, dtype=dtype)
    return np.asfortranarray(a)


def mask_u8(mask) -> np.ndarray:
    """Column-major uint8 0/1 array of mask != 0 for the library's mask
    arguments: a view of a column-major bool mask, else one converted
    copy."""
    m = np.asarray(mask)
    if m.dtype != np.bool_:
        m = m != 0
    return np.asfortranarray(m).view(np.uint8)
