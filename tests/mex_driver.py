"""Drives the MEX drop-in wrappers (qsm.m_amd/mex/*_mex.c) the way MATLAB
does, through the CI build libqsm_mex_ci.so (wrappers + the mex.h shim):
prhs are mxArrays, plhs come back as mxArrays, mexErrMsgTxt raises MexError
and mexPrintf output is collected."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "qsm.m_amd" / "lib" / "libqsm_mex_ci.so"

CLASS = {np.dtype(np.float64): 6, np.dtype(np.float32): 7, np.dtype(np.uint8): 9,
         np.dtype(np.bool_): 3, np.dtype(np.uint32): 13, np.dtype(np.int32): 12}
DTYPE = {6: np.float64, 7: np.float32, 9: np.uint8, 3: np.uint8, 13: np.uint32, 12: np.int32}

MEX_FN = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_void_p), C.c_int, C.POINTER(C.c_void_p))

_lib = None


class MexError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        # the wrappers resolve libqsm_hip.so through their rpath ($ORIGIN)
        L = C.CDLL(str(LIB))
        L.mxCreateNumericArray.restype = C.c_void_p
        L.mxCreateNumericArray.argtypes = [C.c_size_t, C.POINTER(C.c_size_t), C.c_int, C.c_int]
        L.mxGetData.restype = C.c_void_p
        L.mxGetData.argtypes = [C.c_void_p]
        L.mxGetClassID.restype = C.c_int
        L.mxGetClassID.argtypes = [C.c_void_p]
        L.mxGetNumberOfDimensions.restype = C.c_size_t
        L.mxGetNumberOfDimensions.argtypes = [C.c_void_p]
        L.mxGetDimensions.restype = C.POINTER(C.c_size_t)
        L.mxGetDimensions.argtypes = [C.c_void_p]
        L.mxDestroyArray.argtypes = [C.c_void_p]
        L.shim_invoke.restype = C.c_int
        L.shim_invoke.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_int,
                                  C.POINTER(C.c_void_p)]
        L.shim_last_error.restype = C.c_char_p
        L.shim_output.restype = C.c_char_p
        _lib = L
    return _lib


class MxArray:
    """An mxArray owning a copy of a numpy array (column-major)."""

    def __init__(self, a=None, ptr=None):
        L = lib()
        if ptr is not None:
            self.ptr = ptr
            return
        if isinstance(a, (int, float)) and not isinstance(a, bool):
            a = np.float64(a)  # MATLAB literals are double
        a = np.asarray(a)
        if a.dtype == np.bool_:
            cls = 3
            a = a.astype(np.uint8)
        else:
            cls = CLASS[a.dtype]
        dims = a.shape if a.ndim else (1, 1)
        d = (C.c_size_t * len(dims))(*dims)
        self.ptr = L.mxCreateNumericArray(len(dims), d, cls, 0)
        buf = np.asfortranarray(a).reshape(-1, order="F")
        C.memmove(L.mxGetData(self.ptr), buf.ctypes.data, buf.nbytes)

    def numpy(self):
        L = lib()
        cls = L.mxGetClassID(self.ptr)
        nd = L.mxGetNumberOfDimensions(self.ptr)
        dims = tuple(L.mxGetDimensions(self.ptr)[i] for i in range(nd))
        n = int(np.prod(dims))
        dt = np.dtype(DTYPE[cls])
        raw = (C.c_char * (n * dt.itemsize)).from_address(L.mxGetData(self.ptr))
        return np.frombuffer(raw, dtype=dt).copy().reshape(dims, order="F")

    def free(self):
        if self.ptr:
            lib().mxDestroyArray(self.ptr)
            self.ptr = None


def call(name: str, nlhs: int, *args):
    """Calls <name>'s mexFunction with prhs = args (numpy arrays / scalars or
    MxArray); returns (plhs as numpy arrays, printed text).  MxArray
    arguments are left alive so in-place outputs can be read back."""
    L = lib()
    fn = C.cast(getattr(L, f"{name}_mexFunction"), C.c_void_p)
    owned, prhs = [], []
    for a in args:
        if isinstance(a, MxArray):
            prhs.append(a.ptr)
        else:
            m = MxArray(a)
            owned.append(m)
            prhs.append(m.ptr)
    pr = (C.c_void_p * max(1, len(prhs)))(*prhs)
    pl = (C.c_void_p * max(1, nlhs))()
    L.shim_clear_output()
    rc = L.shim_invoke(fn, nlhs, pl, len(prhs), pr)
    out = L.shim_output().decode()
    for m in owned:
        m.free()
    if rc:
        raise MexError(L.shim_last_error().decode())
    res = []
    for i in range(max(1, nlhs)):
        if pl[i]:
            m = MxArray(ptr=pl[i])
            res.append(m.numpy())
            m.free()
    return res, out
