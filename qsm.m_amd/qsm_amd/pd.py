"""Primal-dual TV/TGV dipole inversion on the GPU.

`pd_tgv_mex` / `pd_tv_mex` mirror the reference MEX entry points
(src/inversion/mex/pd_tgv_mex.c:104-135, pd_tv_mex.c:100-130): same
positional arguments, same meaning, a fresh chi array returned, the input
left untouched, the precision chosen by the input's dtype (single/double).

`Solver` exposes the resident handle (include/qsm_hip.h qsm_pd_*) for callers
that keep the state in HBM between calls (benchmark, batching, serving).
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _lib
from ._lib import check, dtype_code, fortran, h3, lib, mask_u8, ptr, sz3


def _printer(verbose):
    if not verbose:
        return _lib.PRINT_FN(), None

    def emit(line, _ctx):
        sys.stdout.write(line.decode())

    cb = _lib.PRINT_FN(emit)
    return cb, cb


def _inputs(d2f, mask):
    f = np.asarray(d2f)
    if f.dtype not in (np.float32, np.float64):
        f = f.astype(np.float64)
    f = fortran(f)
    m = mask_u8(mask)
    if m.shape != f.shape:
        raise ValueError(f"mask shape {m.shape} != d2f shape {f.shape}")
    return f, m


def pd_tgv_mex(d2f, mask, vsz, lam, alpha, tol, maxiter, verbose=0):
    """chi = pd_tgv_mex(d2f, mask, vsz, lam, alpha, tol, maxiter, verbose)

    TGV primal-dual inversion (alpha1 = lam, alpha0 = alpha*lam) of the
    Laplacian-domain input d2f = -lapw(phase) on the mask.  Arrays are 3-D,
    column-major (x fastest) as in MATLAB.
    """
    f, m = _inputs(d2f, mask)
    chi = np.zeros(f.shape, dtype=f.dtype, order="F")
    cb, keep = _printer(verbose)
    check(lib().qsm_pd_tgv(dtype_code(f.dtype), ptr(f), ptr(m), sz3(f.shape), h3(vsz),
                           float(lam), float(alpha), float(tol), int(maxiter),
                           int(verbose), ptr(chi), cb, None))
    del keep
    return chi


def pd_tv_mex(d2f, mask, vsz, lam, tol, maxiter, verbose=0):
    """chi = pd_tv_mex(d2f, mask, vsz, lam, tol, maxiter, verbose)"""
    f, m = _inputs(d2f, mask)
    chi = np.zeros(f.shape, dtype=f.dtype, order="F")
    cb, keep = _printer(verbose)
    check(lib().qsm_pd_tv(dtype_code(f.dtype), ptr(f), ptr(m), sz3(f.shape), h3(vsz),
                          float(lam), float(tol), int(maxiter), int(verbose),
                          ptr(chi), cb, None))
    del keep
    return chi


def solve_static(kind: int, d2f, mask, shape, vsz, lam, alpha, tol, maxiter, chi, stream=None):
    """qsm_pd_solve_static: the solve of pd_tgv_mex (kind TGV) or pd_tv_mex
    (kind TV) on torch device tensors holding column-major volumes of
    `shape` (d2f and chi of one float dtype, mask uint8), with the solver
    state in the library's static work space (no device allocation).
    Asynchronous on `stream` (a torch.cuda.Stream, or None: the null
    stream); returns chi."""
    import torch

    if d2f.dtype not in (torch.float32, torch.float64) or chi.dtype != d2f.dtype:
        raise ValueError("d2f and chi must share one float dtype")
    n = int(np.prod(shape))
    if d2f.numel() != n or chi.numel() != n or mask.numel() != n or mask.dtype != torch.uint8:
        raise ValueError("d2f, mask (uint8) and chi must hold prod(shape) elements")
    code = _lib.F32 if d2f.dtype == torch.float32 else _lib.F64
    sp = stream.cuda_stream if stream is not None else None
    check(lib().qsm_pd_solve_static(int(kind), code, d2f.data_ptr(), mask.data_ptr(),
                                    sz3(shape), h3(vsz), float(lam), float(alpha), float(tol),
                                    int(maxiter), chi.data_ptr(), sp))
    return chi


def step_sizes(vsz):
    """(tau, sigma) of pd_tgv_mex.c:272-278."""
    t, s = C.c_double(), C.c_double()
    lib().qsm_step_sizes(h3(vsz), C.byref(t), C.byref(s))
    return t.value, s.value


def footprint(kind: int, dtype, shape) -> int:
    """Device bytes a resident solve of this size allocates."""
    return int(lib().qsm_pd_footprint(kind, dtype_code(dtype), sz3(shape)))


class Solver:
    """Resident primal-dual solve (qsm_pd_* handle).

    Inputs may be numpy arrays (host) or torch CUDA tensors (device); the
    state stays in HBM until close().  `stream` is a torch.cuda.Stream or
    None (null stream).
    """

    def __init__(self, kind: int, dtype, shape, vsz, lam, alpha=0.0, stream=None):
        self.kind, self.shape = kind, tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self._stream = stream
        sp = stream.cuda_stream if stream is not None else None
        h = C.c_void_p()
        check(lib().qsm_pd_create(C.byref(h), kind, dtype_code(self.dtype), sz3(self.shape),
                                  h3(vsz), float(lam), float(alpha), sp))
        self._h = h

    @property
    def handle(self):
        return self._h

    def load(self, d2f, mask):
        if hasattr(d2f, "data_ptr"):
            self._keep = (d2f, mask)
            check(lib().qsm_pd_load(self._h, d2f.data_ptr(), mask.data_ptr(), 1))
        else:
            f, m = _inputs(d2f, mask)
            if f.dtype != self.dtype:
                f = f.astype(self.dtype, order="F")
            check(lib().qsm_pd_load(self._h, ptr(f), ptr(m), 0))
        return self

    def set_schedule(self, schedule: int):
        """0: two-iteration passes (default), 1: the reference's two sweeps,
        2: one-iteration passes only; identical results."""
        check(lib().qsm_pd_set_schedule(self._h, int(schedule)))
        return self

    def set_tiles(self, tx: int, ty: int, kz: int):
        """Tile shape of the one-iteration fused pass."""
        check(lib().qsm_pd_set_tiles(self._h, int(tx), int(ty), int(kz)))
        return self

    def set_kz2(self, kz: int):
        """z-chunk depth of the two-iteration fused pass."""
        check(lib().qsm_pd_set_kz2(self._h, int(kz)))
        return self

    def set_active_items(self, on: int):
        """Work list of active tiles for masked volumes: 0 off, 1 where the
        cost model favours it (default), 2 always; identical results."""
        check(lib().qsm_pd_set_active_items(self._h, int(on)))
        return self

    def active_items(self):
        """(items, chunk depth) of the current work list; (0, 0): uniform grid."""
        n, kz = C.c_int(), C.c_int()
        check(lib().qsm_pd_active_items(self._h, C.byref(n), C.byref(kz)))
        return n.value, kz.value

    def launch_counts(self) -> dict:
        """Kernel launches enqueued since creation, by kind."""
        c = (C.c_uint64 * 4)()
        check(lib().qsm_pd_launch_counts(self._h, c))
        return {"pass2": int(c[0]), "pass1": int(c[1]), "sweeps": int(c[2]), "checks": int(c[3])}

    def autotune(self):
        """Applies the tuned choices of this size class (qsm_amd.autotune
        table): the one-iteration pass's (tx, ty, kz), and, where measured,
        the two-iteration pass's kz2 and the faster schedule.  Results are
        bitwise independent of the choices."""
        from . import autotune

        self.set_tiles(*autotune.best(self.kind, self.dtype, self.shape))
        b2 = autotune.best2(self.kind, self.dtype, self.shape)
        if b2 is not None:
            self.set_kz2(b2[0])
            self.set_schedule(b2[1])
        return self

    def tiles(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().qsm_pd_get_tiles(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def iterate(self, niter: int, tol: float = 0.0):
        check(lib().qsm_pd_iterate(self._h, int(niter), float(tol)))
        return self

    def status(self):
        it, st = C.c_uint32(), C.c_int()
        nr = (C.c_double * 2)()
        check(lib().qsm_pd_status(self._h, C.byref(it), C.byref(st), nr))
        return it.value, bool(st.value), (nr[0], nr[1])

    def norm_history(self, cap: int = 4096):
        buf = (C.c_double * (2 * cap))()
        n = lib().qsm_pd_norm_history(self._h, buf, cap)
        return np.frombuffer(buf, dtype=np.float64, count=2 * n).reshape(n, 2).copy()

    def chi(self, out=None):
        """chi as a numpy array, or into a torch device tensor `out`."""
        if out is not None:
            check(lib().qsm_pd_get_chi(self._h, out.data_ptr(), 1))
            return out
        a = np.empty(self.shape, dtype=self.dtype, order="F")
        check(lib().qsm_pd_get_chi(self._h, ptr(a), 0))
        return a

    def close(self):
        if getattr(self, "_h", None):
            lib().qsm_pd_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
