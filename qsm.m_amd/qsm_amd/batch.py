"""Batched primal-dual inversions on one GPU (qsm_batch_* in
include/qsm_hip.h): many uneven TV/TGV volumes advanced by one persistent
launch per iteration, iterations replayed from captured hipGraphs, volumes
admitted and retired between chunks against an HBM budget."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, dtype_code, fortran, h3, lib, mask_u8, ptr, sz3

_p = C.c_void_p
_sz = C.c_size_t
_lib.register("qsm_batch_create", C.c_int, [C.POINTER(_p), C.c_int, C.c_int, C.c_int, _sz, C.c_int, _p])
_lib.register("qsm_batch_volume_bytes", _sz, [C.c_int, C.c_int, C.POINTER(_sz), C.c_int, C.c_int, C.c_int, C.c_uint32])
_lib.register("qsm_batch_add", C.c_int, [_p, C.c_int, C.POINTER(_sz), C.POINTER(C.c_double), C.c_double,
                                          C.c_double, _p, _p, C.c_int, C.c_uint32, C.c_double, C.c_int,
                                          C.POINTER(C.c_int)])
_lib.register("qsm_batch_step", C.c_int, [_p, C.c_uint32])
_lib.register("qsm_batch_sync", C.c_int, [_p])
_lib.register("qsm_batch_state", C.c_int, [_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_uint32)])
_lib.register("qsm_batch_result", C.c_int, [_p, C.c_int, _p, C.POINTER(C.c_uint32), C.POINTER(C.c_int),
                                             C.POINTER(C.c_double)])
_lib.register("qsm_batch_remove", C.c_int, [_p, C.c_int])
_lib.register("qsm_batch_set_graph", C.c_int, [_p, C.c_int])
_lib.register("qsm_batch_used_bytes", _sz, [_p])
_lib.register("qsm_batch_destroy", None, [_p])


@dataclass
class SlotResult:
    chi: np.ndarray
    iters: int
    stopped: bool
    norms: tuple


class Batch:
    """A set of resident inversions advanced together.

    tiles=(0, 0) (the default) runs the two-iteration march (tiles of 32x16
    voxels in FP32, 16x16 in FP64); a supported one-iteration shape such as
    (64, 8) runs the one-iteration march instead.  Results are bitwise the
    same either way."""

    def __init__(self, dtype=np.float32, tiles=(0, 0), hbm_budget=64 << 30,
                 max_volumes=64, stream=None):
        self.dtype = np.dtype(dtype)
        self.tiles = tuple(tiles)
        self.budget = int(hbm_budget)
        h = C.c_void_p()
        sp = stream.cuda_stream if stream is not None else None
        check(lib().qsm_batch_create(C.byref(h), dtype_code(self.dtype), int(tiles[0]), int(tiles[1]),
                                     int(hbm_budget), int(max_volumes), sp))
        self._h = h
        self.shapes = {}

    def volume_bytes(self, kind, shape, kz, maxiter):
        return int(lib().qsm_batch_volume_bytes(kind, dtype_code(self.dtype), sz3(shape),
                                                self.tiles[0], self.tiles[1], int(kz), int(maxiter)))

    def kz_for(self, kind, shape):
        """The tuned z-chunk depth of this size class for the batch's march."""
        from . import autotune

        if tuple(self.tiles) == (0, 0):
            return autotune.batch_kz2(shape)
        return autotune.best(kind, self.dtype, shape, tiles=self.tiles)[2]

    def add(self, kind, d2f, mask, vsz, lam, alpha=0.0, maxiter=1000, tol=0.0, kz=None):
        """Admits one inversion; returns its slot (raises QsmAdmissionError
        when the HBM budget or the slots are exhausted).  kz=None takes the
        tuned depth of the size class."""
        slot = C.c_int()
        if kz is None:
            kz = self.kz_for(kind, tuple(int(s) for s in d2f.shape))
        if hasattr(d2f, "data_ptr"):
            shape = tuple(int(s) for s in d2f.shape)
            check(lib().qsm_batch_add(self._h, kind, sz3(shape), h3(vsz), float(lam), float(alpha),
                                      d2f.data_ptr(), mask.data_ptr(), 1, int(maxiter), float(tol),
                                      int(kz), C.byref(slot)))
        else:
            f = fortran(np.asarray(d2f), self.dtype)
            m = mask_u8(mask)
            shape = f.shape
            check(lib().qsm_batch_add(self._h, kind, sz3(shape), h3(vsz), float(lam), float(alpha),
                                      ptr(f), ptr(m), 0, int(maxiter), float(tol), int(kz),
                                      C.byref(slot)))
        self.shapes[slot.value] = shape
        return slot.value

    def step(self, chunk=10):
        check(lib().qsm_batch_step(self._h, int(chunk)))

    def sync(self):
        check(lib().qsm_batch_sync(self._h))

    def state(self, slot):
        u, l, it = C.c_int(), C.c_int(), C.c_uint32()
        check(lib().qsm_batch_state(self._h, int(slot), C.byref(u), C.byref(l), C.byref(it)))
        return bool(u.value), bool(l.value), it.value

    def result(self, slot, with_chi=True):
        it, st = C.c_uint32(), C.c_int()
        nr = (C.c_double * 2)()
        chi = np.empty(self.shapes[slot], dtype=self.dtype, order="F") if with_chi else None
        check(lib().qsm_batch_result(self._h, int(slot), ptr(chi) if with_chi else None,
                                     C.byref(it), C.byref(st), nr))
        return SlotResult(chi, it.value, bool(st.value), (nr[0], nr[1]))

    def remove(self, slot):
        check(lib().qsm_batch_remove(self._h, int(slot)))
        self.shapes.pop(slot, None)

    def set_graph(self, on: bool):
        check(lib().qsm_batch_set_graph(self._h, 1 if on else 0))

    @property
    def used_bytes(self):
        return int(lib().qsm_batch_used_bytes(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().qsm_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
