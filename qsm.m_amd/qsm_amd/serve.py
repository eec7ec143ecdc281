"""Serving loop: inversion requests of mixed sizes arrive over time and are
served by continuous batching on one GPU.

* Requests wait in arrival order; at every chunk boundary (a fixed number of
  iterations, one captured hipGraph replay) the loop admits waiting requests
  whose resident footprint fits the HBM budget (FIFO; a request that does
  not fit blocks the ones behind it, so large requests are not starved).
* Optionally the live work is capped (max_live_voxels): enough resident
  voxels to keep the GPU saturated, not more.  All live requests share the
  device, so admitting everything at once would make every request finish
  near the end of the burst; the cap lets requests complete in arrival
  order, which lowers the median latency at the same throughput.
* Admitted requests keep field, mask, primal and dual state resident; all
  live requests advance together in the same launches, so work that
  arrives while others run overlaps with them on the device.
* A request leaves when its iteration budget is spent (or its tolerance
  test stops it); its chi is copied back and its HBM released.
* Latency = time its chi is on the host - arrival time.

`Server.run` drives a pre-built request list against a wall clock (arrival
times are offsets from the start), which is how the benchmark replays a
Poisson stream.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from ._lib import QsmAdmissionError
from .batch import Batch

QSM_EFULL = 5


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    kind: int
    d2f: np.ndarray
    mask: np.ndarray
    vsz: tuple
    lam: float
    alpha: float
    maxiter: int
    tol: float = 0.0
    kz: int | None = None    # z-chunk depth; None: the tuned one
    # filled in by the server
    admitted: float = float("nan")
    done: float = float("nan")
    iters: int = 0
    chi: np.ndarray | None = field(default=None, repr=False)

    @property
    def latency(self):
        return self.done - self.arrival


@dataclass
class ServeStats:
    n: int
    wall: float
    voxel_iters: int
    p50: float
    p99: float
    mean_wait: float
    peak_live: int

    @property
    def voxel_iters_per_s(self):
        return self.voxel_iters / self.wall if self.wall > 0 else 0.0


class Server:
    def __init__(self, batch: Batch, chunk: int = 10, keep_chi: bool = True,
                 max_live_voxels: int | None = None):
        self.b = batch
        self.chunk = chunk
        self.keep_chi = keep_chi
        self.max_live_voxels = max_live_voxels

    def run(self, requests, clock=time.perf_counter, sleep=time.sleep):
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = collections.deque(reqs)
        waiting = collections.deque()
        active = {}  # slot -> request
        live_vox = 0
        peak = 0
        t0 = clock()
        while pending or waiting or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admit in arrival order while the head fits
            while waiting:
                r = waiting[0]
                if r.kz is None:
                    r.kz = self.b.kz_for(r.kind, r.d2f.shape)
                need = self.b.volume_bytes(r.kind, r.d2f.shape, r.kz, r.maxiter)
                if need > self._budget():
                    raise QsmAdmissionError(QSM_EFULL, f"request {r.rid} needs {need} bytes, "
                                            f"more than the whole budget {self._budget()}")
                if self.b.used_bytes + need > self._budget():
                    break
                vox = int(np.prod(r.d2f.shape))
                if (self.max_live_voxels is not None and active
                        and live_vox + vox > self.max_live_voxels):
                    break
                try:
                    slot = self.b.add(r.kind, r.d2f, r.mask, r.vsz, r.lam, r.alpha,
                                      r.maxiter, r.tol, r.kz)
                except QsmAdmissionError:
                    if not active:
                        raise  # refused with nothing resident: it never fits
                    break
                waiting.popleft()
                r.admitted = clock() - t0
                active[slot] = r
                live_vox += vox
            peak = max(peak, len(active))
            if not active:
                if pending:
                    sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                continue
            self.b.step(self.chunk)
            self.b.sync()
            for slot in list(active):
                used, live, it = self.b.state(slot)
                if live:
                    continue
                r = active.pop(slot)
                live_vox -= int(np.prod(r.d2f.shape))
                res = self.b.result(slot, with_chi=True)
                r.done = clock() - t0
                r.iters = res.iters
                r.chi = res.chi if self.keep_chi else None
                self.b.remove(slot)
        wall = clock() - t0
        lat = np.array([r.latency for r in reqs]) if reqs else np.zeros(1)
        vox = sum(int(np.prod(r.d2f.shape)) * r.iters for r in reqs)
        wait = np.array([r.admitted - r.arrival for r in reqs]) if reqs else np.zeros(1)
        return ServeStats(n=len(reqs), wall=wall, voxel_iters=vox,
                          p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                          mean_wait=float(wait.mean()), peak_live=peak)

    def _budget(self):
        return getattr(self.b, "budget", 1 << 62)


def request_from_nii(rid: int, arrival: float, phase_file, mask_file, kind: int, lam: float,
                     alpha: float, maxiter: int, tol: float = 0.0, dtype=np.float32) -> Request:
    """A request whose payload is a pair of NIfTI files, as a user of
    sstgv.m holds them: the local-field phase and its mask (loadNii.m), the
    voxel size from the phase's header.  d2f = -lapw(phase) on the device
    (sstgv.m:67), in the solver's precision."""
    from . import fd
    from .nifti import loadNii

    ph = loadNii(phase_file)
    mk = loadNii(mask_file)
    phase = np.asfortranarray(np.asarray(ph["img"]).astype(dtype))
    if phase.ndim != 3 or mk["img"].shape != phase.shape:
        raise ValueError("phase must be 3-D and the mask of the same size")
    vsz = tuple(float(v) for v in ph["hdr"]["dime"]["pixdim"][1:4])
    d2f = -fd.lapw(phase, vsz)
    mask = np.asfortranarray(np.asarray(mk["img"]) != 0).astype(np.uint8)
    return Request(rid=rid, arrival=arrival, kind=kind, d2f=np.asfortranarray(d2f), mask=mask,
                   vsz=vsz, lam=lam, alpha=alpha, maxiter=maxiter, tol=tol)


def save_chi(r: Request, filename):
    """Write a finished request's susceptibility map (saveNii.m) with its
    voxel size."""
    from .nifti import saveNii

    if r.chi is None:
        raise ValueError(f"request {r.rid} holds no chi (not finished, or keep_chi off)")
    return saveNii(filename, r.chi, r.vsz)


def poisson_arrivals(rate: float, n: int, rng) -> np.ndarray:
    """n arrival times of a Poisson process of the given rate (1/s)."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
