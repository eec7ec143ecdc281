"""Autotuner for the fused primal-dual passes: per size class (kind,
precision, volume shape) it picks the one-iteration pass's tile shape
(tx x ty) and z-chunk depth kz, the two-iteration pass's chunk depth kz2,
and the faster of the two schedules, by timing candidates on the device
with HIP events, and keeps the winners in a JSON table.  Results never
depend on the choice (every choice is bitwise identical), only speed does.

    from qsm_amd import autotune
    autotune.tune(TGV, np.float32, (256, 256, 176))      # measure, store
    autotune.best(TGV, np.float32, (256, 256, 176))      # (tx, ty, kz)
    autotune.best2(TGV, np.float32, (256, 256, 176))     # (kz2, schedule)

The packaged table `tuning_mi355x.json` holds the MI355X measurements made
with `python -m qsm_amd.autotune`; tune() adds to a user table (path in
QSM_TUNING, default ~/.cache/qsm_amd/tuning.json) that takes precedence.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from . import _lib

SHAPES = [(64, 4), (64, 8), (64, 16), (32, 8), (32, 16), (32, 32)]
PACKAGED = Path(__file__).resolve().parent / "tuning_mi355x.json"
USER = Path(os.environ.get("QSM_TUNING", Path.home() / ".cache" / "qsm_amd" / "tuning.json"))
DEFAULT = (64, 8, 16)


def key(kind, dtype, shape):
    k = "tgv" if kind == _lib.TGV else "tv"
    return f"{k}/{np.dtype(dtype).name}/{shape[0]}x{shape[1]}x{shape[2]}"


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except Exception:
        return {}


def table() -> dict:
    t = _load(PACKAGED)
    t.update(_load(USER))
    return t


def kz_candidates(nz: int, n_tiles_xy: int, ncu: int = 256):
    """Chunk depths worth timing: short chunks (many blocks) to whole-column
    marches that leave exactly one or two blocks per CU."""
    inner = max(1, nz - 2)
    c = {8, 12, 16, 22, 32, 44, 58, 88}
    for blocks in (ncu, 2 * ncu, 4 * ncu):
        nch = max(1, -(-blocks // max(1, n_tiles_xy)))
        c.add(-(-inner // nch))
    return sorted(k for k in c if 1 <= k <= inner)


# owned tile of the two-iteration pass per precision (the library's choice)
TILE2 = {"float32": (32, 16), "float64": (16, 16)}


def kz2_candidates(shape, dtype, ncu: int = 256):
    """Chunk depths of the two-iteration pass worth timing: one to four
    blocks per CU, and the whole column."""
    tx, ty = TILE2[np.dtype(dtype).name]
    nxy = -(-shape[0] // tx) * -(-shape[1] // ty)
    inner = max(1, shape[2] - 2)
    c = {inner}
    for blocks in (ncu, 2 * ncu, 3 * ncu, 4 * ncu):
        nch = max(1, blocks // max(1, nxy))
        c.add(-(-inner // nch))
    return sorted(k for k in c if 1 <= k <= inner)


def _timed(s, st, iters, repeats):
    import torch

    s.iterate(4)
    st.synchronize()
    ts = []
    for _ in range(repeats):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        s.iterate(iters)
        e1.record(st)
        e1.synchronize()
        ts.append(e0.elapsed_time(e1) / iters)
    return float(np.median(ts))


def tune(kind, dtype, shape, iters=40, repeats=3, shapes=SHAPES, store=True, log=None,
         one=True):
    """Times every (tx, ty, kz) candidate of the one-iteration pass (unless
    one=False: then the stored entry's are kept) and every kz2 of the
    two-iteration pass on the current device; returns the table entry
    dict(tx, ty, kz, ms, kz2, ms2, schedule), schedule 0 when the
    two-iteration pass is faster, else 2."""
    import torch

    from .pd import Solver

    dtype = np.dtype(dtype)
    rng = np.random.default_rng(0)
    d2f = np.asfortranarray((rng.standard_normal(shape) * 1e-2).astype(dtype))
    m = np.ones(shape, np.uint8, order="F")
    st = torch.cuda.Stream()
    k = key(kind, dtype, shape)
    best = dict(table().get(k, {})) if not one else None
    with Solver(kind, dtype, shape, (1, 1, 1), 1e-3, 2.0, stream=st) as s:
        s.load(d2f, m)
        s.set_schedule(2)
        if one or not best:
            best = None
            for tx, ty in shapes:
                nxy = -(-shape[0] // tx) * -(-shape[1] // ty)
                for kz in kz_candidates(shape[2], nxy):
                    try:
                        s.set_tiles(tx, ty, kz)
                    except _lib.QsmError:
                        continue
                    ms = _timed(s, st, iters, repeats)
                    if log:
                        log(f"{k} one-iteration {tx}x{ty} kz={kz}: {ms:.4f} ms")
                    if best is None or ms < best["ms"]:
                        best = {"tx": tx, "ty": ty, "kz": kz, "ms": ms}
    # the two-iteration pass (even counts: every launch a pair)
    with Solver(kind, dtype, shape, (1, 1, 1), 1e-3, 2.0, stream=st) as s:
        s.load(d2f, m)
        for kz2 in kz2_candidates(shape, dtype):
            try:
                s.set_kz2(kz2)
            except _lib.QsmError:
                continue
            ms = _timed(s, st, iters, repeats)
            if log:
                log(f"{k} two-iteration kz2={kz2}: {ms:.4f} ms")
            if "ms2" not in best or ms < best["ms2"]:
                best["kz2"], best["ms2"] = kz2, ms
    if "ms2" in best:
        best["schedule"] = 0 if best["ms2"] < best["ms"] else 2
    if store and best is not None:
        t = _load(USER)
        t[k] = best
        USER.parent.mkdir(parents=True, exist_ok=True)
        USER.write_text(json.dumps(t, indent=1, sort_keys=True))
    return best


def default_kz2(dtype, shape, ncu: int = 256):
    """The library's default chunk depth of the two-iteration pass: at
    most one block per CU."""
    tx, ty = TILE2[np.dtype(dtype).name]
    nxy = -(-shape[0] // tx) * -(-shape[1] // ty)
    inner = max(1, shape[2] - 2)
    nch = max(1, ncu // nxy)
    return max(1, -(-inner // nch))


def batch_kz2(shape):
    """Chunk depth of the two-iteration march in a batch: half the column
    (two tiles deep).  The batch claims tiles dynamically, so the one block
    per CU of a lone volume no longer matters; halves measured fastest on
    the 48-volume mix of 96^3 to 256x256x176 (8.9e10 voxel-iters/s FP32,
    against 8.5e10 for 64 planes, 8.3e10 for the lone-volume depths and
    8.1e10 for whole columns)."""
    return max(1, -(-max(1, shape[2] - 2) // 2))


def best2(kind, dtype, shape):
    """(kz2, schedule) of this size class's entry, or None when the table
    has no two-iteration measurement for it (the library's defaults then
    apply: schedule 0, at most one block per CU)."""
    e = table().get(key(kind, dtype, shape))
    if e and "kz2" in e:
        return int(e["kz2"]), int(e.get("schedule", 0))
    return None


def best(kind, dtype, shape, tiles=None):
    """(tx, ty, kz) for this size class: the exact entry, else the entry of
    the same kind/precision with the nearest voxel count (kz clipped), else
    the default.  With tiles=(tx, ty) only entries of that shape count."""
    t = table()
    k = key(kind, dtype, shape)
    e = t.get(k)
    if e and (tiles is None or (e["tx"], e["ty"]) == tuple(tiles)):
        return e["tx"], e["ty"], e["kz"]
    prefix = k.rsplit("/", 1)[0] + "/"
    n = shape[0] * shape[1] * shape[2]
    cands = []
    for kk, v in t.items():
        if not kk.startswith(prefix):
            continue
        if tiles is not None and (v["tx"], v["ty"]) != tuple(tiles):
            continue
        dims = [int(x) for x in kk.rsplit("/", 1)[1].split("x")]
        cands.append((abs(np.log(dims[0] * dims[1] * dims[2] / n)), v, dims))
    if cands:
        _, v, dims = min(cands, key=lambda c: c[0])
        kz = max(1, min(v["kz"], shape[2] - 2)) if shape[2] > 2 else 1
        return v["tx"], v["ty"], kz
    if tiles is not None:
        return tiles[0], tiles[1], DEFAULT[2]
    return DEFAULT


def main(argv=None):
    import argparse
    import sys

    ap = argparse.ArgumentParser(description="tune the fused pass per size class")
    ap.add_argument("--out", default=None, help="write the table here (default: user table)")
    ap.add_argument("--sizes", default="96x96x96,128x128x128,160x160x96,224x224x128,256x256x176,320x320x192")
    ap.add_argument("--iters", type=int, default=40)
    ap.add_argument("--two-only", action="store_true",
                    help="keep the table's one-iteration entries, time only kz2")
    a = ap.parse_args(argv)
    res = {}
    for sz in a.sizes.split(","):
        shape = tuple(int(x) for x in sz.split("x"))
        for kind in (_lib.TGV, _lib.TV):
            for dt in (np.float32, np.float64):
                b = tune(kind, dt, shape, iters=a.iters, store=False, one=not a.two_only,
                         log=lambda s: print(s, file=sys.stderr, flush=True))
                res[key(kind, dt, shape)] = b
                print(key(kind, dt, shape), b, flush=True)
    if a.out:
        Path(a.out).write_text(json.dumps(res, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
