#!/usr/bin/env python3
"""bench.py -- TGV-QSM voxel-iterations/s at 256x256x176 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (C3, the size the metric is quoted at): seeded 60-ellipsoid phantom,
256x256x176 at 1 mm, B0 = z, full-box mask, TGV alpha1 = 1e-3,
alpha0 = 2*alpha1, tol = 0.  A step is one primal-dual iteration over the whole
volume (half a two-iteration fused pass; the every-10th-iteration
convergence check is included).  The input is resident in HBM before timing
starts.

Reported: FP32 voxel-iters/s (value) and the FP64 drop-in beside it, chi
error of FP32 against FP64 (the FP64 path is bitwise the oracle's), the
roofline of the fused kernel measured with HIP events on its stream, and the
CPU baseline (the oracle, the reference's C loop restated, on a bounded
sample on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "qsm.m_amd", ROOT):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

METRIC = ("TGV-QSM voxel-iters/s at 256²×176 (FP32 vs FP64, same GPU); "
          "p99 solve latency")
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = ROOT / "profiles" / "traffic_fused2_tgv_c3.json"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_input(cfg_name: str):
    from qsm_amd import fd, phantom
    from qsm_amd.configs import CONFIGS

    cfg = CONFIGS[cfg_name]
    t = time.perf_counter()
    ph = phantom.make(cfg["spec"], field="gpu")  # GPU dipole convolution
    d2f = -fd.lapw(ph["field"], cfg["spec"].vsz)  # GPU lapw
    log(f"[bench] phantom {cfg['spec'].shape} ready in {time.perf_counter() - t:.1f}s")
    return (cfg, np.asfortranarray(d2f), ph["mask"].astype(np.uint8, order="F"),
            np.asfortranarray(ph["field"]))


def time_solver(kind, dtype, cfg, d2f, mask, steps, warmup, settle=500):
    """Times `steps` iterations after `warmup` untimed ones, with HIP events
    on the solver's stream.  Before the warmup the same solve runs `settle`
    (or settle + 1) untimed iterations: the GPU's clocks settle over the
    first tens of milliseconds under load (launch times of the C3 pass vary
    by +-8 % there), which a short timed region would otherwise sample
    instead of the steady state a 1000-iteration solve runs at.  The count's
    parity makes settle + warmup even, so the timed region starts right
    after a pair of iterations and every convergence check falls on the
    second iteration of a two-iteration launch: the region holds
    two-iteration launches only (plus one single iteration at the end when
    `steps` is odd)."""
    import torch
    from qsm_amd import pd

    shape = d2f.shape
    stream = torch.cuda.Stream()
    s = pd.Solver(kind, dtype, shape, cfg["spec"].vsz, cfg["lam"], cfg["alpha"], stream=stream)
    s.autotune()
    tiles = s.tiles()
    from qsm_amd import autotune
    b2 = autotune.best2(kind, dtype, shape)
    s.load(d2f.astype(dtype, order="F"), mask)
    items = s.active_items()
    settled = settle + ((settle + warmup) & 1)
    s.iterate(settled, 0.0)
    s.iterate(warmup, 0.0)
    stream.synchronize()
    torch.cuda.synchronize()
    c0 = s.launch_counts()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    e0.record(stream)
    s.iterate(steps, 0.0)
    e1.record(stream)
    e1.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    gpu = e0.elapsed_time(e1) / 1e3
    c1 = s.launch_counts()
    it, stopped, _ = s.status()
    chi = s.chi()
    s.close()
    assert it == settled + warmup + steps and not stopped
    launches = {k: c1[k] - c0[k] for k in c1}
    return {"gpu_s": gpu, "wall_s": wall, "chi": chi, "tiles": tiles, "launches": launches,
            "kz2": b2[0] if b2 else None, "schedule": b2[1] if b2 else 0,
            "work_items": items[0], "work_item_kz": items[1], "settle_iterations": settled,
            "iterations": settled + warmup + steps}


def stream_ceiling_gbs(n=256 * 256 * 176, narr=15):
    """Bandwidth of the library's 15-arrays-in, 15-out streaming kernel
    (qsm_stream_dev) at the C3 size: the ceiling of the fused pass's access
    pattern (its state traffic with none of its arithmetic)."""
    import ctypes as C

    import torch
    from qsm_amd import _lib
    from qsm_amd._lib import check, lib

    _lib.register("qsm_stream_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int,
                                              C.c_void_p])
    a = torch.zeros(narr * n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        check(lib().qsm_stream_dev(b.data_ptr(), a.data_ptr(), n, narr, None))
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        check(lib().qsm_stream_dev(b.data_ptr(), a.data_ptr(), n, narr, None))
    e1.record()
    e1.synchronize()
    dt = e0.elapsed_time(e1) / 1e3 / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * narr * n * 4 / dt / 1e9


def brain_leg(steps=200, warmup=20):
    """C3 with the ellipsoidal brain mask (semi-axes 0.40, 0.45, 0.40 of the
    volume): FP32 TGV on the same 256x256x176 grid, algorithmic bytes
    N_active * 41 words + N mask bytes per iteration."""
    from qsm_amd import _lib

    cfg, d2f, mask, _ = make_input("c3_brain")
    r = time_solver(_lib.TGV, np.float32, cfg, d2f, mask, steps, warmup)
    N = int(np.prod(d2f.shape))
    act = int(np.count_nonzero(mask))
    t = r["gpu_s"] / steps
    return {"ms_per_iter": t * 1e3, "voxel_iters_per_s": N / t, "n_active": act, "n": N,
            "active_fraction": act / N, "algorithmic_gbs": (act * 41 * 4 + N) / t / 1e9,
            "frac": (act * 41 * 4 + N) / t / 1e9 / HBM_PEAK_GBS,
            "work_items": r["work_items"], "work_item_kz": r["work_item_kz"],
            "tiles": r["tiles"], "steps": steps}


def copy_peak_gbs():
    """Device-to-device bandwidth of the library's plain 16-byte vector copy
    kernel (qsm_copy_dev) over 2 GiB: the achievable-peak reference."""
    import torch
    from qsm_amd._lib import check, lib

    n = 2 ** 29  # float32 elements -> 2 GiB
    a = torch.empty(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    torch.cuda.synchronize()
    nbytes = n * 4
    for _ in range(3):
        check(lib().qsm_copy_dev(b.data_ptr(), a.data_ptr(), nbytes, None))
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        check(lib().qsm_copy_dev(b.data_ptr(), a.data_ptr(), nbytes, None))
    e1.record()
    e1.synchronize()
    dt = e0.elapsed_time(e1) / 1e3 / reps
    del a, b
    torch.cuda.empty_cache()
    return 2 * n * 4 / dt / 1e9


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cfg, d2f, mask, iters, runs=3, iters_1t=4):
    """The oracle's pd_tgv (the C restatement of mx_pdd) on the C3 input:
    median of `runs` runs on all the host threads (OMP_NUM_THREADS or the
    affinity set), and one run of iters_1t iterations on one thread."""
    from oracle import oracle as orc

    x = d2f.astype(np.float64, order="F")
    N = int(np.prod(d2f.shape))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    loops, walls = [], []
    chi = None
    for _ in range(runs):
        t = time.perf_counter()
        chi = orc.pd_tgv(x, mask, cfg["spec"].vsz, cfg["lam"], cfg["alpha"], 0.0, iters)[0]
        walls.append(time.perf_counter() - t)
        loops.append(orc.last_loop_seconds())
    loop = float(np.median(loops))
    wall = float(np.median(walls))
    one = None
    if iters_1t > 0:
        orc.set_threads(1)
        try:
            ones = []
            for _ in range(runs):
                orc.pd_tgv(x, mask, cfg["spec"].vsz, cfg["lam"], cfg["alpha"], 0.0, iters_1t)
                ones.append(orc.last_loop_seconds())
            one = N * iters_1t / float(np.median(ones))
        finally:
            orc.set_threads(threads)
    return {
        "value": N * iters / loop,
        "unit": "voxel-iters/s",
        "cores": threads,
        "kind": "port",
        "value_1_thread": one,
        "host_cpu": _cpu_model(),
        "sample": (f"oracle pd_tgv float64 (C11+OpenMP restatement of pd_tgv_mex.c, "
                   f"-O2, no FMA), C3 input {d2f.shape[0]}x{d2f.shape[1]}x{d2f.shape[2]} "
                   f"full-box mask, {iters} iterations, tol=0; median of {runs} runs: "
                   f"iteration-loop time {loop:.2f}s (end-to-end {wall:.2f}s), {threads} "
                   f"OpenMP threads; 1 thread: median of {runs} runs of {iters_1t} "
                   f"iterations"),
        "_chi": chi,
        "_iters": iters,
    }


def fp64_agreement(cfg, d2f, mask, cpu_chi, iters):
    """The FP64 drop-in against the CPU path on the same input: chi of the
    GPU solver after the same iterations as the CPU baseline's run
    (pd_tgv_mex.c:130 returns that chi)."""
    from qsm_amd import _lib, pd

    with pd.Solver(_lib.TGV, np.float64, d2f.shape, cfg["spec"].vsz, cfg["lam"],
                   cfg["alpha"]) as s:
        s.load(d2f.astype(np.float64, order="F"), mask)
        s.iterate(iters)
        g = s.chi()
    n = float(np.linalg.norm(cpu_chi))
    return {"iterations": iters, "rel_l2": float(np.linalg.norm(g - cpu_chi) / n) if n else 0.0,
            "max_abs": float(np.max(np.abs(g - cpu_chi))),
            "bitwise_equal": bool(np.array_equal(g, cpu_chi)),
            "cpu": "oracle pd_tgv float64 (the CPU baseline's last run)"}


def accuracy_leg(cfg, d2f, mask, iters=1000):
    """The precision variant's accuracy at the config's own iteration count,
    independent of --steps: FP32 chi against FP64 chi (the FP64 GPU chi is
    bitwise the oracle's) after `iters` iterations, TGV and TV."""
    from qsm_amd import _lib, pd

    out = {"iterations": iters}
    for name, kind, alpha in (("tgv", _lib.TGV, cfg["alpha"]), ("tv", _lib.TV, 0.0)):
        chis = {}
        for dt in (np.float64, np.float32):
            with pd.Solver(kind, dt, d2f.shape, cfg["spec"].vsz, cfg["lam"], alpha) as s:
                s.autotune()
                s.load(d2f.astype(dt, order="F"), mask)
                s.iterate(iters)
                chis[dt] = s.chi().astype(np.float64)
        c64, c32 = chis[np.float64], chis[np.float32]
        n = float(np.linalg.norm(c64))
        out[name] = {"rel_l2": float(np.linalg.norm(c32 - c64) / n) if n else 0.0,
                     "max_abs": float(np.max(np.abs(c32 - c64))),
                     "max_abs_chi": float(np.max(np.abs(c64)))}
    return out


def fd_leg(shape=(256, 256, 176), reps=20):
    """The finite-difference entry points at the C3 size on device arrays
    (both precisions), against their algorithmic bytes: gradients read 1
    and write 3 arrays, the adjoints read 3 and write 1, lapw reads 1 and
    writes 1, lap1 (on a full mask) reads 1 array and the 1-byte mask and
    writes 1.  lapw is timed on what it is given in use, a wrapped phase
    (values in [-pi, pi)), and as lapw_jumps on samples 9 standard normals
    wide, whose differences mostly exceed 3 pi and take the fmod path; the
    other operators run on standard normals."""
    import ctypes as C

    import torch
    from qsm_amd import _lib

    L = _lib.lib()
    N = int(np.prod(shape))
    sz, h = _lib.sz3(shape), _lib.h3((1.0, 1.0, 1.0))
    st = torch.cuda.Stream()
    out = {"shape": list(shape)}
    for dt, name in ((torch.float32, "f32"), (torch.float64, "f64")):
        s = 4 if dt == torch.float32 else 8
        code = _lib.F32 if dt == torch.float32 else _lib.F64
        with torch.cuda.stream(st):
            a = [torch.randn(N, dtype=dt, device="cuda") for _ in range(4)]
            ph = torch.remainder(4 * torch.randn(N, dtype=dt, device="cuda"), 2 * np.pi) - np.pi
            jumps = 9 * torch.randn(N, dtype=dt, device="cuda")
            G = torch.ones(N, dtype=torch.uint8, device="cuda")
        st.synchronize()
        sp = C.c_void_p(st.cuda_stream)
        ops = {
            "gradfp": (lambda: L.qsm_gradfp_dev(code, a[1].data_ptr(), a[2].data_ptr(), a[3].data_ptr(),
                                                a[0].data_ptr(), sz, h, sp), 4),
            "gradfp_adj": (lambda: L.qsm_gradfp_adj_dev(code, a[0].data_ptr(), a[1].data_ptr(),
                                                        a[2].data_ptr(), a[3].data_ptr(), sz, h, sp), 4),
            "gradf": (lambda: L.qsm_gradf_dev(code, a[1].data_ptr(), a[2].data_ptr(), a[3].data_ptr(),
                                              a[0].data_ptr(), sz, h, sp), 4),
            "gradf_adj": (lambda: L.qsm_gradf_adj_dev(code, a[0].data_ptr(), a[1].data_ptr(),
                                                      a[2].data_ptr(), a[3].data_ptr(), sz, h, sp), 4),
            "lapw": (lambda: L.qsm_lapw_dev(code, a[1].data_ptr(), ph.data_ptr(), sz, h, sp), 2),
            "lapw_jumps": (lambda: L.qsm_lapw_dev(code, a[1].data_ptr(), jumps.data_ptr(), sz, h, sp), 2),
            "gradc": (lambda: L.qsm_gradc_dev(code, a[1].data_ptr(), a[2].data_ptr(), a[3].data_ptr(),
                                              a[0].data_ptr(), sz, h, sp), 4),
            "lap1": (lambda: L.qsm_lap1_dev(code, a[1].data_ptr(), a[0].data_ptr(), G.data_ptr(),
                                            sz, h, sp), 2 + 1 / s),
        }
        res = {}
        for op, (fn, words) in ops.items():
            _lib.check(fn())
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(reps):
                _lib.check(fn())
            e1.record(st)
            e1.synchronize()
            t = e0.elapsed_time(e1) / 1e3 / reps
            res[op] = {"ms": t * 1e3, "gbs": words * N * s / t / 1e9}
        out[name] = res
        del a, ph, jumps, G
    torch.cuda.empty_cache()
    return out


def _timed_batch(vols, kinds, inputs, iters, warm):
    """Advances the volumes together (FP32, captured chunks of 10); returns
    (seconds for `iters` iterations after `warm`, resident bytes)."""
    from qsm_amd.batch import Batch

    with Batch(np.float32, hbm_budget=160 << 30, max_volumes=64) as b:
        for v, k, (d2f, m) in zip(vols, kinds, inputs):
            b.add(k, d2f, m, v["spec"].vsz, v["lam"], v["alpha"], maxiter=warm + iters)
        resident = b.used_bytes
        b.step(warm)
        b.sync()
        t0 = time.perf_counter()
        for _ in range(iters // 10):
            b.step(10)
        b.sync()
        dt = time.perf_counter() - t0
        done = [b.state(s)[2] for s in range(len(vols))]
    assert all(d == warm + iters for d in done), done
    return dt, resident


def batch_leg(iters: int, warm: int = 10):
    """C4: 48 uneven TV/TGV volumes advanced together (FP32, captured
    chunks of 10 iterations); compared with the same volumes solved one
    after another by the single-volume solver, for the whole mix and for
    each size class batched on its own."""
    import torch
    from qsm_amd import _lib, configs, pd, phantom

    vols = configs.c4_volumes()
    t = time.perf_counter()
    inputs = phantom.solver_inputs([v["spec"] for v in vols], dtype=np.float32)
    log(f"[bench] C4: {len(vols)} phantoms in {time.perf_counter() - t:.1f}s")
    kinds = [_lib.TGV if v["kind"] == "tgv" else _lib.TV for v in vols]
    run = 10 * (iters // 10)
    vox = sum(int(np.prod(v["spec"].shape)) for v in vols)
    dt, resident = _timed_batch(vols, kinds, inputs, run, warm)
    # the same volumes one after another (per size class and kind, tuned tiles)
    st = torch.cuda.Stream()
    per = {}
    for v, k, (d2f, m) in zip(vols, kinds, inputs):
        key = (v["spec"].shape, k)
        if key in per:
            continue
        with pd.Solver(k, np.float32, v["spec"].shape, v["spec"].vsz, v["lam"], v["alpha"],
                       stream=st) as s:
            s.autotune()
            s.load(d2f, m)
            s.iterate(warm)
            st.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            s.iterate(run)
            e1.record(st)
            e1.synchronize()
            per[key] = e0.elapsed_time(e1) / 1e3 / run
    seq = sum(per[(v["spec"].shape, k)] for v, k in zip(vols, kinds)) * run
    classes = {}
    for shape in configs.C4_SIZES:
        idx = [i for i, v in enumerate(vols) if v["spec"].shape == shape]
        if not idx:
            continue
        cv = [vols[i] for i in idx]
        ck = [kinds[i] for i in idx]
        ci = [inputs[i] for i in idx]
        cdt, _ = _timed_batch(cv, ck, ci, run, warm)
        cseq = sum(per[(v["spec"].shape, k)] for v, k in zip(cv, ck)) * run
        cvox = len(idx) * int(np.prod(shape))
        classes["x".join(map(str, shape))] = {
            "volumes": len(idx), "tgv": int(sum(k == _lib.TGV for k in ck)),
            "batch_voxel_iters_per_s": cvox * run / cdt,
            "sequential_voxel_iters_per_s": cvox * run / cseq,
            "speedup_vs_sequential": cseq / cdt,
        }
    return {
        "workload": ("C4 48 TV/TGV inversions, sizes uniform over 96^3, 128^3, 160x160x96, "
                     "224x224x128, 256x256x176, own phantom/mask/B0 each, FP32"),
        "volumes": len(vols), "voxels": vox, "iterations": run,
        "voxel_iters_per_s": vox * run / dt, "wall_s": dt,
        "sequential_voxel_iters_per_s": vox * run / seq,
        "speedup_vs_sequential": seq / dt,
        "resident_gib": resident / 2 ** 30,
        "per_size_class": classes,
    }


def _serve_once(specs, inputs, rate, max_live_voxels, chunk):
    from qsm_amd import _lib
    from qsm_amd.batch import Batch
    from qsm_amd.serve import Request, Server

    reqs = [Request(i, r["arrival"], _lib.TGV if r["kind"] == "tgv" else _lib.TV, d2f, m,
                    r["spec"].vsz, r["lam"], r["alpha"], r["maxiter"])
            for i, (r, (d2f, m)) in enumerate(zip(specs, inputs))]
    with Batch(np.float32, hbm_budget=160 << 30, max_volumes=64) as b:
        stats = Server(b, chunk=chunk, keep_chi=False, max_live_voxels=max_live_voxels).run(reqs)
    assert all(r.iters == s["maxiter"] for r, s in zip(reqs, specs))
    span = float(max(r.arrival for r in reqs))
    work = sum(int(np.prod(r.d2f.shape)) * r.maxiter for r in reqs)
    return {
        "requests": stats.n, "rate_per_s": rate, "arrival_span_s": span,
        "p50_latency_s": stats.p50, "p99_latency_s": stats.p99,
        "mean_wait_s": stats.mean_wait, "peak_live": stats.peak_live,
        "offered_voxel_iters_per_s": work / span if span > 0 else None,
        "voxel_iters_per_s": stats.voxel_iters_per_s, "wall_s": stats.wall,
    }


def serve_leg(n: int, rate: float, sustained_rate: float | None = None,
              max_live_voxels: int | None = 60_000_000, chunk: int = 20):
    """C5: Poisson arrivals served by continuous batching, each request with
    its own seeded phantom; latency from arrival to chi on the host.  Two
    rates over the same requests: `rate` (20/s, more than one GPU serves:
    the latency is mostly queueing) and a sustainable one that offers 0.8 of
    the throughput measured at `rate` (or `sustained_rate` if given)."""
    from qsm_amd import configs, phantom

    specs = configs.c5_requests(n=n, rate=rate)
    t = time.perf_counter()
    inputs = phantom.solver_inputs([r["spec"] for r in specs], dtype=np.float32)
    log(f"[bench] C5: {n} requests, own phantom each, prepared in "
        f"{time.perf_counter() - t:.1f}s")
    over = _serve_once(specs, inputs, rate, max_live_voxels, chunk)
    log(f"[bench] C5 at {rate}/s: p50 {over['p50_latency_s']:.2f}s p99 "
        f"{over['p99_latency_s']:.2f}s, served {over['voxel_iters_per_s']:.3e} of offered "
        f"{over['offered_voxel_iters_per_s']:.3e} voxel-iters/s")
    mean_work = sum(int(np.prod(r["spec"].shape)) * r["maxiter"] for r in specs) / n
    if sustained_rate is None:
        sustained_rate = 0.8 * over["voxel_iters_per_s"] / mean_work
    # the same requests (sizes, kinds, budgets and phantoms do not depend on
    # the rate: arrivals are scaled exponential draws)
    specs2 = configs.c5_requests(n=n, rate=sustained_rate)
    assert all(a["spec"] == b["spec"] and a["maxiter"] == b["maxiter"]
               for a, b in zip(specs, specs2))
    sus = _serve_once(specs2, inputs, sustained_rate, max_live_voxels, chunk)
    log(f"[bench] C5 at {sustained_rate:.2f}/s: p50 {sus['p50_latency_s']:.2f}s p99 "
        f"{sus['p99_latency_s']:.2f}s")
    return {
        "workload": ("C5 Poisson arrivals, TV/TGV 50/50, 70% 224x224x128, 20% 256x256x176, "
                     "10% 320x320x192, budgets 300-1000 iterations, FP32, own seeded phantom "
                     "per request"),
        "max_live_voxels": max_live_voxels, "chunk_iterations": chunk,
        "mean_voxel_iters_per_request": mean_work,
        "overload": over,
        "sustained": dict(sus, utilisation=(sus["offered_voxel_iters_per_s"]
                                            / over["voxel_iters_per_s"])),
        # the overload figures at the top level, as before
        "p50_latency_s": over["p50_latency_s"], "p99_latency_s": over["p99_latency_s"],
        "voxel_iters_per_s": over["voxel_iters_per_s"],
    }


def dipole_leg(shape=(256, 256, 176), reps=10):
    """FFT dipole convolution of a C3-sized volume (zero padding to
    nextfastfft(2*dims)), both precisions, timed with HIP events on its
    stream over `reps` calls after warm-up (plans cached, inputs resident).

    hbm_bytes: what the five passes move (the input read and the output
    written once; the half spectrum along x, H = M0/2 + 1 columns, read and
    written by each pass over the rows and planes that hold data) -- the
    FETCH_SIZE/WRITE_SIZE counters agree with it.  complex_equiv_bytes: the
    bytes of two full complex 3D transforms of the padded volume, each axis
    pass reading and writing it once (2 x 3 x 2*c*M), which the real-input
    half spectrum and the
    unmaterialised padding avoid; its rate can exceed HBM peak."""
    import torch
    from qsm_amd import dipole

    pad = tuple(n // 2 for n in shape)
    M, _ = dipole.pad_size(shape, pad)
    Mt = int(np.prod(M))
    N = int(np.prod(shape))
    out = {"shape": list(shape), "padded": list(M)}
    st = torch.cuda.Stream()
    for dt, name in ((torch.float32, "f32"), (torch.float64, "f64")):
        s = 4 if dt == torch.float32 else 8
        c = 2 * s
        with torch.cuda.stream(st):
            chi = torch.randn(shape[::-1], dtype=dt, device="cuda").permute(2, 1, 0)
            fld = torch.empty_like(chi)
            work = torch.empty(int(dipole.work_bytes(dt, shape, pad)), dtype=torch.uint8,
                               device="cuda")
        st.synchronize()
        args = ((1.0, 1.0, 1.0), (0.0, 0.0, 1.0), pad, fld, work, st)
        dipole.dipole_conv_dev(chi, *args)
        st.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            dipole.dipole_conv_dev(chi, *args)
        e1.record(st)
        e1.synchronize()
        t = e0.elapsed_time(e1) / 1e3 / reps
        H = M[0] // 2 + 1
        rows = H * shape[1] * shape[2]      # half-spectrum rows that hold data
        planes = H * M[1] * shape[2]        # data planes, all rows
        hbm = 2 * N * s + 4 * rows * c + 4 * planes * c
        alg = 12 * c * Mt                    # 2 transforms x 3 axis passes x 2*c*M
        out[name] = {"ms": t * 1e3, "hbm_bytes": hbm, "hbm_gbs": hbm / t / 1e9,
                     "complex_equiv_bytes": alg, "complex_equiv_gbs": alg / t / 1e9}
        del chi, fld, work
    torch.cuda.empty_cache()
    return out


def admm_leg(cfg, field, mask, iters=20, cpu_iters=2):
    """tv.m (TV inversion by ADMM) and rts.m (LSMR in k-space, then the same
    ADMM) at the C3 size: the phantom's local field and mask, both
    precisions, `iters` ADMM iterations (the default maxit) with the
    stopping tolerances at 0 so that every run does them all; one warm-up
    solve, then one solve timed with HIP events on its stream (the loop is
    decided on the device: one synchronisation per solve).  default_opts:
    one more timed solve with the method's default tolerances, which stop
    the loop early; the launches enqueued after the stop return at once.

    hbm_bytes per ADMM iteration: right-hand side 7 N s (v, l read with -1
    neighbours, r written), the half-spectrum solve 2 N s + 9 S (S = H ny nz
    c, the spectrum; passes: x 1+1, y 2, z 3 with F, y 2, x 1+1), the v/l
    update with both residuals' sums and the next right-hand side at the
    same rho 14 N s (u, v, l read; the other v, l set and r written; the -1
    neighbours' terms recomputed, not stored).  The right-hand side runs
    (7 N s) only in the first iteration and after those whose rule changed
    rho, counted from the returned trace.
    rts adds its step 1 once (4 LSMR iterations and two full complex 3D
    FFTs).  cpu: the oracle (numpy restatement, one thread for its FFTs) on
    cpu_iters ADMM iterations of the same input, setup included."""
    import torch
    from qsm_amd import rts as grts
    from qsm_amd import tv as gtv

    shape = field.shape
    N = int(np.prod(shape))
    vsz, bdir = cfg["spec"].vsz, cfg["spec"].bdir
    opts = {"maxit": iters, "rtol": 0.0, "atol": 0.0, "verbose": 0}
    out = {"shape": list(shape), "iterations": iters,
           "opts": "tv.m / rts.m defaults, rtol = atol = 0"}
    st = torch.cuda.Stream()
    for kind, mod, run in (("tv", gtv, gtv.tv_dev), ("rts", grts, grts.rts_dev)):
        res = {}
        for dt, name in ((torch.float32, "f32"), (torch.float64, "f64")):
            s = 4 if dt == torch.float32 else 8
            with torch.cuda.stream(st):
                f = torch.from_numpy(np.ascontiguousarray(field.reshape(-1, order="F")))
                f = f.to("cuda", dt).view(shape[::-1]).permute(2, 1, 0)
                m = torch.from_numpy(np.ascontiguousarray(mask.reshape(-1, order="F"))).to("cuda")
                m = m.view(shape[::-1]).permute(2, 1, 0)
                x = torch.empty_like(f)
                work = torch.empty(mod.work_bytes(dt, shape), dtype=torch.uint8, device="cuda")
            st.synchronize()
            run(f, m, vsz, bdir, opts, x, work=work, stream=st)  # warm-up (plans)
            st.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            tr = np.zeros(5 * iters, dtype=np.float64)
            e0.record(st)
            it = run(f, m, vsz, bdir, opts, x, work=work, stream=st, trace=tr)
            e1.record(st)
            e1.synchronize()
            t = e0.elapsed_time(e1) / 1e3
            H = shape[0] // 2 + 1
            S = H * shape[1] * shape[2] * 2 * s
            rho = np.concatenate(([mod.make_opts(opts).rho], tr[4::5][:it]))
            n_rhs = 1 + int(np.count_nonzero(rho[1:it] != rho[:it - 1]))
            hbm = ((2 + 14) * N * s + 9 * S) * it + 7 * N * s * n_rhs
            res[name] = {"ms_total": t * 1e3, "ms_per_iter": t / it * 1e3,
                         "voxel_iters_per_s": N * it / t, "rhs_launches_run": n_rhs,
                         "hbm_bytes_per_iter": hbm / it,
                         "hbm_gbs_per_iter": hbm / t / 1e9}
            # the defaults' tolerances: the rule stops the loop early and the
            # remaining maxit - iterations launches return at once
            e0.record(st)
            itd = run(f, m, vsz, bdir, {"verbose": 0}, x, work=work, stream=st)
            e1.record(st)
            e1.synchronize()
            res[name]["default_opts"] = {"ms_total": e0.elapsed_time(e1),
                                         "iterations": itd,
                                         "maxit": mod.make_opts(None).maxit}
            del f, m, x, work
        torch.cuda.empty_cache()
        if cpu_iters > 0:
            from oracle import tv_admm as tva

            o = dict(opts, maxit=cpu_iters)
            t = time.perf_counter()
            (tva.tv if kind == "tv" else tva.rts)(field.astype(np.float32), mask, vsz, bdir, o)
            c = time.perf_counter() - t
            res["cpu"] = {"value": N * cpu_iters / c, "unit": "voxel-iters/s", "cores": 1,
                          "kind": "port",
                          "sample": (f"oracle {kind} (numpy restatement of {kind}.m, float32, "
                                     f"numpy FFT), {shape[0]}x{shape[1]}x{shape[2]}, "
                                     f"{cpu_iters} ADMM iterations incl. setup, {c:.2f}s")}
        out[kind] = res
    return out


def bgremove_leg(cfg, field, mask, reps=3, cpu=True):
    """SHARP (sharp.m, r = 9) and V-SHARP (vsharp.m, radii 9:-2:2) at the C3
    size on the phantom's field and mask, both precisions.  device_ms: the
    device form on the cropped volume resident in HBM, HIP events on its
    stream; host_ms: the drop-in's form (host arrays in and out: crop,
    copies, device convolutions, uncrop) timed on the host; both over
    `reps` calls after a warm-up.  cpu: the oracle (numpy restatement, one
    thread for its FFTs) on the same input, FP64."""
    import torch
    from qsm_amd import bgremove as gbg

    vsz = cfg["spec"].vsz
    rs = gbg.default_radii(vsz)
    out = {"shape": list(field.shape), "mask_voxels": int(np.count_nonzero(mask)),
           "vsharp_radii": [float(r) for r in rs]}
    ix, iy, iz = (np.nonzero(np.asarray(mask).any(axis=a))[0] for a in ((1, 2), (0, 2), (0, 1)))
    crop = np.ix_(ix, iy, iz)
    st = torch.cuda.Stream()
    for name, fn, dev in (("sharp", lambda f: gbg.sharp(f, mask, vsz, 9),
                           lambda fc, mc, o, mo, w: gbg.sharp_dev(fc, mc, vsz, 9, 0.05, o, mo, w, st)),
                          ("vsharp", lambda f: gbg.vsharp(f, mask, vsz),
                           lambda fc, mc, o, mo, w: gbg.vsharp_dev(fc, mc, vsz, rs, 0.05, o, mo, w, st))):
        res = {}
        for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
            f = np.asfortranarray(field.astype(dt))
            fn(f)
            t = time.perf_counter()
            for _ in range(reps):
                fn(f)
            host_ms = (time.perf_counter() - t) / reps * 1e3
            fcrop = np.asfortranarray(f[crop])
            shp = fcrop.shape
            with torch.cuda.stream(st):
                fc = torch.from_numpy(np.ascontiguousarray(fcrop.reshape(-1, order="F"))).to("cuda")
                fc = fc.view(shp[::-1]).permute(2, 1, 0)
                mcp = np.ascontiguousarray(np.asarray(mask)[crop].astype(np.uint8).reshape(-1, order="F"))
                mc = torch.from_numpy(mcp).to("cuda").view(shp[::-1]).permute(2, 1, 0)
                o = torch.empty_like(fc)
                mo = torch.empty_like(mc)
            w = dev(fc, mc, o, mo, None)
            st.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(reps):
                dev(fc, mc, o, mo, w)
            e1.record(st)
            e1.synchronize()
            res[tag] = {"device_ms": e0.elapsed_time(e1) / reps, "host_ms": host_ms}
            del fc, mc, o, mo, w
        torch.cuda.empty_cache()
        if cpu:
            from oracle import bgremove as obg

            t = time.perf_counter()
            (obg.sharp(field, mask, vsz, 9) if name == "sharp" else obg.vsharp(field, mask, vsz))
            res["cpu"] = {"ms": (time.perf_counter() - t) * 1e3, "cores": 1, "kind": "port",
                          "sample": f"oracle {name} (numpy restatement), float64, same input"}
        out[name] = res
    return out


def _wrapped_phase(field, mask):
    """A wrapped phase on the volume: the field scaled to 8 rad peak inside
    the mask plus a smooth quadratic background of about 40 rad across the
    volume, wrapped to (-pi, pi]."""
    shp = field.shape
    g = np.meshgrid(*(np.linspace(-1.0, 1.0, n) for n in shp), indexing="ij")
    m = np.asarray(mask) != 0
    peak = float(np.abs(field[m]).max()) if m.any() else 1.0
    phi = 8.0 * field / (peak if peak > 0 else 1.0) + 20.0 * (g[0] ** 2 + 0.5 * g[1] * g[2] - 0.3 * g[2])
    return np.asfortranarray(np.angle(np.exp(1j * phi)))


def unwrap_leg(cfg, field, mask, reps=3, cpu=True):
    """unwrapLaplacian.m at the C3 size on a wrapped phase made from the
    phantom's field (_wrapped_phase) with the phantom's mask, both
    precisions.  device_ms: the device form (lapw + MGPCG) on the crop (the
    mask's box grown by one voxel) resident in HBM, HIP events on its
    stream; host_ms: the drop-in's form (host arrays in and out); iterations:
    PCG iterations of the device form.  cpu: the oracle's numpy port of the
    reference's algorithm (mgpcg_mex.c's MGPCG, FP64, unwrapLaplacian's
    options) for a fixed 4 PCG iterations on the centre 128^3 block with
    the mask restricted to it, per iteration, beside the device form's time
    per iteration on that same block.  A fixed count: on a mask that
    touches the volume's faces (the phantom's covers the whole volume) the
    reference's red-black loops, which skip a row end of the second colour,
    make its V-cycle a poor preconditioner and PCG would run for hundreds of
    iterations; the port's residual after the 4 is reported."""
    import torch
    from qsm_amd import unwrap as uw

    vsz = cfg["spec"].vsz
    t = time.perf_counter()
    phase = _wrapped_phase(field, mask)
    m = np.asarray(mask) != 0
    log(f"[bench] unwrap: wrapped phase ready in {time.perf_counter() - t:.1f}s")
    out = {"shape": list(field.shape), "mask_voxels": int(np.count_nonzero(m))}
    nz = np.argwhere(m)
    lo = np.maximum(nz.min(axis=0) - 1, 0)
    hi = np.minimum(nz.max(axis=0) + 1, np.array(m.shape) - 1)
    crop = tuple(slice(a, b + 1) for a, b in zip(lo, hi))
    st = torch.cuda.Stream()

    def on_device(u, mk, dt):
        shp = u.shape
        with torch.cuda.stream(st):
            tu = torch.from_numpy(np.ascontiguousarray(u.astype(dt).reshape(-1, order="F"))).to("cuda")
            tu = tu.view(shp[::-1]).permute(2, 1, 0)
            tm = torch.from_numpy(np.ascontiguousarray(mk.astype(np.uint8).reshape(-1, order="F")))
            tm = tm.to("cuda").view(shp[::-1]).permute(2, 1, 0)
            up, d2 = torch.empty_like(tu), torch.empty_like(tu)
            w = torch.empty(uw.work_bytes(tu.dtype, shp), dtype=torch.uint8, device="cuda")
        it = uw.unwrap_laplacian_dev(tu, tm, vsz, up, d2, w, st)
        st.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            uw.unwrap_laplacian_dev(tu, tm, vsz, up, d2, w, st)
        e1.record(st)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del tu, tm, up, d2, w
        return ms, it

    for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
        p = np.asfortranarray(phase.astype(dt))
        t = time.perf_counter()
        uw.unwrap_laplacian(p, m, vsz)
        log(f"[bench] unwrap {tag}: host form warm-up {time.perf_counter() - t:.2f}s")
        t = time.perf_counter()
        for _ in range(reps):
            uw.unwrap_laplacian(p, m, vsz)
        host_ms = (time.perf_counter() - t) / reps * 1e3
        dev_ms, it = on_device(np.asfortranarray(p[crop]), m[crop], dt)
        log(f"[bench] unwrap {tag}: host form {host_ms:.1f} ms, device form {dev_ms:.1f} ms, "
            f"{it} PCG iterations")
        out[tag] = {"device_ms": dev_ms, "host_ms": host_ms, "iterations": it,
                    "crop": [int(b - a + 1) for a, b in zip(lo, hi)]}
    torch.cuda.empty_cache()
    if cpu:
        from oracle import poisson as opo

        c = [n // 2 for n in field.shape]
        blk = tuple(slice(x - 64, x + 64) for x in c)
        pb, mb = np.asfortranarray(phase[blk]), m[blk]
        from oracle import oracle as orc

        rhs = np.asfortranarray(-(mb * orc.lapw(pb, vsz)))
        t = time.perf_counter()
        _, cpu_it, hist = opo.mgpcg_port(rhs, mb, vsz, 1e-8, 4,
                                         {"maxit": 2, "npre": 2, "npost": 2, "nboundary": 2})
        cpu_ms = (time.perf_counter() - t) * 1e3
        log(f"[bench] unwrap cpu port (128^3 block): {cpu_ms / 1e3:.1f}s for {cpu_it} PCG iterations")
        dev_ms, dev_it = on_device(pb, mb, np.float64)
        fn = float(np.linalg.norm(rhs))
        out["cpu"] = {"ms_per_iteration": cpu_ms / max(cpu_it, 1), "iterations_run": cpu_it,
                      "rel_residual_after": float(hist[-1] / fn) if cpu_it and fn > 0 else 0.0,
                      "cores": 1, "kind": "port",
                      "device_ms_same_block_f64": dev_ms, "device_iterations_same_block": dev_it,
                      "device_ms_per_iteration_same_block": dev_ms / max(dev_it, 1),
                      "sample": "oracle port of mgpcg_mex.c (numpy, float64), unwrapLaplacian's "
                                "options, 4 PCG iterations on the centre 128^3 block of the same "
                                "phase and mask"}
    return out


def nifti_leg(field, mask, vsz, reps=3):
    """saveNii.m / loadNii.m at the C3 size (host IO, the request payloads):
    the field as float64 .nii and float32 .nii.gz, the mask as uint8 .nii,
    written and read back in a temporary folder; MB/s of image data."""
    import shutil
    import tempfile

    from qsm_amd.nifti import loadNii, saveNii

    out = {}
    d = tempfile.mkdtemp(prefix="qsm_nii_")
    try:
        for tag, arr, name in (("f64_nii", field.astype(np.float64), "f.nii"),
                               ("f32_nii_gz", field.astype(np.float32), "f.nii.gz"),
                               ("mask_u8_nii", (np.asarray(mask) != 0).astype(np.uint8), "m.nii")):
            fn = os.path.join(d, name)
            t = time.perf_counter()
            for _ in range(reps):
                saveNii(fn, arr, vsz)
            w = (time.perf_counter() - t) / reps
            t = time.perf_counter()
            for _ in range(reps):
                img = loadNii(fn)["img"]
            r = (time.perf_counter() - t) / reps
            assert img.shape == arr.shape
            mb = arr.nbytes / 1e6
            out[tag] = {"mb": mb, "file_mb": os.path.getsize(fn) / 1e6, "save_ms": w * 1e3,
                        "load_ms": r * 1e3, "save_mb_s": mb / w, "load_mb_s": mb / r}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return out


def _schedule_text(r):
    tx, ty, kz = r["tiles"]
    l = r["launches"]
    if l["pass2"]:
        return (f"two-iteration fused pass (primal n, dual n+1, primal n+1, dual n+2 per "
                f"launch), z-chunks of {r['kz2']}; {l['pass2']} two-iteration launches, "
                f"{l['pass1']} one-iteration, {l['checks']} convergence checks in the timed "
                f"region (autotuned)")
    return (f"one-iteration fused pass, {tx}x{ty} tiles, z-chunks of {kz}; {l['pass1']} "
            f"launches, {l['checks']} convergence checks in the timed region (autotuned)")


def roofline(r, N, n_active, s):
    """Roofline of the dominant kernel of the timed region.

    Algorithmic bytes per voxel-iteration: 41 words of state + 1 mask byte
    for TGV, the reference's minimal single-pass traffic with the
    extrapolated chi stored (pd_tgv_mex.c:331-977).  A two-iteration launch
    processes two voxel-iterations per voxel, so its algorithmic bytes are
    twice one iteration's.  The launch duration is the HIP-event time of the
    timed region on the solver's stream over the dominant kernel's launch
    count (the convergence-check kernels, about 3 us each every tenth
    iteration, are inside it).  `traffic` is the HBM bytes one launch of that
    kernel moves, from the committed FETCH_SIZE/WRITE_SIZE summary."""
    l = r["launches"]
    two = l["pass2"] > 0
    per_it = n_active * 41 * s + N
    if two:
        # iterations of the region in two-iteration launches; a trailing
        # single iteration (odd step count) is priced at the same time per
        # iteration
        its, nl, kern = 2, l["pass2"], "fused2_kernel"
        share = 2 * l["pass2"] / (2 * l["pass2"] + l["pass1"])
    else:
        its, nl, kern = 1, max(1, l["pass1"] + l["sweeps"]), "fused_kernel"
        share = 1.0
    t_launch = r["gpu_s"] * share / nl
    alg = its * per_it
    ach = alg / t_launch / 1e9
    # the bytes the kernel itself cannot avoid: its 31-word read+write of the
    # state (16 words in, 15 out) and the mask, once per launch
    own = n_active * 31 * s + N
    out = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": ach / HBM_PEAK_GBS, "traffic": None, "kernel": kern,
           "iterations_per_launch": its, "launch_ms": t_launch * 1e3,
           "algorithmic_bytes_per_launch": alg,
           "own_minimum_bytes_per_launch": own,
           "own_minimum_gbs": own / t_launch / 1e9}
    key = "f32" if s == 4 else "f64"
    if TRAFFIC_FILE.exists():
        try:
            t = json.loads(TRAFFIC_FILE.read_text()).get(key, {})
            if t.get("iterations_per_launch") == its:
                out["traffic"] = t["bytes_per_launch"]
                out["traffic_gbs"] = t["bytes_per_launch"] / t_launch / 1e9
                out["traffic_frac"] = out["traffic_gbs"] / HBM_PEAK_GBS
        except Exception:
            pass
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[1])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--cpu-iters", type=int, default=20,
                    help="iterations of the CPU baseline sample (0: skip)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--accuracy-iters", type=int, default=1000,
                    help="iterations of the FP32-vs-FP64 chi accuracy leg (0: skip)")
    ap.add_argument("--batch-iters", type=int, default=100,
                    help="iterations of the C4 batch leg (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=200,
                    help="requests of the C5 serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-rate-sustained", type=float, default=None,
                    help="rate of the second serving leg (default: 0.8 of what the GPU "
                         "served at --serve-rate)")
    ap.add_argument("--no-bgremove", dest="bgremove", action="store_false",
                    help="skip the SHARP / V-SHARP leg")
    ap.add_argument("--no-brain", dest="brain", action="store_false",
                    help="skip the C3 brain-mask variant")
    ap.add_argument("--no-nifti", dest="nifti", action="store_false",
                    help="skip the NIfTI IO leg (saveNii / loadNii at C3)")
    ap.add_argument("--no-unwrap", dest="unwrap", action="store_false",
                    help="skip the unwrapLaplacian (lapw + MGPCG) leg")
    ap.add_argument("--admm-iters", type=int, default=20,
                    help="ADMM iterations of the tv.m / rts.m leg (0: skip)")
    ap.add_argument("--no-dipole", dest="dipole", action="store_false",
                    help="skip the finite-difference and FFT dipole convolution legs")
    a = ap.parse_args(argv)

    if a.gpus != 1:
        print(f"bench.py measures one GPU only (the inversion of one volume runs on "
              f"one MI355X); --gpus {a.gpus} is not supported", file=sys.stderr)
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs as a single process", file=sys.stderr)
        return 2

    import torch
    from qsm_amd import _lib

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    _lib.lib()

    cfg, d2f, mask, field = make_input(a.workload)
    shape = d2f.shape
    N = int(np.prod(shape))
    n_active = int(np.count_nonzero(mask))

    r32 = time_solver(_lib.TGV, np.float32, cfg, d2f, mask, a.steps, a.warmup)
    log(f"[bench] fp32: {r32['gpu_s'] / a.steps * 1e3:.4f} ms/iter")
    r64 = None
    if not a.no_fp64:
        r64 = time_solver(_lib.TGV, np.float64, cfg, d2f, mask, a.steps, a.warmup)
        log(f"[bench] fp64: {r64['gpu_s'] / a.steps * 1e3:.4f} ms/iter")

    v32 = N * a.steps / r32["gpu_s"]
    out = {
        "metric": METRIC,
        "value": v32,
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r32["wall_s"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 60-ellipsoid phantom, dipole field + noise, -lapw)",
        "config": {
            "workload": (f"{a.workload.upper()} TGV primal-dual inversion "
                         f"{shape[0]}x{shape[1]}x{shape[2]}, full-box mask, "
                         f"alpha1={cfg['lam']}, alpha0={cfg['alpha']}*alpha1, tol=0"),
            "voxels": N,
            "n_active": n_active,
            "schedule": _schedule_text(r32),
            "settle": ("before the warmup the same solve runs {} untimed iterations, so "
                       "the timed region sees the GPU's steady clocks and starts on a "
                       "two-iteration boundary").format(r32["settle_iterations"]),
        },
    }
    out["roofline"] = roofline(r32, N, n_active, 4)
    ach = out["roofline"]["achieved"]
    rf = out["roofline"]
    log(f"[bench] roofline FP32 {rf['kernel']}: {rf['launch_ms']:.4f} ms/launch, "
        f"{rf['iterations_per_launch']} iterations/launch, algorithmic {ach:.0f} GB/s "
        f"(frac {rf['frac']:.3f}), own minimum {rf['own_minimum_gbs']:.0f} GB/s"
        + (f", HBM traffic {rf['traffic_gbs']:.0f} GB/s" if rf.get("traffic_gbs") else ""))
    if r64 is not None:
        v64 = N * a.steps / r64["gpu_s"]
        c32 = r32["chi"].astype(np.float64)
        c64 = r64["chi"]
        n64 = float(np.linalg.norm(c64))
        out["fp64"] = {
            "value": v64,
            "ms_per_step": r64["wall_s"] / a.steps * 1e3,
            "roofline": roofline(r64, N, n_active, 8),
            "schedule": _schedule_text(r64),
        }
        out["fp32_over_fp64"] = v32 / v64
        out["chi_err_fp32_vs_fp64"] = {
            "rel_l2": float(np.linalg.norm(c32 - c64) / n64) if n64 > 0 else 0.0,
            "max_abs": float(np.max(np.abs(c32 - c64))),
            "iterations": r32["iterations"],
        }
    try:
        out["copy_peak_gbs"] = copy_peak_gbs()
        out["roofline"]["frac_of_copy_peak"] = ach / out["copy_peak_gbs"]
    except Exception as e:  # measurement aid only
        log(f"[bench] copy peak failed: {e}")
    try:
        out["stream15_ceiling_gbs"] = stream_ceiling_gbs()
        if out["roofline"].get("traffic_gbs"):
            out["roofline"]["traffic_frac_of_stream15"] = (out["roofline"]["traffic_gbs"]
                                                           / out["stream15_ceiling_gbs"])
    except Exception as e:  # measurement aid only
        log(f"[bench] stream ceiling failed: {e}")
    if a.dipole:
        out["fd"] = fd_leg()
        log("[bench] fd f32: " + ", ".join(f"{k} {v['gbs']:.0f} GB/s" for k, v in out["fd"]["f32"].items()))
        out["dipole_conv"] = dipole_leg()
        log(f"[bench] dipole conv f32 {out['dipole_conv']['f32']['ms']:.3f} ms, "
            f"f64 {out['dipole_conv']['f64']['ms']:.3f} ms")
    if a.admm_iters > 0:
        out["admm"] = admm_leg(cfg, field, mask, a.admm_iters, 2 if a.cpu_iters > 0 else 0)
        for k in ("tv", "rts"):
            r32, r64 = out['admm'][k]['f32'], out['admm'][k]['f64']
            log(f"[bench] {k}.m (ADMM) f32 {r32['ms_total']:.2f} ms, "
                f"f64 {r64['ms_total']:.2f} ms for {a.admm_iters} iterations; default "
                f"tolerances f32 {r32['default_opts']['ms_total']:.2f} ms "
                f"({r32['default_opts']['iterations']} of {r32['default_opts']['maxit']}), "
                f"f64 {r64['default_opts']['ms_total']:.2f} ms "
                f"({r64['default_opts']['iterations']} of {r64['default_opts']['maxit']})")
    if a.bgremove:
        out["bgremove"] = bgremove_leg(cfg, field, mask, cpu=a.cpu_iters > 0)
        log(f"[bench] sharp f32 {out['bgremove']['sharp']['f32']['device_ms']:.2f} ms on device, "
            f"vsharp f32 {out['bgremove']['vsharp']['f32']['device_ms']:.2f} ms")
    if a.brain:
        out["c3_brain"] = brain_leg()
        log(f"[bench] C3 brain mask ({out['c3_brain']['active_fraction']:.0%} active): FP32 TGV "
            f"{out['c3_brain']['ms_per_iter']:.3f} ms/iter, "
            f"{out['c3_brain']['algorithmic_gbs']:.0f} GB/s algorithmic")
    if a.unwrap:
        out["unwrap"] = unwrap_leg(cfg, field, mask, cpu=a.cpu_iters > 0)
        log(f"[bench] unwrapLaplacian f32 {out['unwrap']['f32']['device_ms']:.2f} ms on device "
            f"({out['unwrap']['f32']['iterations']} PCG iterations), "
            f"f64 {out['unwrap']['f64']['device_ms']:.2f} ms")
    if a.nifti:
        out["nifti"] = nifti_leg(field, mask, cfg["spec"].vsz)
        log("[bench] NIfTI IO: " + ", ".join(f"{k} load {v['load_mb_s']:.0f} MB/s save "
                                             f"{v['save_mb_s']:.0f} MB/s" for k, v in out["nifti"].items()))
    if a.batch_iters > 0:
        out["batch"] = batch_leg(a.batch_iters)
        log(f"[bench] C4 batch: {out['batch']['voxel_iters_per_s']:.3e} voxel-iters/s")
    if a.serve_requests > 0:
        out["serving"] = serve_leg(a.serve_requests, a.serve_rate, a.serve_rate_sustained)
    if a.accuracy_iters > 0:
        out["accuracy"] = accuracy_leg(cfg, d2f, mask, a.accuracy_iters)
        log(f"[bench] FP32 vs FP64 chi after {a.accuracy_iters} iterations: TGV rel L2 "
            f"{out['accuracy']['tgv']['rel_l2']:.2e}, TV {out['accuracy']['tv']['rel_l2']:.2e}")
    if a.cpu_iters > 0:
        log(f"[bench] CPU baseline: {a.cpu_iters} iterations of the oracle ...")
        cb = cpu_baseline(cfg, d2f, mask, a.cpu_iters)
        cpu_chi, cpu_it = cb.pop("_chi"), cb.pop("_iters")
        out["cpu_baseline"] = cb
        out["fp64_gpu_vs_cpu"] = fp64_agreement(cfg, d2f, mask, cpu_chi, cpu_it)
        log(f"[bench] FP64 GPU vs CPU chi after {cpu_it} iterations: rel L2 "
            f"{out['fp64_gpu_vs_cpu']['rel_l2']:.2e}, bitwise "
            f"{out['fp64_gpu_vs_cpu']['bitwise_equal']}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
