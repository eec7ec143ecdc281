"""The benchmark workloads themselves on the GPU, at their full sizes.

* C2 (224x224x128 TGV, 0.9x0.9x1.2 mm, ellipsoidal mask): all 1000
  iterations of the loop (pd_tgv_mex.c:296-320), FP64 chi bitwise the
  oracle's.
* C3 (256x256x176, full-box mask): FP64 TGV and TV bitwise the oracle's
  after 20 iterations (two convergence checks); the precision variant, FP32
  against FP64 after the config's 1000 iterations, relative L2 <= 1e-4
  (the bar on chi) for both kinds.
* C4: all 48 volumes advanced together in one batch for 20 iterations;
  every volume bitwise equal to its solo solve, and one volume of each size
  class bitwise the oracle's.
* C5: eight requests of the stream (one of them 320x320x192) served by
  continuous batching at their full iteration budgets; every result bitwise
  equal to the request solved alone, and the shortest one the oracle's.
* A bank above 2 GiB: FP64 TGV at 320x320x192 (2.36 GB per bank, byte
  offsets past 2^31) through the two- and one-iteration passes, the
  two-sweep schedule, both batch marches and the oracle, bitwise.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from qsm_amd import _lib, configs, pd, phantom
from qsm_amd.batch import Batch
from qsm_amd.serve import Request, Server

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[workloads]", *a, flush=True)


def _kind(name):
    return _lib.TGV if name == "tgv" else _lib.TV


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _solo(kind, dt, shape, vsz, lam, alpha, d2f, m, iters):
    with pd.Solver(kind, dt, shape, vsz, lam, alpha) as s:
        s.load(np.asarray(d2f, dtype=dt, order="F"), m)
        s.iterate(iters)
        return s.chi()


def _oracle(kind, d2f, m, vsz, lam, alpha, iters):
    if kind == _lib.TGV:
        return orc.pd_tgv(d2f, m, vsz, lam, alpha, 0.0, iters)[0]
    return orc.pd_tv(d2f, m, vsz, lam, 0.0, iters)[0]


@pytest.fixture(scope="module")
def c3_input(cuda):
    _say("C3 phantom")
    return phantom.solver_input(configs.CONFIGS["c3"]["spec"])


@pytest.mark.timeout(600)
def test_c2_tgv_full_1000_iterations(cuda):
    cfg = configs.CONFIGS["c2"]
    spec = cfg["spec"]
    d2f, m = phantom.solver_input(spec)
    _say("C2 oracle, 1000 iterations")
    ref, it, _ = orc.pd_tgv(d2f, m, spec.vsz, cfg["lam"], cfg["alpha"], 0.0, cfg["iters"])
    assert it == 1000
    _say("C2 GPU drop-in, 1000 iterations")
    got = pd.pd_tgv_mex(d2f, m, spec.vsz, cfg["lam"], cfg["alpha"], 0.0, cfg["iters"], 0)
    assert _rel(got, ref) <= 1e-10
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("kind", ["tgv", "tv"])
def test_c3_fp64_matches_oracle(c3_input, kind):
    cfg = configs.CONFIGS["c3"]
    d2f, m = c3_input
    vsz, lam, alpha = cfg["spec"].vsz, cfg["lam"], cfg["alpha"] if kind == "tgv" else 0.0
    _say(f"C3 {kind} FP64 oracle, 20 iterations")
    ref = _oracle(_kind(kind), d2f, m, vsz, lam, alpha, 20)
    _say(f"C3 {kind} FP64 GPU, 20 iterations")
    got = _solo(_kind(kind), np.float64, d2f.shape, vsz, lam, alpha, d2f, m, 20)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("kind", ["tgv", "tv"])
def test_c3_fp32_within_1e4_of_fp64_after_1000(c3_input, kind):
    """The precision variant's bar: FP32 chi within 1e-4 relative L2 of the
    FP64 chi (bitwise the oracle's, test above) after 1000 iterations."""
    cfg = configs.CONFIGS["c3"]
    d2f, m = c3_input
    vsz, lam, alpha = cfg["spec"].vsz, cfg["lam"], cfg["alpha"] if kind == "tgv" else 0.0
    _say(f"C3 {kind} FP64 and FP32, 1000 iterations")
    c64 = _solo(_kind(kind), np.float64, d2f.shape, vsz, lam, alpha, d2f, m, 1000)
    c32 = _solo(_kind(kind), np.float32, d2f.shape, vsz, lam, alpha, d2f, m, 1000)
    err = _rel(c32, c64)
    _say(f"C3 {kind} FP32 vs FP64 rel L2 {err:.3e}, max abs "
         f"{np.abs(c32.astype(np.float64) - c64).max():.3e}")
    assert np.all(np.isfinite(c32))
    assert err <= 1e-4


@pytest.mark.timeout(600)
def test_c4_all_48_volumes_batched(cuda):
    vols = configs.c4_volumes()
    assert len(vols) == 48
    iters = 20
    _say("C4 phantoms")
    inputs = phantom.solver_inputs([v["spec"] for v in vols])
    kinds = [_kind(v["kind"]) for v in vols]
    _say("C4 batch of 48, 20 iterations")
    with Batch(np.float32, hbm_budget=160 << 30, max_volumes=64) as b:
        slots = [b.add(k, d2f, m, v["spec"].vsz, v["lam"], v["alpha"], maxiter=iters)
                 for v, k, (d2f, m) in zip(vols, kinds, inputs)]
        b.step(10)
        b.step(10)
        b.sync()
        res = [b.result(s) for s in slots]
    assert all(r.iters == iters for r in res)
    _say("C4 solo solves")
    first = {}
    for i, (v, k, (d2f, m), r) in enumerate(zip(vols, kinds, inputs, res)):
        solo = _solo(k, np.float32, v["spec"].shape, v["spec"].vsz, v["lam"], v["alpha"],
                     d2f, m, iters)
        np.testing.assert_array_equal(r.chi, solo, err_msg=f"volume {i} {v['spec'].shape}")
        first.setdefault(v["spec"].shape, i)
    assert len(first) == len(configs.C4_SIZES)
    for shape, i in sorted(first.items()):
        v, k, (d2f, m) = vols[i], kinds[i], inputs[i]
        _say(f"C4 oracle, volume {i} {shape} {v['kind']}")
        ref = _oracle(k, d2f.astype(np.float32, order="F"), m, v["spec"].vsz, v["lam"],
                      v["alpha"], iters)
        np.testing.assert_array_equal(res[i].chi, ref, err_msg=f"volume {i} {shape}")


@pytest.mark.timeout(600)
def test_c5_requests_equal_solo_solves(cuda):
    specs = configs.c5_requests(n=40)
    big = next(i for i, r in enumerate(specs) if r["spec"].shape == (320, 320, 192))
    pick = sorted(set(range(7)) | {big})
    assert len(pick) == 8
    reqs = []
    _say("C5 phantoms")
    for j, i in enumerate(pick):
        r = specs[i]
        d2f, m = phantom.solver_input(r["spec"])
        # arrivals kept relative to each other, compressed so the test is short
        reqs.append(Request(j, 0.05 * j, _kind(r["kind"]), d2f.astype(np.float32, order="F"), m,
                            r["spec"].vsz, r["lam"], r["alpha"], r["maxiter"]))
    _say("C5 serving, budgets", [r.maxiter for r in reqs])
    with Batch(np.float32, hbm_budget=160 << 30, max_volumes=16) as b:
        stats = Server(b, chunk=25, max_live_voxels=60_000_000).run(reqs)
    assert stats.n == 8 and stats.peak_live >= 2
    for r in reqs:
        assert r.iters == r.maxiter
        _say(f"C5 solo request {r.rid} {r.d2f.shape} {r.maxiter} iterations")
        solo = _solo(r.kind, np.float32, r.d2f.shape, r.vsz, r.lam, r.alpha, r.d2f, r.mask,
                     r.maxiter)
        np.testing.assert_array_equal(r.chi, solo, err_msg=f"request {r.rid}")
    short = min((r for r in reqs if r.d2f.shape == (224, 224, 128)), key=lambda r: r.maxiter)
    _say(f"C5 oracle, request {short.rid} ({short.maxiter} iterations)")
    ref = _oracle(short.kind, short.d2f, short.mask, short.vsz, short.lam, short.alpha,
                  short.maxiter)
    np.testing.assert_array_equal(short.chi, ref)


@pytest.mark.timeout(600)
def test_fp64_tgv_bank_above_2gib(cuda):
    """FP64 TGV at 320x320x192: 15 arrays x 19.66 M voxels x 8 B = 2.36 GB
    per bank, so array offsets pass 2^31 bytes.  Fused pass (default
    schedule), two-sweep schedule, batch and oracle agree bitwise after 12
    iterations (one convergence check)."""
    shape = (320, 320, 192)
    bank = 15 * int(np.prod(shape)) * 8
    assert bank > 2 ** 31
    spec = configs._spec(shape, (1.0, 1.0, 1.0), (0.0, 0.0, 1.0), 7777)
    d2f, m = phantom.solver_input(spec)
    lam, alpha, iters = 1e-3, 2.0, 12
    _say("bank > 2 GiB: fused")
    fused = _solo(_lib.TGV, np.float64, shape, spec.vsz, lam, alpha, d2f, m, iters)
    _say("bank > 2 GiB: two sweeps")
    with pd.Solver(_lib.TGV, np.float64, shape, spec.vsz, lam, alpha) as s:
        s.set_schedule(1)
        s.load(d2f, m)
        s.iterate(iters)
        two = s.chi()
    np.testing.assert_array_equal(fused, two)
    for tiles in ((64, 8), (0, 0)):
        _say(f"bank > 2 GiB: batch, tiles {tiles}")
        with Batch(np.float64, tiles=tiles, hbm_budget=160 << 30, max_volumes=2) as b:
            slot = b.add(_lib.TGV, d2f, m, spec.vsz, lam, alpha, maxiter=iters)
            b.step(10)
            b.step(10)
            b.sync()
            got = b.result(slot)
        assert got.iters == iters
        np.testing.assert_array_equal(got.chi, fused)
    _say("bank > 2 GiB: one-iteration passes")
    with pd.Solver(_lib.TGV, np.float64, shape, spec.vsz, lam, alpha) as s:
        s.set_schedule(2)
        s.load(d2f, m)
        s.iterate(iters)
        np.testing.assert_array_equal(s.chi(), fused)
    _say("bank > 2 GiB: oracle")
    ref = orc.pd_tgv(d2f, m, spec.vsz, lam, alpha, 0.0, iters)[0]
    np.testing.assert_array_equal(fused, ref)
    assert np.abs(ref).max() > 0
