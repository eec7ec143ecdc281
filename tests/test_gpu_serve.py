"""Serving: requests of mixed sizes arrive while others run; each result is
bitwise the same as the request solved alone (the oracle's solve)."""
import numpy as np
import pytest

from helpers import brainish_mask
from oracle import oracle as orc
from qsm_amd import _lib
from qsm_amd.batch import Batch
from qsm_amd.serve import Request, Server, poisson_arrivals
from test_gpu_pd import _phantom_input

pytestmark = pytest.mark.gpu


def test_serving_results_equal_solo_solves(cuda):
    rng = np.random.default_rng(5)
    shapes = [(28, 24, 20), (36, 30, 18), (22, 22, 22), (40, 20, 16), (30, 30, 24), (26, 34, 21)]
    arrivals = poisson_arrivals(50.0, len(shapes), rng)
    reqs = []
    for i, (shape, t) in enumerate(zip(shapes, arrivals)):
        kind = _lib.TGV if i % 2 == 0 else _lib.TV
        vsz = (1.0, 1.0, 1.0) if i % 3 else (0.9, 0.9, 1.2)
        d2f = _phantom_input(shape, vsz, 40 + i, np.float32)
        m = brainish_mask(shape, seed=50 + i, frac=0.9)
        iters = int(rng.integers(30, 90))
        reqs.append(Request(i, float(t), kind, d2f, m, vsz, 2e-3, 2.0, iters, kz=8))
    # a budget that holds about three of these at once forces queueing
    with Batch(np.float32, max_volumes=8) as probe:
        one = probe.volume_bytes(_lib.TGV, (40, 30, 24), 8, 100)
    with Batch(np.float32, hbm_budget=3 * one, max_volumes=8) as b:
        stats = Server(b, chunk=10).run(reqs)
    assert stats.n == len(reqs) and stats.p99 >= stats.p50 > 0
    for r in reqs:
        if r.kind == _lib.TGV:
            ref, it, _ = orc.pd_tgv(r.d2f, r.mask, r.vsz, r.lam, r.alpha, 0.0, r.maxiter)
        else:
            ref, it, _ = orc.pd_tv(r.d2f, r.mask, r.vsz, r.lam, 0.0, r.maxiter)
        assert r.iters == it == r.maxiter
        np.testing.assert_array_equal(r.chi, ref, err_msg=f"request {r.rid}")


def test_serving_with_live_voxel_cap(cuda):
    """With the live work capped at two volumes' worth, requests still match
    the oracle bitwise, at most two run at once, and they finish in arrival
    order (equal budgets)."""
    rng = np.random.default_rng(11)
    shape = (26, 24, 20)
    reqs = []
    for i, t in enumerate(poisson_arrivals(200.0, 5, rng)):
        kind = _lib.TGV if i % 2 else _lib.TV
        d2f = _phantom_input(shape, (1.0, 1.0, 1.0), 70 + i, np.float32)
        m = brainish_mask(shape, seed=80 + i, frac=0.9)
        reqs.append(Request(i, float(t), kind, d2f, m, (1.0, 1.0, 1.0), 2e-3, 2.0, 40, kz=8))
    with Batch(np.float32, max_volumes=8) as b:
        stats = Server(b, chunk=10, max_live_voxels=2 * int(np.prod(shape))).run(reqs)
    assert stats.peak_live <= 2
    done = [r.done for r in reqs]
    assert done == sorted(done)
    for r in reqs:
        if r.kind == _lib.TGV:
            ref = orc.pd_tgv(r.d2f, r.mask, r.vsz, r.lam, r.alpha, 0.0, r.maxiter)[0]
        else:
            ref = orc.pd_tv(r.d2f, r.mask, r.vsz, r.lam, 0.0, r.maxiter)[0]
        np.testing.assert_array_equal(r.chi, ref, err_msg=f"request {r.rid}")
