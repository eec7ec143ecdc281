"""Batched inversions: every volume of a mixed batch (uneven sizes, TV and
TGV, different iteration budgets and tolerances, volumes joining while
others run) equals the oracle's single solve bit for bit, with and without
graph replay."""
import numpy as np
import pytest

from helpers import brainish_mask
from oracle import oracle as orc
from qsm_amd import _lib
from qsm_amd.batch import Batch
from qsm_amd.pd import Solver
from test_gpu_pd import _phantom_input

pytestmark = pytest.mark.gpu

CASES = [
    # shape, vsz, kind, lam, alpha, iters, tol
    ((30, 26, 22), (1.0, 1.0, 1.0), _lib.TGV, 3e-3, 2.0, 40, 0.0),
    ((45, 33, 19), (0.9, 0.9, 1.2), _lib.TV, 2e-3, 0.0, 57, 0.0),
    ((20, 40, 28), (1.0, 1.1, 1.0), _lib.TGV, 1e-3, 3.0, 33, 0.0),
    ((24, 22, 20), (1.0, 1.0, 1.0), _lib.TGV, 1e-3, 2.0, 2000, 5e-3),  # stops on tol
    ((70, 17, 9), (1.0, 1.0, 1.0), _lib.TV, 5e-3, 0.0, 21, 0.0),
]


def _inputs(dt):
    out = []
    for i, (shape, vsz, kind, lam, alpha, iters, tol) in enumerate(CASES):
        d2f = _phantom_input(shape, vsz, 20 + i, dt)
        m = brainish_mask(shape, seed=30 + i, frac=0.9)
        out.append((d2f, m))
    return out


def _oracle(case, d2f, m):
    shape, vsz, kind, lam, alpha, iters, tol = case
    if kind == _lib.TGV:
        return orc.pd_tgv(d2f, m, vsz, lam, alpha, tol, iters)
    return orc.pd_tv(d2f, m, vsz, lam, tol, iters)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("tiles", [(64, 8), (0, 0)])
def test_batch_mix_matches_oracle(cuda, dt, graph, tiles):
    """(64, 8): the one-iteration march; (0, 0): the two-iteration march."""
    inputs = _inputs(dt)
    with Batch(dt, tiles=tiles, max_volumes=8) as b:
        b.set_graph(graph)
        slots = []
        # first three join now, the last two after two chunks (continuous)
        for case, (d2f, m) in list(zip(CASES, inputs))[:3]:
            shape, vsz, kind, lam, alpha, iters, tol = case
            slots.append(b.add(kind, d2f, m, vsz, lam, alpha, iters, tol, kz=7))
        b.step(10)
        b.step(10)
        for case, (d2f, m) in list(zip(CASES, inputs))[3:]:
            shape, vsz, kind, lam, alpha, iters, tol = case
            slots.append(b.add(kind, d2f, m, vsz, lam, alpha, iters, tol, kz=5))
        for _ in range(300):
            b.step(10)
            b.sync()
            if not any(b.state(s)[1] for s in slots):
                break
        for s, case, (d2f, m) in zip(slots, CASES, inputs):
            ref, it, nr = _oracle(case, d2f, m)
            r = b.result(s)
            assert r.iters == it, (case, r.iters, it)
            np.testing.assert_array_equal(r.chi, ref, err_msg=str(case))


def test_batch_result_same_alone_and_in_mix(cuda):
    inputs = _inputs(np.float32)
    alone = []
    for case, (d2f, m) in zip(CASES, inputs):
        shape, vsz, kind, lam, alpha, iters, tol = case
        with Batch(np.float32, max_volumes=2) as b:
            s = b.add(kind, d2f, m, vsz, lam, alpha, iters, tol, kz=9)
            while b.state(s)[1]:
                b.step(10)
                b.sync()
            alone.append(b.result(s))
    with Batch(np.float32, max_volumes=8) as b:
        slots = [b.add(c[2], d, m, c[1], c[3], c[4], c[5], c[6], kz=9)
                 for c, (d, m) in zip(CASES, inputs)]
        while any(b.state(s)[1] for s in slots):
            b.step(10)
            b.sync()
        for s, a in zip(slots, alone):
            r = b.result(s)
            assert r.iters == a.iters and r.norms == a.norms
            np.testing.assert_array_equal(r.chi, a.chi)


def test_batch_admission_budget(cuda):
    shape = (40, 40, 40)
    d2f = np.zeros(shape, np.float32, order="F")
    m = np.ones(shape, np.uint8, order="F")
    one = None
    with Batch(np.float32, hbm_budget=1) as b:
        with pytest.raises(_lib.QsmAdmissionError):
            b.add(_lib.TGV, d2f, m, (1, 1, 1), 1e-3, 2.0, 10)
    with Batch(np.float32, max_volumes=4) as b:
        one = b.volume_bytes(_lib.TGV, shape, 16, 10)
    with Batch(np.float32, hbm_budget=2 * one + one // 2, max_volumes=4) as b:
        s0 = b.add(_lib.TGV, d2f, m, (1, 1, 1), 1e-3, 2.0, 10, kz=16)
        s1 = b.add(_lib.TGV, d2f, m, (1, 1, 1), 1e-3, 2.0, 10, kz=16)
        assert b.used_bytes == 2 * one
        with pytest.raises(_lib.QsmAdmissionError):
            b.add(_lib.TGV, d2f, m, (1, 1, 1), 1e-3, 2.0, 10, kz=16)
        b.step(10)
        b.sync()
        b.remove(s0)
        s2 = b.add(_lib.TGV, d2f, m, (1, 1, 1), 1e-3, 2.0, 10, kz=16)
        assert s2 == s0 and b.used_bytes == 2 * one


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_two_iteration_batch_uneven_chunks(cuda, dt):
    """The two-iteration march with chunks of 7, 13, 1 and 25 iterations (a
    check then falls on the first iteration of a pair and runs alone, odd
    budgets end with a single iteration), volumes joining between chunks
    and a tolerance stop: every volume bitwise the oracle's, same stopping
    iteration."""
    inputs = _inputs(dt)
    with Batch(dt, max_volumes=8) as b:
        slots = []
        for case, (d2f, m) in list(zip(CASES, inputs))[:2]:
            shape, vsz, kind, lam, alpha, iters, tol = case
            slots.append(b.add(kind, d2f, m, vsz, lam, alpha, iters, tol, kz=6))
        b.step(7)
        for case, (d2f, m) in list(zip(CASES, inputs))[2:]:
            shape, vsz, kind, lam, alpha, iters, tol = case
            slots.append(b.add(kind, d2f, m, vsz, lam, alpha, iters, tol, kz=11))
        for chunk in (13, 1, 25, 7) * 100:
            b.step(chunk)
            b.sync()
            if not any(b.state(s)[1] for s in slots):
                break
        for s, case, (d2f, m) in zip(slots, CASES, inputs):
            ref, it, nr = _oracle(case, d2f, m)
            r = b.result(s)
            assert r.iters == it, (case, r.iters, it)
            np.testing.assert_array_equal(r.chi, ref, err_msg=str(case))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_two_iteration_batch_masked_and_empty_volumes(cuda, dt):
    """The two-iteration march's entries trimmed to each tile's active
    planes: a volume masked to two thin slabs and one with an empty mask
    (nothing to trim to: its uniform entries stay) batched together; the
    slab volume bitwise its solo solve and the oracle, the empty one all
    zeros after its full budget, as alone."""
    shape, vsz = (60, 40, 34), (1.0, 1.0, 1.0)
    rng = np.random.default_rng(17)
    d2f = np.asfortranarray((0.01 * rng.standard_normal(shape)).astype(dt))
    slab = np.zeros(shape, np.uint8)
    slab[4:40, 3:30, 2:9] = 1
    slab[20:58, 10:38, 24:32] = 1
    slab = np.asfortranarray(slab)
    empty = np.zeros(shape, np.uint8, order="F")
    iters = 24
    with Batch(dt, tiles=(0, 0), max_volumes=4) as b:
        s1 = b.add(_lib.TGV, d2f, slab, vsz, 1e-3, 2.0, iters, 0.0, kz=16)
        s2 = b.add(_lib.TV, d2f, empty, vsz, 1e-3, 0.0, iters, 0.0, kz=16)
        while b.state(s1)[1] or b.state(s2)[1]:
            b.step(10)
            b.sync()
        r1, r2 = b.result(s1), b.result(s2)
    assert r1.iters == iters and r2.iters == iters
    with Solver(_lib.TGV, dt, shape, vsz, 1e-3, 2.0) as s:
        s.load(d2f, slab).iterate(iters)
        solo = s.chi()
    np.testing.assert_array_equal(r1.chi, solo)
    np.testing.assert_array_equal(r1.chi, orc.pd_tgv(d2f, slab, vsz, 1e-3, 2.0, 0.0, iters)[0])
    assert np.all(r2.chi == 0)
