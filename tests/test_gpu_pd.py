"""GPU parity of the primal-dual path against the oracle (C restatement of
src/inversion/mex/pd_tgv_mex.c and pd_tv_mex.c).

FP64 and FP32 are compared BIT FOR BIT: the kernels keep the reference's
association order and are built without FMA contraction, so any difference
is a bug.  (The drop-in's tolerances, relative L2 1e-12 per operator and
1e-10 on chi, are asserted as well so a failure message shows both.)
"""
import ctypes as C

import numpy as np
import pytest

from conftest import rel_l2
from helpers import brainish_mask, from_dev, rand_field, to_dev
from oracle import oracle as orc
from qsm_amd import _lib, pd
from qsm_amd._lib import check, h3, lib, sz3

pytestmark = pytest.mark.gpu

SHAPES = [(12, 13, 14), (33, 20, 17)]
HS = [(1.0, 1.0, 1.0), (0.9, 0.9, 1.2)]
DTYPES = [np.float64, np.float32]


def _code(dt):
    return _lib.F64 if np.dtype(dt) == np.float64 else _lib.F32


def _vec(rng, shape, n, dt, scale):
    return np.asfortranarray((rng.standard_normal(shape + (n,)) * scale).astype(dt))


def _assert_same(gpu, ref, what):
    gpu = np.asarray(gpu)
    ref = np.asarray(ref)
    assert rel_l2(gpu, ref) <= 1e-12, f"{what}: rel L2 {rel_l2(gpu, ref):.3e}"
    np.testing.assert_array_equal(gpu, ref, err_msg=what)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("h", HS)
def test_tgv_dual_sweep(cuda, dt, shape, h):
    rng = np.random.default_rng(7)
    m = brainish_mask(shape, seed=1)
    mi = orc.init_mask(m)
    nu = rand_field(rng, shape, dt)
    p = _vec(rng, shape, 3, dt, 0.5)
    q = _vec(rng, shape, 6, dt, 0.5)
    cb = rand_field(rng, shape, dt)
    eb = rand_field(rng, shape, dt)
    wb = _vec(rng, shape, 3, dt, 1.0)
    f = rand_field(rng, shape, dt)
    tau, sigma = orc.step_sizes(h)
    a1, a0 = 0.3, 0.6  # projections active on part of the volume
    dev = [to_dev(a) for a in (nu, p, q, cb, eb, wb, f)]
    dmi = to_dev(mi)
    check(lib().qsm_tgv_dual_dev(_code(dt), *(d.data_ptr() for d in dev), dmi.data_ptr(),
                                 sz3(shape), h3(h), a0, a1, sigma, None))
    for a in (nu, p, q):
        pass
    orc.tgv_dual(nu, p, q, cb, eb, wb, f, mi, h, a0, a1, sigma)
    _assert_same(from_dev(dev[0], shape), nu, "nu")
    _assert_same(from_dev(dev[1], shape + (3,)), p, "p")
    _assert_same(from_dev(dev[2], shape + (6,)), q, "q")


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("h", HS)
def test_tgv_primal_sweep(cuda, dt, shape, h):
    rng = np.random.default_rng(8)
    m = brainish_mask(shape, seed=2)
    chi = rand_field(rng, shape, dt)
    eta = rand_field(rng, shape, dt)
    w = _vec(rng, shape, 3, dt, 1.0)
    cb = np.zeros_like(chi)
    eb = np.zeros_like(eta)
    wb = np.zeros_like(w)
    nu = rand_field(rng, shape, dt)
    p = _vec(rng, shape, 3, dt, 1.0)
    q = _vec(rng, shape, 6, dt, 1.0)
    tau, sigma = orc.step_sizes(h)
    dev = [to_dev(a) for a in (chi, eta, w, cb, eb, wb, nu, p, q)]
    dm = to_dev(m)
    check(lib().qsm_tgv_primal_dev(_code(dt), *(d.data_ptr() for d in dev), dm.data_ptr(),
                                   sz3(shape), h3(h), tau, None))
    orc.tgv_primal(chi, eta, w, cb, eb, wb, nu, p, q, m, h, tau)
    for d, r, s, n in zip(dev[:6], (chi, eta, w, cb, eb, wb),
                          (shape, shape, shape + (3,), shape, shape, shape + (3,)),
                          ("chi", "eta", "w", "chi_", "eta_", "w_")):
        _assert_same(from_dev(d, s), r, n)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape", SHAPES)
def test_tv_sweeps(cuda, dt, shape):
    h = (0.9, 1.0, 1.2)
    rng = np.random.default_rng(9)
    m = brainish_mask(shape, seed=3)
    mi = orc.init_mask(m)
    tau, sigma = orc.step_sizes(h)
    # dual
    nu = rand_field(rng, shape, dt)
    p = _vec(rng, shape, 3, dt, 0.5)
    cb = rand_field(rng, shape, dt)
    eb = rand_field(rng, shape, dt)
    f = rand_field(rng, shape, dt)
    dev = [to_dev(a) for a in (nu, p, cb, eb, f)]
    dmi = to_dev(mi)
    lam = 0.25
    check(lib().qsm_tv_dual_dev(_code(dt), *(d.data_ptr() for d in dev), dmi.data_ptr(),
                                sz3(shape), h3(h), lam, sigma, None))
    orc.tv_dual(nu, p, cb, eb, f, mi, h, lam, sigma)
    _assert_same(from_dev(dev[0], shape), nu, "nu")
    _assert_same(from_dev(dev[1], shape + (3,)), p, "p")
    # primal
    chi = rand_field(rng, shape, dt)
    eta = rand_field(rng, shape, dt)
    cb2 = np.zeros_like(chi)
    eb2 = np.zeros_like(chi)
    dev = [to_dev(a) for a in (chi, eta, cb2, eb2, nu, p)]
    dm = to_dev(m)
    check(lib().qsm_tv_primal_dev(_code(dt), *(d.data_ptr() for d in dev), dm.data_ptr(),
                                  sz3(shape), h3(h), tau, None))
    orc.tv_primal(chi, eta, cb2, eb2, nu, p, m, h, tau)
    for d, r, n in zip(dev[:4], (chi, eta, cb2, eb2), ("chi", "eta", "chi_", "eta_")):
        _assert_same(from_dev(d, shape), r, n)


@pytest.mark.parametrize("shape", SHAPES)
def test_init_mask(cuda, shape):
    m = brainish_mask(shape, seed=4)
    mi = orc.init_mask(m)
    dmi = to_dev(np.zeros_like(m))
    dm = to_dev(m)
    check(lib().qsm_init_mask_dev(dmi.data_ptr(), dm.data_ptr(), sz3(shape), None))
    np.testing.assert_array_equal(from_dev(dmi, shape), mi)


@pytest.mark.parametrize("dt", DTYPES)
def test_convergence_sums(cuda, dt):
    rng = np.random.default_rng(10)
    shape = (30, 21, 17)
    a = [rand_field(rng, shape, dt) for _ in range(4)]
    a[1] = (a[0] + 1e-3 * a[1]).astype(dt)
    a[3] = (a[2] + 1e-3 * a[3]).astype(dt)
    n1, n2, flag = orc.convergence(a[0], a[1], a[2], a[3], 0.1)
    dev = [to_dev(x) for x in a]
    out = (C.c_double * 2)()
    fl = C.c_int()
    check(lib().qsm_convergence_dev(_code(dt), *(d.data_ptr() for d in dev), a[0].size, 0.1,
                                    out, C.byref(fl), None))
    # summation order differs (OpenMP vs tree): agree to rounding
    assert abs(out[0] - n1) <= 1e-13 * abs(n1)
    assert abs(out[1] - n2) <= 1e-13 * abs(n2)
    assert fl.value == flag


def _phantom_input(shape, h, seed, dt):
    rng = np.random.default_rng(seed)
    x = np.linspace(-1, 1, shape[0])[:, None, None]
    y = np.linspace(-1, 1, shape[1])[None, :, None]
    z = np.linspace(-1, 1, shape[2])[None, None, :]
    chi = 0.1 * np.exp(-((x - 0.2) ** 2 + y ** 2 + (z + 0.1) ** 2) * 8) - 0.05 * (x * x + y * y + z * z < 0.2)
    fld = chi * (3 * z * z - 1) + 0.01 * rng.standard_normal(shape)
    d2f = -orc.lapw(np.asfortranarray(fld), h)
    return np.asfortranarray(d2f.astype(dt))


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("kind", ["tgv", "tv"])
@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_whole_solve_fixed_iterations(cuda, dt, kind, schedule):
    shape, h = (26, 23, 21), (0.9, 0.9, 1.2)
    d2f = _phantom_input(shape, h, 11, dt)
    m = brainish_mask(shape, seed=5, frac=0.85)
    iters = 137
    if kind == "tgv":
        ref, it, nr = orc.pd_tgv(d2f, m, h, 3e-3, 2.0, 0.0, iters)
        s = pd.Solver(_lib.TGV, dt, shape, h, 3e-3, 2.0)
    else:
        ref, it, nr = orc.pd_tv(d2f, m, h, 3e-3, 0.0, iters)
        s = pd.Solver(_lib.TV, dt, shape, h, 3e-3)
    with s:
        s.set_schedule(schedule)
        s.load(d2f, m).iterate(iters, 0.0)
        got = s.chi()
        it_g, stopped, nr_g = s.status()
    assert it_g == it == iters and not stopped
    assert rel_l2(got, ref) <= 1e-10
    np.testing.assert_array_equal(got, ref)
    assert abs(nr_g[0] - nr[0]) <= 1e-12 * abs(nr[0])


@pytest.mark.parametrize("dt", DTYPES)
def test_dropin_tolerance_stop(cuda, dt):
    """tol > 0: same stopping iteration and bitwise the same chi."""
    shape, h = (24, 22, 20), (1.0, 1.0, 1.0)
    d2f = _phantom_input(shape, h, 12, dt)
    m = brainish_mask(shape, seed=6, frac=0.85)
    ref, it, nr = orc.pd_tgv(d2f, m, h, 1e-3, 2.0, 5e-3, 2000)
    assert 10 <= it < 2000 and it % 10 == 0
    got = pd.pd_tgv_mex(d2f, m, h, 1e-3, 2.0, 5e-3, 2000, 0)
    np.testing.assert_array_equal(got, ref)


def test_fused_vs_two_sweep_c1_size(cuda):
    """64^3 TV and TGV, 60 iterations: the fused single-pass schedule and the
    reference's two-sweep schedule agree bit for bit (several tile shapes
    are covered by the tile-shape test in test_gpu_tuning)."""
    shape, h = (64, 64, 64), (1.0, 1.0, 1.0)
    d2f = _phantom_input(shape, h, 13, np.float64)
    m = brainish_mask(shape, seed=7, frac=0.9)
    for kind in (_lib.TGV, _lib.TV):
        outs = []
        for sched in (0, 1, 2):
            with pd.Solver(kind, np.float64, shape, h, 1e-3, 2.0) as s:
                s.set_schedule(sched)
                s.load(d2f, m).iterate(60, 0.0)
                outs.append(s.chi())
        np.testing.assert_array_equal(outs[0], outs[1])
        np.testing.assert_array_equal(outs[2], outs[1])


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("kind", ["tgv", "tv"])
def test_fused_tile_shapes(cuda, dt, kind):
    """Every supported fused tile shape, with partial tiles in x, y and z
    (70x45x33) and a holed mask, matches the oracle bit for bit."""
    shape, h = (70, 45, 33), (0.9, 1.0, 1.1)
    d2f = _phantom_input(shape, h, 14, dt)
    m = brainish_mask(shape, seed=8, frac=0.95)
    iters = 23
    if kind == "tgv":
        ref, _, _ = orc.pd_tgv(d2f, m, h, 2e-3, 2.0, 0.0, iters)
        k = _lib.TGV
    else:
        ref, _, _ = orc.pd_tv(d2f, m, h, 2e-3, 0.0, iters)
        k = _lib.TV
    for tx, ty in ((64, 4), (64, 8), (64, 16), (32, 8), (32, 16), (32, 32)):
        for kz in (1, 5, 31):
            with pd.Solver(k, dt, shape, h, 2e-3, 2.0) as s:
                s.set_tiles(tx, ty, kz)
                s.load(d2f, m).iterate(iters, 0.0)
                got = s.chi()
            np.testing.assert_array_equal(got, ref, err_msg=f"tiles {tx}x{ty} kz={kz}")



@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("kind", ["tgv", "tv"])
def test_two_iteration_pass_chunks_and_partial_tiles(cuda, dt, kind):
    """The two-iteration pass with partial tiles in x, y and z (70x45x33, a
    holed mask) and z-chunks of 1, 2, 5 and 31 planes (chunk edges inside
    the four-stage pipeline) matches the oracle bit for bit."""
    shape, h = (70, 45, 33), (0.9, 1.0, 1.1)
    d2f = _phantom_input(shape, h, 15, dt)
    m = brainish_mask(shape, seed=9, frac=0.95)
    iters = 24
    k = _lib.TGV if kind == "tgv" else _lib.TV
    if kind == "tgv":
        ref, _, _ = orc.pd_tgv(d2f, m, h, 2e-3, 2.0, 0.0, iters)
    else:
        ref, _, _ = orc.pd_tv(d2f, m, h, 2e-3, 0.0, iters)
    for kz in (1, 2, 5, 31):
        with pd.Solver(k, dt, shape, h, 2e-3, 2.0) as s:
            s.set_kz2(kz)
            s.load(d2f, m).iterate(iters, 0.0)
            got = s.chi()
        np.testing.assert_array_equal(got, ref, err_msg=f"kz2={kz}")


@pytest.mark.parametrize("dt", DTYPES)
def test_two_iteration_pass_uneven_calls_and_stop(cuda, dt):
    """Iterations enqueued in uneven calls (3, 7, 13, ...: a check then falls
    on the first iteration of a would-be pair, which runs alone) and a
    tolerance stop: same stopping iteration and chi as the oracle."""
    shape, h = (30, 28, 26), (1.0, 1.0, 1.0)
    d2f = _phantom_input(shape, h, 16, dt)
    m = brainish_mask(shape, seed=10, frac=0.9)
    ref, it, _ = orc.pd_tgv(d2f, m, h, 1e-3, 2.0, 0.0, 57)
    with pd.Solver(_lib.TGV, dt, shape, h, 1e-3, 2.0) as s:
        s.load(d2f, m)
        for n in (3, 7, 13, 1, 2, 31):
            s.iterate(n, 0.0)
        got = s.chi()
        assert s.status()[0] == 57
    np.testing.assert_array_equal(got, ref)
    ref, it, nr = orc.pd_tgv(d2f, m, h, 1e-3, 2.0, 5e-3, 2000)
    assert 10 <= it < 2000
    with pd.Solver(_lib.TGV, dt, shape, h, 1e-3, 2.0) as s:
        s.load(d2f, m)
        for n in (9, 4, 1987):
            s.iterate(n, 5e-3)
        got = s.chi()
        it_g, stopped, nr_g = s.status()
    assert stopped and it_g == it
    np.testing.assert_array_equal(got, ref)
    import torch

    with pd.Solver(_lib.TGV, dt, shape, h, 1e-3, 2.0) as s:
        s.load(d2f, m).iterate(2000, 5e-3)
        dev = torch.zeros(int(np.prod(shape)), device="cuda",
                          dtype=torch.float64 if dt == np.float64 else torch.float32)
        s.chi(out=dev)  # device copy of the bank the stop recorded
        np.testing.assert_array_equal(from_dev(dev, shape), ref)


def test_launch_counts_follow_the_schedule(cuda):
    """The handle's launch counters: pairs of iterations share a launch
    unless a check falls on the first of them or the call's count is odd;
    schedule 2 launches once per iteration, schedule 1 sweeps once per
    iteration; one check every tenth iteration."""
    shape, h = (24, 22, 20), (1.0, 1.0, 1.0)
    d2f = _phantom_input(shape, h, 4, np.float32)
    m = np.ones(shape, np.uint8)
    with pd.Solver(_lib.TGV, np.float32, shape, h, 1e-3, 2.0) as s:
        s.load(d2f, m)
        s.iterate(3)       # (1, 2) paired, 3 alone
        assert s.launch_counts() == {"pass2": 1, "pass1": 1, "sweeps": 0, "checks": 0}
        s.iterate(7)       # (4, 5) (6, 7) (8, 9) paired, 10 alone with its check
        assert s.launch_counts() == {"pass2": 4, "pass1": 2, "sweeps": 0, "checks": 1}
        s.iterate(10)      # (11, 12) ... (19, 20), the check of 20 after the last pair
        assert s.launch_counts() == {"pass2": 9, "pass1": 2, "sweeps": 0, "checks": 2}
        s.set_schedule(2).iterate(20)
        assert s.launch_counts() == {"pass2": 9, "pass1": 22, "sweeps": 0, "checks": 4}
        assert s.status()[0] == 40
    with pd.Solver(_lib.TGV, np.float32, shape, h, 1e-3, 2.0) as s:
        s.set_schedule(1).load(d2f, m).iterate(10)
        assert s.launch_counts() == {"pass2": 0, "pass1": 0, "sweeps": 10, "checks": 1}


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("kind", [_lib.TGV, _lib.TV])
def test_active_items_match_the_uniform_grid(cuda, dt, kind):
    """A brain-like mask filling half the box: the two-iteration pass over
    the work list of active tiles (forced on) gives chi bitwise the uniform
    grid's and the oracle's, after an odd count with convergence checks and
    at a tolerance stop (same iteration)."""
    shape, h = (90, 70, 60), (1.0, 0.9, 1.1)
    d2f = _phantom_input(shape, h, 21, dt)
    m = brainish_mask(shape, seed=3, frac=0.5)
    alpha = 2.0 if kind == _lib.TGV else 0.0

    def oracle(tol, iters):
        if kind == _lib.TGV:
            return orc.pd_tgv(d2f, m, h, 1e-3, alpha, tol, iters)
        return orc.pd_tv(d2f, m, h, 1e-3, tol, iters)

    def solve(on, tol, iters, calls=None):
        with pd.Solver(kind, dt, shape, h, 1e-3, alpha) as s:
            s.set_active_items(on)
            s.load(d2f, m)
            n, kz = s.active_items()
            for c in (calls or [iters]):
                s.iterate(c, tol)
            it, stopped, _ = s.status()
            return s.chi(), n, kz, it, stopped

    chi_l, n, kz, it, _ = solve(2, 0.0, 37, [3, 24, 10])
    assert n > 0 and kz >= 8 and it == 37
    chi_u, n0, _, _, _ = solve(0, 0.0, 37)
    assert n0 == 0
    np.testing.assert_array_equal(chi_l, chi_u)
    np.testing.assert_array_equal(chi_l, oracle(0.0, 37)[0])
    ref, it_ref, _ = oracle(5e-3, 2000)
    assert 10 <= it_ref < 2000
    chi_s, _, _, it_s, stopped = solve(2, 5e-3, 2000)
    assert stopped and it_s == it_ref
    np.testing.assert_array_equal(chi_s, ref)


def test_active_items_cost_model_and_edge_masks(cuda):
    """The cost model keeps the uniform grid for a full-box mask (nothing to
    skip) and for an empty one, and picks the list for a small blob in a
    large box; switching the list off and on again on a loaded handle keeps
    the iterates bitwise."""
    shape, h = (160, 160, 120), (1.0, 1.0, 1.0)
    d2f = _phantom_input(shape, h, 5, np.float32)
    full = np.ones(shape, np.uint8)
    empty = np.zeros(shape, np.uint8)
    blob = brainish_mask(shape, seed=4, frac=0.35)
    with pd.Solver(_lib.TGV, np.float32, shape, h, 1e-3, 2.0) as s:
        assert s.load(d2f, full).active_items() == (0, 0)
        assert s.load(d2f, empty).active_items() == (0, 0)
        s.iterate(4)
        assert np.all(s.chi() == 0)
        s.load(d2f, blob)
        n, kz = s.active_items()
        assert n > 0 and kz >= 8
        s.iterate(6)
        s.set_active_items(0)
        assert s.active_items() == (0, 0)
        s.iterate(6)
        s.set_active_items(1)
        assert s.active_items()[0] == n
        s.iterate(9)
        got = s.chi()
    ref = orc.pd_tgv(d2f, blob, h, 1e-3, 2.0, 0.0, 21)[0]
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("dt", DTYPES)
def test_active_items_at_the_volume_ends(cuda, dt):
    """Masks active only near z = 1 and z = nz - 2 (the items' first and
    last planes clip to the interior), a tile active on one plane only, and
    an x/y edge tile: the work list bitwise the uniform grid and the
    oracle."""
    shape, h = (70, 40, 30), (1.0, 1.0, 1.0)
    d2f = _phantom_input(shape, h, 33, dt)
    m = np.zeros(shape, np.uint8)
    m[5:30, 4:20, 0:6] = 1          # bottom slab (z = 0 is outside the interior)
    m[40:69, 25:40, 26:30] = 1      # top slab, touching x and y edges
    m[50, 10, 15] = 1               # one voxel: its tile is active on one plane
    m = np.asfortranarray(m)

    def solve(on):
        with pd.Solver(_lib.TGV, dt, shape, h, 1e-3, 2.0) as s:
            s.set_active_items(on)
            s.load(d2f, m)
            n = s.active_items()[0]
            s.iterate(23)
            return s.chi(), n

    got, n = solve(2)
    assert n > 0
    ref_u, n0 = solve(0)
    assert n0 == 0
    np.testing.assert_array_equal(got, ref_u)
    np.testing.assert_array_equal(got, orc.pd_tgv(d2f, m, h, 1e-3, 2.0, 0.0, 23)[0])
