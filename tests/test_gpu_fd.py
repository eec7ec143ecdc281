"""GPU finite-difference operators (gradfp, gradfp_adj, gradf, gradf_adj,
gradfm, gradfm_adj, lapw, lap1) against the oracle, bitwise, through both the
host-pointer entry points (the MEX shape: outputs overwritten in place) and
the device-pointer *_dev entry points on a torch stream.  Shapes include
degenerate axes (length 1 and 2), which take the reference's special
branches (single-voxel wrap, one-face volumes)."""
import ctypes as C

import numpy as np
import pytest

from helpers import from_dev, rand_field, to_dev
from oracle import oracle as orc
from qsm_amd import _lib, fd

pytestmark = pytest.mark.gpu

SHAPES = [(33, 20, 17), (64, 7, 9), (1, 5, 6), (5, 1, 1), (2, 2, 2), (130, 3, 70)]
H = (0.9, 1.1, 1.3)
DTS = [np.float64, np.float32]


def _mask(shape, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray((rng.random(shape) < 0.7).astype(np.uint8))


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    np.testing.assert_array_equal(np.ascontiguousarray(a).view(np.uint8),
                                  np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dt", DTS)
def test_gradients_host(cuda, shape, dt):
    rng = np.random.default_rng(0)
    u = rand_field(rng, shape, dt)
    G = _mask(shape, 1)
    cases = [(fd.gradfp(u, H), orc.gradfp(u, H)), (fd.gradf(u, G, H), orc.gradfm(u, G, H))]
    if min(shape) >= 2:
        cases.append((fd.gradf(u, None, H), orc.gradf(u, H)))
    else:
        # the unmasked forward/backward pair needs two voxels per axis (the
        # reference reads out of bounds there); rejected, not computed
        with pytest.raises(_lib.QsmError, match="needs >= 2"):
            fd.gradf(u, None, H)
    for g, want in cases:
        for a, b in zip(g, want):
            _same(a, b)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dt", DTS)
def test_adjoints_host(cuda, shape, dt):
    rng = np.random.default_rng(2)
    x, y, z = (rand_field(rng, shape, dt) for _ in range(3))
    G = _mask(shape, 3)
    _same(fd.gradfp_adj(x, y, z, H), orc.gradfp_adj(x, y, z, H))
    _same(fd.gradf_adj(x, y, z, G, H), orc.gradfm_adj(x, y, z, G, H))
    if min(shape) >= 2:
        _same(fd.gradf_adj(x, y, z, None, H), orc.gradf_adj(x, y, z, H))
    else:
        with pytest.raises(_lib.QsmError, match="needs >= 2"):
            fd.gradf_adj(x, y, z, None, H)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dt", DTS)
def test_in_place_outputs_keep_untouched_voxels(cuda, shape, dt):
    """gradfm_mex / gradfm_adj_mex / lapw_mex overwrite only what the
    reference writes; the rest of the caller's array is left as passed."""
    rng = np.random.default_rng(4)
    u = rand_field(rng, shape, dt, 3.0)
    G = _mask(shape, 5)
    init = [rand_field(rng, shape, dt) for _ in range(3)]
    outs = [a.copy(order="F") for a in init]
    fd.gradfm_mex(*outs, u, G, H)
    for a, b in zip(outs, orc.gradfm(u, G, H, init=init)):
        _same(a, b)
    x, y, z = (rand_field(rng, shape, dt) for _ in range(3))
    du = init[0].copy(order="F")
    fd.gradfm_adj_mex(du, x, y, z, G, H)
    _same(du, orc.gradfm_adj(x, y, z, G, H, init=init[0]))
    d2u = init[1].copy(order="F")
    fd.lapw_mex(d2u, u, H)
    _same(d2u, orc.lapw(u, H, init=init[1]))


@pytest.mark.parametrize("dt", DTS)
def test_lapw_wrapped_phase(cuda, dt):
    """lapw of a phase wrapped many times over (values far outside
    [-pi, pi]) -- the wrap of every difference goes through fmod."""
    rng = np.random.default_rng(6)
    shape = (40, 36, 28)
    i, j, k = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    phase = 0.9 * i - 0.7 * j + 0.35 * k + 0.3 * rng.standard_normal(shape)
    u = np.asfortranarray((np.angle(np.exp(1j * phase)) + 2 * np.pi * rng.integers(-3, 4, shape)).astype(dt))
    _same(fd.lapw(u, H), orc.lapw(u, H))


@pytest.mark.parametrize("dt", DTS)
def test_lapw_special_values(cuda, dt):
    """Samples at and next to 0, +/-pi, +/-2pi, +/-3pi (where the wrap's
    branches and its 2-tau range limit sit), large values, +/-Inf and NaN,
    mixed with ordinary ones in the same wavefronts: bitwise where the
    oracle is a number, NaN where it is NaN."""
    rng = np.random.default_rng(11)
    T = np.dtype(dt).type
    base = [T(0), -T(0), T(1e3), T(-1e30), T(np.inf), T(-np.inf), T(np.nan)]
    for m in (1, 2, 3, 4):
        for s in (1, -1):
            v = T(s * m * np.pi)
            base += [v, np.nextafter(v, T(np.inf)), np.nextafter(v, T(-np.inf))]
    shape = (70, 9, 6)
    u = rng.standard_normal(shape).astype(dt)
    pick = rng.random(shape) < 0.5
    u[pick] = rng.choice(np.array(base, dtype=dt), size=int(pick.sum()))
    u = np.asfortranarray(u)
    with np.errstate(invalid="ignore"):
        want = orc.lapw(u, H)
    got = fd.lapw(u, H)
    nan = np.isnan(want)
    assert nan.any() and (~nan).any()
    np.testing.assert_array_equal(np.isnan(got), nan)
    _same(np.where(nan, 0, got).astype(dt), np.where(nan, 0, want).astype(dt))


def _dev_call(name, dt, outs_dev, ins_dev, shape, G_dev=None, stream=None):
    L = _lib.lib()
    args = [_lib.dtype_code(dt), *(t.data_ptr() for t in outs_dev), *(t.data_ptr() for t in ins_dev)]
    if G_dev is not None:
        args.append(C.cast(C.c_void_p(G_dev.data_ptr()), C.POINTER(C.c_uint8)))
    _lib.check(getattr(L, name)(*args, _lib.sz3(shape), _lib.h3(H),
                                C.c_void_p(stream.cuda_stream) if stream is not None else None))


@pytest.mark.parametrize("dt", DTS)
def test_device_entry_points(cuda, dt):
    import torch

    shape = (70, 45, 33)
    rng = np.random.default_rng(7)
    u = rand_field(rng, shape, dt)
    x, y, z = (rand_field(rng, shape, dt) for _ in range(3))
    G = _mask(shape, 8)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        du, dx, dy, dz = (to_dev(np.zeros(shape, dt)) for _ in range(4))
        U, X, Y, Z, Gd = to_dev(u), to_dev(x), to_dev(y), to_dev(z), to_dev(G)
    st.synchronize()

    def grab(*ts):
        st.synchronize()
        return [from_dev(t, shape) for t in ts]

    _dev_call("qsm_gradfp_dev", dt, (dx, dy, dz), (U,), shape, stream=st)
    for a, b in zip(grab(dx, dy, dz), orc.gradfp(u, H)):
        _same(a, b)
    _dev_call("qsm_gradf_dev", dt, (dx, dy, dz), (U,), shape, stream=st)
    for a, b in zip(grab(dx, dy, dz), orc.gradf(u, H)):
        _same(a, b)
    with torch.cuda.stream(st):  # ordered before the entry point on st
        for t in (dx, dy, dz):
            t.zero_()
    _dev_call("qsm_gradfm_dev", dt, (dx, dy, dz), (U,), shape, Gd, stream=st)
    for a, b in zip(grab(dx, dy, dz), orc.gradfm(u, G, H)):
        _same(a, b)
    _dev_call("qsm_gradfp_adj_dev", dt, (du,), (X, Y, Z), shape, stream=st)
    _same(grab(du)[0], orc.gradfp_adj(x, y, z, H))
    _dev_call("qsm_gradf_adj_dev", dt, (du,), (X, Y, Z), shape, stream=st)
    _same(grab(du)[0], orc.gradf_adj(x, y, z, H))
    with torch.cuda.stream(st):
        du.zero_()
    _dev_call("qsm_gradfm_adj_dev", dt, (du,), (X, Y, Z), shape, Gd, stream=st)
    _same(grab(du)[0], orc.gradfm_adj(x, y, z, G, H))
    with torch.cuda.stream(st):
        du.zero_()
    _dev_call("qsm_lapw_dev", dt, (du,), (U,), shape, stream=st)
    _same(grab(du)[0], orc.lapw(u, H))


def test_gradfp_adjoint_pair_on_gpu(cuda):
    """<gradfp u, p> = <u, gradfp_adj p> for the GPU pair (float64)."""
    rng = np.random.default_rng(9)
    shape = (31, 26, 19)
    u = rand_field(rng, shape, np.float64)
    p = [rand_field(rng, shape, np.float64) for _ in range(3)]
    g = fd.gradfp(u, H)
    du = fd.gradfp_adj(*p, H)
    lhs = sum((a * b).sum() for a, b in zip(g, p))
    assert abs(lhs - (u * du).sum()) <= 1e-12 * (abs(lhs) + 1)


BC_SHAPES = [(33, 20, 17), (64, 7, 9), (2, 2, 2), (5, 3, 2), (130, 3, 70), (1, 5, 6)]


@pytest.mark.parametrize("shape", BC_SHAPES)
@pytest.mark.parametrize("dt", DTS)
def test_backward_and_central_differences(cuda, shape, dt):
    """gradb, gradbm, gradb_adj, gradbm_adj, gradc, gradcm (src/utils/fd/mex)
    bitwise against the oracle, in place on outputs with random contents:
    voxels the reference leaves alone (outside G, the corner gradb_adj and
    gradc never write) keep them; the unmasked forms need two voxels per
    axis."""
    rng = np.random.default_rng(6)
    u = rand_field(rng, shape, dt, 2.0)
    G = _mask(shape, 7)
    x, y, z = (rand_field(rng, shape, dt) for _ in range(3))
    init = [rand_field(rng, shape, dt) for _ in range(3)]
    for mex, oracle_fn, masked in ((fd.gradbm_mex, orc.gradbm, True), (fd.gradcm_mex, orc.gradcm, True),
                                   (fd.gradb_mex, orc.gradb, False), (fd.gradc_mex, orc.gradc, False)):
        outs = [a.copy(order="F") for a in init]
        args = (u, G, H) if masked else (u, H)
        if not masked and min(shape) < 2:
            with pytest.raises(_lib.QsmError, match="needs >= 2"):
                mex(*outs, *args)
            continue
        mex(*outs, *args)
        for a, b in zip(outs, oracle_fn(*args, init=init)):
            _same(a, b)
    for mex, oracle_fn, masked in ((fd.gradbm_adj_mex, orc.gradbm_adj, True),
                                   (fd.gradb_adj_mex, orc.gradb_adj, False)):
        du = init[0].copy(order="F")
        args = (x, y, z, G, H) if masked else (x, y, z, H)
        if not masked and min(shape) < 2:
            with pytest.raises(_lib.QsmError, match="needs >= 2"):
                mex(du, *args)
            continue
        mex(du, *args)
        _same(du, oracle_fn(*args, init=init[0]))


@pytest.mark.parametrize("dt", DTS)
def test_backward_and_central_allocating_forms(cuda, dt):
    """gradb.m / gradbAdj.m / gradc.m: zeros, then the unmasked MEX for an
    all-true mask, the masked one otherwise."""
    shape = (24, 18, 12)
    rng = np.random.default_rng(8)
    u = rand_field(rng, shape, dt)
    G = _mask(shape, 9)
    for a, b in zip(fd.gradb(u, None, H), orc.gradb(u, H)):
        _same(a, b)
    for a, b in zip(fd.gradc(u, np.ones(shape), H), orc.gradc(u, H)):
        _same(a, b)
    for a, b in zip(fd.gradc(u, G, H), orc.gradcm(u, G, H)):
        _same(a, b)
    x, y, z = (rand_field(rng, shape, dt) for _ in range(3))
    _same(fd.gradb_adj(x, y, z, None, H), orc.gradb_adj(x, y, z, H))
    _same(fd.gradb_adj(x, y, z, G, H), orc.gradbm_adj(x, y, z, G, H))


LAP1_SHAPES = [(33, 20, 17), (64, 7, 9), (2, 2, 2), (5, 3, 2), (130, 3, 70), (4, 9, 4)]


@pytest.mark.parametrize("shape", LAP1_SHAPES)
@pytest.mark.parametrize("dt", DTS)
@pytest.mark.parametrize("frac", [0.5, 0.9, 1.0])
def test_lap1_host(cuda, shape, dt, frac):
    """lap1_mex (src/utils/fd/mex/lap1_mex.c) bitwise against the oracle, in
    place on an output with random contents (voxels outside G keep them);
    random masks of several densities pick every branch of the per-axis rule
    (central, one-sided of third / second / first order, none)."""
    rng = np.random.default_rng(int(frac * 10) + shape[0])
    u = rand_field(rng, shape, dt, 2.0)
    G = np.asfortranarray((rng.random(shape) < frac).astype(np.uint8))
    init = rand_field(rng, shape, dt)
    d = init.copy(order="F")
    fd.lap1_mex(d, u, G, H)
    _same(d, orc.lap1(u, G, H, init=init))


def test_lap1_rejects_unit_extent_and_allocating_form(cuda):
    rng = np.random.default_rng(12)
    u = rand_field(rng, (1, 5, 6), np.float64)
    with pytest.raises(_lib.QsmError, match="needs >= 2"):
        fd.lap(u)
    shape = (40, 31, 22)
    for dt in DTS:
        u = rand_field(rng, shape, dt)
        G = _mask(shape, 13)
        _same(fd.lap(u, G, H), orc.lap1(u, G, H))
        _same(fd.lap(u, None, H), orc.lap1(u, np.ones(shape, bool), H))


@pytest.mark.parametrize("dt", DTS)
def test_lap1_device_entry_point(cuda, dt):
    import torch

    shape = (70, 45, 33)
    rng = np.random.default_rng(14)
    u = rand_field(rng, shape, dt)
    G = _mask(shape, 15)
    init = rand_field(rng, shape, dt)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        du, U, Gd = to_dev(init), to_dev(u), to_dev(G)
    st.synchronize()
    _dev_call("qsm_lap1_dev", dt, (du,), (U,), shape, Gd, stream=st)
    st.synchronize()
    _same(from_dev(du, shape), orc.lap1(u, G, H, init=init))
