"""tv.m (TV dipole inversion by ADMM) and rts.m (rapid two-step: LSMR in
k-space, then the same ADMM) on the GPU against the oracle's restatement
(oracle/tv_admm.py) on phantom fields.

The FFTs make bitwise parity impossible (MATLAB's FFT is closed; the oracle
uses numpy's), so the bar is a tolerance: FP64 relative L2 of x <= 1e-9 and
the same residual trajectory (iteration count and every rho decision equal,
norms to 1e-8); FP32 relative L2 <= 1e-3 with the same iteration count.
Sizes include lengths with factors 11 and 13 (radix-11/13 FFT stages) and
a tilted B0 (Nyquist planes where D(k) != D(-k))."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import tv_admm as tva
from qsm_amd import QsmError, phantom
from qsm_amd import rts as grts
from qsm_amd import tv as gtv

pytestmark = pytest.mark.gpu

CASES = [
    # shape, vsz, bdir
    ((32, 28, 22), (1.0, 1.0, 1.0), (0.1, -0.2, 0.97)),
    ((26, 13, 15), (0.9, 1.1, 1.2), (0.0, 0.0, 1.0)),
    ((20, 18, 16), (1.0, 1.0, 1.5), (0.3, 0.0, 0.95)),
]


def _field(shape, vsz, bdir, seed):
    spec = phantom.PhantomSpec(shape=shape, vsz=vsz, n_ellipsoids=6, semi_axes=(2, 6),
                               mask="sphere", mask_param=(min(shape) / 2 - 2,), noise=0.01,
                               seed=seed, bdir=bdir)
    ph = phantom.make(spec)
    return ph["field"], ph["mask"]


@pytest.mark.parametrize("mtv", [1, 2])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_tv_matches_oracle(cuda, case, dt, mtv):
    shape, vsz, bdir = CASES[case]
    f, m = _field(shape, vsz, bdir, seed=10 + case)
    f = f.astype(dt)
    opts = {"tv": mtv, "verbose": 0}
    want, wtr = tva.tv(f, m, vsz, bdir, opts)
    got, gtr = gtv.tv(f, m, vsz, bdir, opts, return_trace=True)
    assert got.dtype == f.dtype and got.shape == f.shape
    assert len(gtr) == len(wtr) > 2
    if dt == np.float64:
        assert rel_l2(got, want) < 1e-9
        np.testing.assert_array_equal(gtr[:, 4], wtr[:, 4])      # every rho decision
        np.testing.assert_allclose(gtr[:, :4], wtr[:, :4], rtol=1e-8)
    else:
        assert rel_l2(got, want) < 1e-3
        np.testing.assert_allclose(gtr[:, :4], wtr[:, :4], rtol=1e-2)
    assert np.all(got[m == 0] == 0)


def test_tv_multi_echo_is_per_echo(cuda):
    shape, vsz, bdir = CASES[1]
    f1, m = _field(shape, vsz, bdir, seed=3)
    f2, _ = _field(shape, vsz, bdir, seed=4)
    f = np.stack([f1, 2.0 * f2], axis=3)
    x = gtv.tv(f, m, vsz, bdir, {"verbose": 0})
    for t in range(2):
        np.testing.assert_array_equal(x[..., t], gtv.tv(f[..., t], m, vsz, bdir, {"verbose": 0}))


@pytest.mark.parametrize("dt", ["float64", "float32"])
def test_tv_device_form_on_side_stream(cuda, dt):
    """qsm_tv_dev on a torch side stream (work space not cleared, reused)
    matches the host form bitwise, twice in a row."""
    import torch

    shape, vsz, bdir = CASES[0]
    f, m = _field(shape, vsz, bdir, seed=7)
    f = f.astype(dt)
    want = gtv.tv(f, m, vsz, bdir, {"verbose": 0})
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        tf = torch.from_numpy(np.ascontiguousarray(f.reshape(-1, order="F"))).to("cuda")
        tf = tf.view(shape[::-1]).permute(2, 1, 0)
        tm = torch.from_numpy(np.ascontiguousarray(m.astype(np.uint8).reshape(-1, order="F")))
        tm = tm.to("cuda").view(shape[::-1]).permute(2, 1, 0)
        out = torch.empty_like(tf)
        work = torch.full((gtv.work_bytes(tf.dtype, shape),), 255, dtype=torch.uint8, device="cuda")
    for _ in range(2):
        gtv.tv_dev(tf, tm, vsz, bdir, {"verbose": 0}, out, work=work, stream=st)
        st.synchronize()
        got = out.permute(2, 1, 0).contiguous().cpu().numpy().reshape(-1).reshape(shape, order="F")
        np.testing.assert_array_equal(got, want)


def test_tv_rejects_lengths_with_large_prime_factors(cuda):
    f = np.zeros((17, 8, 8))
    with pytest.raises(QsmError, match="prime factor"):
        gtv.tv(f, np.ones(f.shape), (1, 1, 1), None, {"verbose": 0})


@pytest.mark.parametrize("mtv", [1, 2])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_rts_matches_oracle(cuda, case, dt, mtv):
    """rts.m: step 1 (4 LSMR iterations on the full spectrum) and step 2."""
    shape, vsz, bdir = CASES[case]
    f, m = _field(shape, vsz, bdir, seed=30 + case)
    f = f.astype(dt)
    opts = {"tv": mtv, "verbose": 0}
    want, wtr = tva.rts(f, m, vsz, bdir, opts)
    got, gtr = grts.rts(f, m, vsz, bdir, opts, return_trace=True)
    assert got.dtype == f.dtype and got.shape == f.shape
    assert len(gtr) == len(wtr) >= 1
    if dt == np.float64:
        assert rel_l2(got, want) < 1e-9
        np.testing.assert_array_equal(gtr[:, 4], wtr[:, 4])
        np.testing.assert_allclose(gtr[:, :4], wtr[:, :4], rtol=1e-8)
    else:
        assert rel_l2(got, want) < 1e-3
        np.testing.assert_allclose(gtr[:, :4], wtr[:, :4], rtol=1e-2)
    assert np.all(got[m == 0] == 0)


def test_rts_tolls_zero_and_zero_field(cuda):
    """tolls = 0 gives x1 = 0 (LSMR does no iteration); a zero field gives
    x = 0."""
    shape, vsz, bdir = CASES[2]
    f, m = _field(shape, vsz, bdir, seed=5)
    for opts, fld in (({"tolls": 0, "verbose": 0}, f), ({"verbose": 0}, np.zeros_like(f))):
        want, _ = tva.rts(fld, m, vsz, bdir, opts)
        got = grts.rts(fld, m, vsz, bdir, opts)
        assert rel_l2(got, want) < 1e-9 or (np.abs(want).max() == 0 and np.abs(got).max() == 0)


@pytest.mark.parametrize("dt", ["float64", "float32"])
def test_rts_device_form_on_side_stream(cuda, dt):
    import torch

    shape, vsz, bdir = CASES[0]
    f, m = _field(shape, vsz, bdir, seed=9)
    f = f.astype(dt)
    want = grts.rts(f, m, vsz, bdir, {"verbose": 0})
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        tf = torch.from_numpy(np.ascontiguousarray(f.reshape(-1, order="F"))).to("cuda")
        tf = tf.view(shape[::-1]).permute(2, 1, 0)
        tm = torch.from_numpy(np.ascontiguousarray(m.astype(np.uint8).reshape(-1, order="F")))
        tm = tm.to("cuda").view(shape[::-1]).permute(2, 1, 0)
        out = torch.empty_like(tf)
        work = torch.full((grts.work_bytes(tf.dtype, shape),), 255, dtype=torch.uint8, device="cuda")
    for _ in range(2):
        grts.rts_dev(tf, tm, vsz, bdir, {"verbose": 0}, out, work=work, stream=st)
        st.synchronize()
        got = out.permute(2, 1, 0).contiguous().cpu().numpy().reshape(-1).reshape(shape, order="F")
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(24, 256, 176), (192, 20, 16)])
def test_tv_register_pass_lengths(cuda, dt, shape):
    """tv.m on unpadded grids whose lengths run the register-resident FFT
    passes: y (256 = 16x16) and z (176 = 16x11), or x (192 = 16x12, the
    real-row x passes); the oracle to the tolerance with the same rho
    decisions (FP64)."""
    vsz, bdir = (1.0, 1.0, 1.0), (0.1, -0.2, 0.97)
    f, m = _field(shape, vsz, bdir, seed=31)
    f = f.astype(dt)
    opts = {"tv": 1, "verbose": 0}
    want, wtr = tva.tv(f, m, vsz, bdir, opts)
    got, gtr = gtv.tv(f, m, vsz, bdir, opts, return_trace=True)
    assert len(gtr) == len(wtr) > 2
    if dt == np.float64:
        assert rel_l2(got, want) < 1e-9
        np.testing.assert_array_equal(gtr[:, 4], wtr[:, 4])
    else:
        assert rel_l2(got, want) < 1e-3
