"""sstgv (src/inversion/sstgv.m:1-89) on the GPU against the oracle composed
the way the reference composes it: crop to the mask's bounding box +/- 2
voxels (sstgv.m:57-66), per echo lapw then pd_tgv_mex / pd_tv_mex on
-lapw(u) (sstgv.m:76-87), uncropped into zeros -- bitwise."""
import numpy as np
import pytest

from oracle import oracle as orc
from qsm_amd.sstgv import sstgv

pytestmark = pytest.mark.gpu


def _reference(phas, mask, vsz, lam, alpha, tol, maxit):
    x = np.zeros(phas.shape, dtype=phas.dtype, order="F")
    idx = [np.nonzero(mask.any(axis=ax))[0] for ax in ((1, 2), (0, 2), (0, 1))]
    sl = tuple(slice(max(0, i[0] - 2), min(n, i[-1] + 3)) for i, n in zip(idx, mask.shape))
    m = np.asfortranarray(mask[sl].astype(np.uint8))
    for t in range(phas.shape[3]):
        u = np.asfortranarray(phas[sl + (t,)])
        d2u = orc.lapw(u, vsz)
        f = np.asfortranarray(-d2u)
        if alpha > np.finfo(np.float64).eps:
            chi = orc.pd_tgv(f, m, vsz, lam, alpha, tol, maxit)[0]
        else:
            chi = orc.pd_tv(f, m, vsz, lam, tol, maxit)[0]
        x[sl + (t,)] = chi
    return x


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("alpha", [2.0, 0.0])
def test_sstgv_matches_oracle_composition(cuda, dt, alpha):
    rng = np.random.default_rng(21)
    shape = (26, 22, 20)
    c = [(n - 1) / 2 for n in shape]
    g = np.meshgrid(*(np.arange(n) - cc for n, cc in zip(shape, c)), indexing="ij")
    mask = ((g[0] / 8.0) ** 2 + ((g[1] - 1) / 7.0) ** 2 + (g[2] / 6.0) ** 2) <= 1.0
    phas = np.stack([0.8 * np.sin(0.3 * g[0] + 0.2 * g[1] * (t + 1)) + 0.1 * g[2]
                     + 0.05 * rng.standard_normal(shape) for t in range(2)], axis=3)
    phas = np.asfortranarray(np.angle(np.exp(1j * phas)).astype(dt))
    vsz = (0.9, 1.0, 1.1)
    got = sstgv(phas, mask, vsz, 2e-3, alpha, 1e-3, 120, False)
    want = _reference(phas, mask, vsz, 2e-3, alpha, 1e-3, 120)
    assert got.dtype == phas.dtype
    np.testing.assert_array_equal(got, want)
    # outside the cropped box (bounding box + 2) the result is zero
    idx = [np.nonzero(mask.any(axis=ax))[0] for ax in ((1, 2), (0, 2), (0, 1))]
    box = np.zeros(shape, bool)
    box[tuple(slice(max(0, i[0] - 2), min(n, i[-1] + 3)) for i, n in zip(idx, shape))] = True
    assert got[~box].size > 0 and not got[~box].any()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_sstgv_mask_on_the_faces_and_single_echo(cuda, dt):
    """A mask that reaches the volume's low faces on x and y and its high
    face on z: the box +/- 2 is clamped to the volume (sstgv.m:57-66).  A
    3-D phase is one echo."""
    rng = np.random.default_rng(5)
    shape = (21, 18, 16)
    mask = np.zeros(shape, bool)
    mask[0:12, 0:15, 5:16] = rng.random((12, 15, 11)) < 0.9
    mask[0, 0, 15] = mask[11, 14, 5] = True
    phas = np.asfortranarray(np.angle(np.exp(1j * rng.uniform(-4, 4, shape))).astype(dt))
    vsz = (1.0, 1.1, 0.9)
    got = sstgv(phas, mask, vsz, 3e-3, 2.0, 0.0, 40, False)
    want = _reference(phas[..., None], mask, vsz, 3e-3, 2.0, 0.0, 40)[..., 0]
    assert got.shape == shape and got.dtype == phas.dtype
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_sstgv_empty_mask_is_zero(cuda, dt):
    """An empty mask: nothing to invert, chi = 0 (sstgv.m's crop is empty)."""
    shape = (12, 10, 8)
    phas = np.asfortranarray(np.random.default_rng(2).uniform(-3, 3, shape + (2,)).astype(dt))
    got = sstgv(phas, np.zeros(shape, bool), (1, 1, 1), 3e-3, 2.0, 0.0, 10, False)
    assert got.shape == phas.shape and got.dtype == phas.dtype and not got.any()
