"""SHARP / V-SHARP (src/bgremove/sharp.m, vsharp.m) on the GPU against the
oracle's restatement (oracle/bgremove.py): eroded masks equal, local fields
to 1e-10 (FP64) / 1e-4 (FP32) relative -- the FFTs make bitwise parity
impossible (MATLAB's FFT is closed).  Includes anisotropic voxels, odd and
even crop sizes, a mask with an empty interior slice (cropIndices skips
it), 4-D echoes, an odd number of V-SHARP radii (the returned-mask rule)
and the harmonic-background property on the device."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import bgremove as obg
from qsm_amd import QsmError
from qsm_amd import bgremove as gbg

pytestmark = pytest.mark.gpu


def _case(shape, seed, hole=False, bg=1.0):
    rng = np.random.default_rng(seed)
    c = [(n - 1) / 2 for n in shape]
    g = np.meshgrid(*(np.arange(n) - c[a] for a, n in enumerate(shape)), indexing="ij")
    r = min(shape) / 2 - 3
    mask = (g[0] ** 2 + (1.1 * g[1]) ** 2 + (0.9 * g[2]) ** 2) <= r * r
    if hole:
        mask[:, :, shape[2] // 2] = False
    f = np.zeros(shape)
    for _ in range(6):
        p = rng.uniform(-r / 2, r / 2, 3)
        f += rng.uniform(-0.1, 0.1) * np.exp(-sum((g[a] - p[a]) ** 2 for a in range(3)) / 6.0)
    f += bg * (0.01 * (g[0] ** 2 - g[1] ** 2) + 0.02 * g[2]) + 0.004 * rng.standard_normal(shape)
    return np.asfortranarray(f), np.asfortranarray(mask)


CASES = [((40, 36, 30), (1.0, 1.0, 1.0), 5, False), ((37, 33, 28), (0.9, 1.0, 1.2), 4, False),
         ((34, 34, 32), (1.0, 1.0, 1.0), 3, True)]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_sharp_matches_oracle(cuda, case, dt):
    shape, vsz, r, hole = CASES[case]
    f, m = _case(shape, case, hole)
    f = f.astype(dt)
    wf, wm = obg.sharp(f, m, vsz, r)
    gf, gm = gbg.sharp(f, m, vsz, r)
    np.testing.assert_array_equal(gm, wm)
    assert gf.dtype == f.dtype
    assert rel_l2(gf, wf) < (1e-10 if dt == np.float64 else 1e-4)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("rs", [None, [6, 4, 2]])
def test_vsharp_matches_oracle(cuda, dt, rs):
    shape, vsz, _, _ = CASES[0]
    f, m = _case(shape, 9)
    f = f.astype(dt)
    if rs is None:
        rs_used = gbg.default_radii(vsz)
        assert list(rs_used) == [9.0, 7.0, 5.0, 3.0]
    wf, wm = obg.vsharp(f, m, vsz, rs)
    gf, gm = gbg.vsharp(f, m, vsz, rs)
    np.testing.assert_array_equal(gm, wm)
    assert rel_l2(gf, wf) < (1e-10 if dt == np.float64 else 1e-4)


def test_sharp_multi_echo_and_harmonic_background(cuda):
    shape, vsz, r, _ = CASES[1]
    f, m = _case(shape, 2)
    c = [(n - 1) / 2 for n in shape]
    g = np.meshgrid(*(np.arange(n) - c[a] for a, n in enumerate(shape)), indexing="ij")
    b = 0.02 * (g[0] ** 2 - g[2] ** 2) + 0.01 * g[0] * g[1] + 0.03 * g[2]  # harmonic
    f4 = np.stack([f, f + b], axis=3)
    fl, mo = gbg.sharp(f4, m, (1.0, 1.0, 1.0), r)
    for t in range(2):
        one, _ = gbg.sharp(f4[..., t], m, (1.0, 1.0, 1.0), r)
        np.testing.assert_array_equal(fl[..., t], one)
    # the harmonic part is removed: the two echoes agree to rounding
    assert np.abs(fl[..., 1] - fl[..., 0]).max() < 1e-12 * np.abs(b).max()


def test_vsharp_rejects_a_single_radius(cuda):
    f, m = _case((20, 20, 20), 1)
    with pytest.raises(QsmError, match="two radii"):
        gbg.vsharp(f, m, (1, 1, 1), [3])


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_vsharp_spectrum_reuse_is_bitwise(cuda, dt):
    """One echo runs every radius from the stored spectra of f and of the
    mask (op + inverse passes only); several echoes transform each echo per
    radius.  Both give the same bits: each echo of a 4-D call equals the
    same echo alone, and the eroded mask too."""
    shape, vsz, _, _ = CASES[1]
    f, m = _case(shape, 12)
    f4 = np.stack([f, 0.5 * f[::-1, :, :] + 0.01], axis=3).astype(dt)
    fl, mo = gbg.vsharp(f4, m, vsz, [7, 5, 4, 2])
    for t in range(2):
        one, mone = gbg.vsharp(np.asfortranarray(f4[..., t]), m, vsz, [7, 5, 4, 2])
        np.testing.assert_array_equal(fl[..., t], one)
        np.testing.assert_array_equal(mo, mone)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_vsharp_register_pass_lengths(cuda, dt):
    """A 20x168x168 mask box and radii up to 12 mm pad y and z to 192, a
    length the register-resident FFT passes take: the oracle to the
    tolerance, and one echo (stored spectra, op + inverse passes) bitwise
    each echo of a 4-D call (every transform per radius).  The harmonic
    background is scaled to the grid (0.01 (x^2 - y^2) would reach 74 at
    172 voxels, 700x the local field, and FP32's high-pass cancellation
    then measures conditioning, not the transforms)."""
    shape, vsz, rs = (24, 172, 172), (1.0, 1.0, 1.0), [12, 8, 4]
    f, _ = _case(shape, 21, bg=0.05)
    m = np.zeros(shape, bool)
    m[2:22, 2:170, 2:170] = True
    m = np.asfortranarray(m)
    f = f.astype(dt)
    wf, wm = obg.vsharp(f, m, vsz, rs)
    gf, gm = gbg.vsharp(f, m, vsz, rs)
    np.testing.assert_array_equal(gm, wm)
    assert rel_l2(gf, wf) < (1e-10 if dt == np.float64 else 1e-4)
    f4 = np.stack([f, 0.5 * f[::-1, :, :] + 0.01], axis=3).astype(dt)
    fl, mo = gbg.vsharp(f4, m, vsz, rs)
    np.testing.assert_array_equal(fl[..., 0], gf)
    np.testing.assert_array_equal(mo, gm)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_empty_mask_gives_zeros(cuda, dt):
    """An empty mask: no slice to crop, the local field and the eroded mask
    are all zeros (sharp.m:51-56 crops to nothing), one echo or two."""
    shape = (20, 16, 12)
    f = np.random.default_rng(4).standard_normal(shape + (2,)).astype(dt)
    m = np.zeros(shape, bool)
    for fn in (lambda a: gbg.sharp(a, m, (1, 1, 1)), lambda a: gbg.vsharp(a, m, (1, 1, 1))):
        for a in (f[..., 0], f):
            fl, mo = fn(np.asfortranarray(a))
            assert fl.shape == a.shape and fl.dtype == a.dtype and not fl.any()
            assert mo.shape == shape and not mo.any()
