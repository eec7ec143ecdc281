"""Pins of the sharp.m / vsharp.m restatement (oracle/bgremove.py) by what
the mathematics fixes:

* the spherical mean value property: a harmonic background (here degree-2
  harmonic polynomials, whose mean over a centrally symmetric ball is their
  centre value) is removed to rounding, and the methods are linear in f;
* the eroded mask equals a brute-force erosion by the ball (for balls of
  fewer than 1000 samples the 0.999 threshold means the whole ball inside);
* nextfastfft against its definition, padfastfft/unpadfastfft round trips
  and their odd-size rule, cropIndices on a mask with a gap.
"""
import numpy as np
import pytest

from oracle import bgremove as bg


def _phantom(shape, seed):
    rng = np.random.default_rng(seed)
    c = [(n - 1) / 2 for n in shape]
    g = np.meshgrid(*(np.arange(n) - c[a] for a, n in enumerate(shape)), indexing="ij")
    r = min(shape) / 2 - 3
    mask = (g[0] ** 2 + g[1] ** 2 + g[2] ** 2) <= r * r
    f = np.zeros(shape)
    for _ in range(5):
        p = rng.uniform(-r / 2, r / 2, 3)
        f += rng.uniform(-0.1, 0.1) * np.exp(-sum((g[a] - p[a]) ** 2 for a in range(3)) / 8.0)
    harm = 0.01 * (g[0] ** 2 - g[1] ** 2) + 0.003 * g[0] * g[2] + 0.02 * g[1] - 0.005 * g[2] + 0.3
    return np.asfortranarray(f), np.asfortranarray(mask), np.asfortranarray(harm)


@pytest.mark.parametrize("vsz", [(1.0, 1.0, 1.0), (1.0, 1.0, 1.5)])
def test_sharp_removes_harmonic_background_and_is_linear(vsz):
    f, m, b = _phantom((36, 34, 30), 1)
    fl, _ = bg.sharp(f + b, m, vsz, 4)
    f0, _ = bg.sharp(f, m, vsz, 4)
    fb, _ = bg.sharp(b, m, vsz, 4)
    assert np.abs(fb).max() < 1e-12 * np.abs(b).max()
    assert np.abs(fl - f0 - fb).max() < 1e-13
    assert np.abs(f0).max() > 1e-3


def test_vsharp_removes_harmonic_background():
    f, m, b = _phantom((36, 34, 30), 2)
    fl, _ = bg.vsharp(f + b, m, (1, 1, 1), [6, 4, 2])
    f0, _ = bg.vsharp(f, m, (1, 1, 1), [6, 4, 2])
    assert np.abs(fl - f0).max() < 1e-12 * np.abs(b).max()


def _erode(mask, r, vsz):
    """Brute force: voxels whose whole ball (smvKernel grid) lies in mask."""
    offs = []
    R = [int(np.ceil(r / h)) for h in vsz]
    for i in range(-R[0], R[0] + 1):
        for j in range(-R[1], R[1] + 1):
            for k in range(-R[2], R[2] + 1):
                X, Y, Z = i * vsz[0], j * vsz[1], k * vsz[2]
                if Y * Y + X * X + Z * Z <= r * r:
                    offs.append((i, j, k))
    out = np.zeros_like(mask)
    P = np.pad(mask, [(R[0], R[0]), (R[1], R[1]), (R[2], R[2])])
    n = mask.shape
    acc = np.ones(n, bool)
    for (i, j, k) in offs:
        acc &= P[R[0] + i:R[0] + i + n[0], R[1] + j:R[1] + j + n[1], R[2] + k:R[2] + k + n[2]]
    out[...] = acc & mask
    return out, len(offs)


@pytest.mark.parametrize("r,vsz", [(3, (1.0, 1.0, 1.0)), (4, (1.0, 1.2, 0.9))])
def test_eroded_mask_is_ball_erosion(r, vsz):
    _, m, _ = _phantom((30, 28, 26), 3)
    m[10:14, 5:9, 12:15] = False  # a hole
    _, mo = bg.sharp(np.zeros(m.shape), m, vsz, r)
    want, K = _erode(m, r, vsz)
    assert K < 1000
    np.testing.assert_array_equal(mo, want)


def test_nextfastfft_and_padding():
    for n in range(1, 600):
        k = bg.nextfastfft(n)
        r = k
        for p in (2, 3, 5, 7):
            while r % p == 0:
                r //= p
        assert r == 1 and k >= n
        assert all(bg.nextfastfft(j) != j for j in range(n, k))
    # odd size, odd pad: the extra sample goes before (padfastfft.m:47-56)
    u = np.arange(7 * 4 * 3, dtype=float).reshape((7, 4, 3), order="F")
    p = bg.padfastfft(u, [2, 1, 0])
    assert p.shape == (bg.nextfastfft(11), bg.nextfastfft(6), 3) == (12, 6, 3)
    assert p[3, 1, 0] == u[0, 0, 0]          # pre = 5//2 + 1 = 3 on the odd axis
    np.testing.assert_array_equal(bg.unpadfastfft(p, u.shape, [2, 1, 0]), u)


def test_crop_indices_skip_empty_slices():
    m = np.zeros((6, 5, 4), bool)
    m[1, 2, 1] = m[4, 3, 2] = True
    ix, iy, iz = bg.crop_indices(m)
    assert list(ix) == [1, 4] and list(iy) == [2, 3] and list(iz) == [1, 2]
