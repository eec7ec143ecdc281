"""Host-side logic: nextfastfft / padfastfft / unpadfastfft / cropIndices /
dipole kernel (src/utils/padfastfft.m, unpadfastfft.m,
third_party/nextfastfft.m, src/utils/cropIndices.m,
src/utils/kernels/dipoleKernel.m) and the phantom harness.

The MATLAB reference cannot run here; these are pinned by brute force
(7-smooth search), by the formulas' special values and symmetries, and by
round trips.
"""
import numpy as np
import pytest

from qsm_amd import fastfft, phantom


def _smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def test_nextfastfft_brute_force():
    want = {}
    nxt = None
    for n in range(4200, 0, -1):
        if _smooth7(n):
            nxt = n
        want[n] = nxt
    got = fastfft.nextfastfft(np.arange(1, 4097))
    np.testing.assert_array_equal(got, [want[n] for n in range(1, 4097)])


def test_nextfastfft_config_sizes_are_7_smooth():
    # sizes the dipole convolution pads to at the benchmark shapes
    for n, m in ((64 * 2, 128), (224 * 2, 448), (128 * 2, 256), (176 * 2, 360), (352, 360),
                 (63, 63), (211, 216), (1023, 1024)):
        assert fastfft.nextfastfft(n) == m
    assert fastfft.nextfastfft(0) == 1


def test_nextfastfft_large_reduction():
    # above 2^20 the input is halved until in range, looked up, scaled back
    n = 2 ** 20 + 1
    r = fastfft.nextfastfft(n)
    assert r % 2 == 0 and r >= n and _smooth7(r)
    assert fastfft.nextfastfft(2 ** 21) == 2 ** 21


@pytest.mark.parametrize("shape,pads", [((7, 8, 9), (3, 4, 5)), ((10, 11, 1), (5, 0, 0)),
                                        ((6, 6, 6), (0, 0, 0)), ((13, 2, 5), (7, 1, 2))])
def test_padfastfft_roundtrip(shape, pads):
    rng = np.random.default_rng(0)
    u = rng.standard_normal(shape)
    up = fastfft.padfastfft(u, pads)
    for d, (s, p) in enumerate(zip(shape, pads)):
        want = s if p == 0 else fastfft.nextfastfft(s + 2 * p)
        assert up.shape[d] == want
    np.testing.assert_array_equal(fastfft.unpadfastfft(up, shape, pads), u)
    # zeros outside the copy
    mask = np.zeros(up.shape, bool)
    mask[fastfft.unpad_slices(up.shape, shape, pads)] = True
    assert np.all(up[~mask] == 0)


@pytest.mark.parametrize("s,p,M,pre", [
    (5, 1, 7, 1), (5, 2, 9, 2), (6, 3, 12, 3), (4, 9, 24, 10),
    # odd total pad: the extra voxel goes before an odd size, after an even one
    (5, 3, 12, 4), (5, 9, 24, 10), (11, 1, 14, 2), (9, 2, 14, 3), (24, 1, 27, 1),
])
def test_padfastfft_split_rule(s, p, M, pre):
    """padfastfft.m:34-51 ('both'): total = nextfastfft(s+2p) - s, split in
    halves, an odd remainder before odd sizes and after even ones."""
    up = fastfft.padfastfft(np.ones((s, 1, 1)), (p, 0, 0))
    assert up.shape[0] == M
    nz = np.nonzero(up[:, 0, 0])[0]
    assert nz[0] == pre and nz[-1] == pre + s - 1


@pytest.mark.parametrize("n", [8, 9, 10, 15, 16])
def test_kgrid_ifftshift_layout(n):
    h = 0.7
    g = fastfft.kgrid(n, h)
    # ifftshift of the centred grid (-floor(n/2) .. ceil(n/2)-1)*dk
    low = n // 2
    centred = (np.arange(n) - low) / (n * h)
    np.testing.assert_allclose(g, np.fft.ifftshift(centred), rtol=1e-15, atol=1e-18)
    np.testing.assert_allclose(g, np.fft.fftfreq(n, d=h), rtol=1e-12, atol=1e-15)


def test_dipole_kernel_special_values():
    sz, vsz = (16, 15, 14), (1.0, 0.9, 1.2)
    D = fastfft.dipole_kernel(sz, vsz, (0, 0, 1))
    assert D[0, 0, 0] == 0.0
    assert D.min() >= -2 / 3 - 1e-15 and D.max() <= 1 / 3 + 1e-15
    # on the B0 axis: -2/3; in the perpendicular plane: 1/3
    np.testing.assert_allclose(D[0, 0, 1:], -2 / 3, rtol=1e-14)
    np.testing.assert_allclose(D[1:, 1:, 0], 1 / 3, rtol=1e-14)
    # even symmetry D(k) = D(-k) on the FFT grid
    Dm = D[(-np.arange(16)) % 16][:, (-np.arange(15)) % 15][:, :, (-np.arange(14)) % 14]
    np.testing.assert_allclose(D, Dm, rtol=1e-14, atol=1e-16)


def test_dipole_kernel_tilted_and_normalised():
    b = np.array([0.0, -np.sin(np.radians(15)), np.cos(np.radians(15))])
    D1 = fastfft.dipole_kernel((12, 12, 12), 1.0, b)
    D2 = fastfft.dipole_kernel((12, 12, 12), 1.0, 3.0 * b)  # renormalised
    np.testing.assert_allclose(D1, D2, rtol=1e-14, atol=1e-15)
    # trace identity: sum over the three axis-aligned b of D = 3*(1/3) - 1 = 0
    Ds = sum(fastfft.dipole_kernel((10, 11, 12), 1.0, e) for e in np.eye(3))
    Ds[0, 0, 0] = 0.0
    np.testing.assert_allclose(Ds, 0.0, atol=1e-14)


def test_crop_indices():
    m = np.zeros((9, 8, 7), bool)
    m[2:5, 1, 3:7] = True
    ix, iy, iz = fastfft.crop_indices(m)
    np.testing.assert_array_equal(ix, [2, 3, 4])
    np.testing.assert_array_equal(iy, [1])
    np.testing.assert_array_equal(iz, [3, 4, 5, 6])


def test_phantom_is_seeded_and_field_is_a_dipole_convolution():
    spec = phantom.PhantomSpec(shape=(24, 22, 20), n_ellipsoids=5, semi_axes=(2, 5),
                               mask="sphere", mask_param=(9.0,), noise=0.0, seed=3)
    a = phantom.make(spec)
    b = phantom.make(spec)
    np.testing.assert_array_equal(a["chi"], b["chi"])
    np.testing.assert_array_equal(a["field"], b["field"])
    # brute force: full complex FFT convolution on the padded grid
    up = fastfft.padfastfft(a["chi"], [s // 2 for s in spec.shape])
    D = fastfft.dipole_kernel(up.shape, spec.vsz, spec.bdir)
    f = np.real(np.fft.ifftn(D * np.fft.fftn(up)))
    f = f[fastfft.unpad_slices(up.shape, spec.shape, [s // 2 for s in spec.shape])]
    np.testing.assert_allclose(a["field"], f, atol=1e-12)


def test_harness_field_is_the_toolbox_idiom_for_tilted_b0():
    """phantom.dipole_field = real(ifft3(D .* fft3(padfastfft(chi)))) cropped:
    with a B0 that has two non-zero components D is not symmetric under
    k -> -k on the Nyquist planes of even axes, so a real-to-complex
    shortcut would differ (by ~1e-2 here)."""
    from qsm_amd import fastfft, phantom

    spec = phantom.PhantomSpec(shape=(20, 18, 16), n_ellipsoids=5, semi_axes=(2, 5), mask="sphere",
                               mask_param=(8.0,), noise=0.0, seed=4, bdir=(0.1, 0.0, 0.99))
    ph = phantom.make(spec)
    pad = [s // 2 for s in spec.shape]
    up = fastfft.padfastfft(ph["chi"], pad)
    D = fastfft.dipole_kernel(up.shape, spec.vsz, spec.bdir)
    ref = np.real(np.fft.ifftn(D * np.fft.fftn(up)))[fastfft.unpad_slices(up.shape, spec.shape, pad)]
    assert np.linalg.norm(ph["field"] - ref) <= 1e-13 * np.linalg.norm(ref)


def test_sstgv_rejects_bad_arguments_before_any_device_work():
    """sstgv.m:92-136 checks, restated: every rejection happens on the host."""
    from qsm_amd.sstgv import sstgv

    ph = np.zeros((6, 6, 6))
    m = np.ones((6, 6, 6), bool)
    with pytest.raises(TypeError):
        sstgv(ph.astype(np.int32), m, (1, 1, 1))
    with pytest.raises(ValueError, match="mask"):
        sstgv(ph, np.ones((6, 6, 5)), (1, 1, 1))
    with pytest.raises(ValueError, match="binary"):
        sstgv(ph, 2 * np.ones((6, 6, 6)), (1, 1, 1))
    with pytest.raises(ValueError, match="vsz"):
        sstgv(ph, m, (1, 1, 0))
    with pytest.raises(ValueError, match="lambda"):
        sstgv(ph, m, (1, 1, 1), lam=-1.0)
    with pytest.raises(ValueError, match="maxit"):
        sstgv(ph, m, (1, 1, 1), maxit=0)
    with pytest.raises(ValueError, match="finite"):
        bad = ph.copy()
        bad[0, 0, 0] = np.nan
        sstgv(bad, m, (1, 1, 1))
    # an empty mask returns zeros without touching the device
    out = sstgv(ph, np.zeros((6, 6, 6), bool), (1, 1, 1))
    assert out.shape == ph.shape and not out.any()


def test_mask_argument_is_a_zero_copy_view_when_it_can_be():
    """Mask arguments reach the library as column-major uint8 0/1: a
    column-major bool mask by a view of its memory, anything else by one
    converted copy (row-major, or nonzero values other than 1)."""
    from qsm_amd._lib import mask_u8

    m = np.asfortranarray(np.random.default_rng(0).random((5, 4, 3)) < 0.5)
    v = mask_u8(m)
    assert v.dtype == np.uint8 and v.flags.f_contiguous and np.shares_memory(v, m)
    w = mask_u8(np.ascontiguousarray(m))
    assert w.flags.f_contiguous and np.array_equal(w, m.astype(np.uint8))
    z = mask_u8((m * 255).astype(np.int32))
    assert z.dtype == np.uint8 and np.array_equal(z, m.astype(np.uint8))
