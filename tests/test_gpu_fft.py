"""GPU FFT and FFT dipole convolution (fft3.m / ifft3.m, dipoleKernel.m,
padfastfft.m) against numpy's pocketfft (MATLAB's FFT is closed: parity
with it is unpinned; pocketfft and closed forms pin the transform) and the
harness's convolution."""
import numpy as np
import pytest

from qsm_amd import dipole, fastfft, phantom

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(8, 9, 10), (28, 30, 14), (45, 7, 12), (64, 1, 3), (49, 25, 16), (2, 3, 5),
                                   (11, 13, 22), (176, 26, 33), (143, 2, 1)])
@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_fft3_matches_pocketfft(cuda, shape, dt):
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dt)
    tol = 1e-13 if dt == np.complex128 else 2e-6
    X = dipole.fft3(x)
    ref = np.fft.fftn(x.astype(np.complex128))
    assert np.linalg.norm(X - ref) / np.linalg.norm(ref) < tol
    xi = dipole.ifft3(X)
    assert np.linalg.norm(xi - x) / np.linalg.norm(x) < tol


def test_fft3_impulse_and_constant(cuda):
    x = np.zeros((12, 10, 14), complex)
    x[0, 0, 0] = 1.0
    np.testing.assert_allclose(dipole.fft3(x), np.ones(x.shape), atol=1e-15)
    X = np.full(x.shape, 3.0 + 0j)
    y = dipole.ifft3(X)
    want = np.zeros(x.shape, complex)
    want[0, 0, 0] = 3.0
    np.testing.assert_allclose(y, want, atol=1e-14)


def test_fft3_parseval_radix7(cuda):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((49, 14, 7)) + 0j
    X = dipole.fft3(x)
    assert abs((np.abs(X) ** 2).sum() / x.size - (np.abs(x) ** 2).sum()) < 1e-10 * (np.abs(x) ** 2).sum()


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("shape,vsz,bdir,pad", [
    ((24, 22, 20), (1.0, 1.0, 1.0), (0, 0, 1), (12, 11, 10)),
    ((30, 17, 21), (0.9, 0.9, 1.2), (0, -0.258819, 0.965926), (15, 8, 10)),
    ((16, 16, 16), (1.0, 1.0, 1.0), (1, 1, 1), (0, 0, 0)),
    # odd padded x length, odd row count (the last row of the x passes has
    # no partner), Nyquist planes on y and z with a tilted B0
    ((15, 9, 7), (1.0, 1.1, 0.9), (0.3, -0.2, 0.93), (0, 2, 3)),
    # a single x column (half spectrum of length 1)
    ((1, 12, 10), (1.0, 1.0, 1.0), (0.2, 0.1, 0.97), (0, 3, 2)),
])
def test_dipole_conv_matches_fft_reference(cuda, dt, shape, vsz, bdir, pad):
    rng = np.random.default_rng(2)
    chi = np.asfortranarray(rng.standard_normal(shape).astype(dt))
    got = dipole.dipole_conv(chi, vsz, bdir, pad)
    up = fastfft.padfastfft(chi.astype(np.float64), pad)
    D = fastfft.dipole_kernel(up.shape, vsz, bdir)
    ref = np.real(np.fft.ifftn(D * np.fft.fftn(up)))[fastfft.unpad_slices(up.shape, shape, pad)]
    tol = 1e-12 if dt == np.float64 else 1e-5
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("L", [176, 192, 256, 320, 360, 384, 448, 512, 640])
def test_dipole_conv_register_pass_lengths(cuda, dt, L):
    """y and z lengths that run the register-resident four-step passes
    (fft_reg.h: 176 = 16x11 unpadded, the others padded 2x): the y forward,
    the fused z forward / multiplier / inverse and the y inverse against
    numpy's FFT.  A tilted B0 puts D(k) != D(-k) on the Nyquist planes; the
    half spectrum's 21 x positions are not a multiple of the pencils per
    block, so blocks straddle rows."""
    rng = np.random.default_rng(L)
    n, p = (L, 0) if L == 176 else (L // 2, L // 4)
    shape, pad = (40, n, n - 2 * (L % 3 == 0)), (0, p, p + (L % 3 == 0))
    M, _ = dipole.pad_size(shape, pad)
    assert M[1:] == (L, L)
    chi = np.asfortranarray(rng.standard_normal(shape).astype(dt))
    vsz, bdir = (1.0, 0.9, 1.1), (0.3, -0.2, 0.93)
    got = dipole.dipole_conv(chi, vsz, bdir, pad)
    up = fastfft.padfastfft(chi.astype(np.float64), pad)
    D = fastfft.dipole_kernel(up.shape, vsz, bdir)
    ref = np.real(np.fft.ifftn(D * np.fft.fftn(up)))[fastfft.unpad_slices(up.shape, shape, pad)]
    tol = 1e-12 if dt == np.float64 else 1e-5
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("L", [176, 192, 256, 320, 360, 384, 448, 512, 640])
def test_dipole_conv_register_x_pass_lengths(cuda, dt, L):
    """x lengths that run the register-resident x passes (two real rows per
    pencil, the half-spectrum split and its rebuild): 7 x 5 = 35 rows, so
    the last pencil carries one row, and an odd leading pad."""
    rng = np.random.default_rng(L + 1)
    n, p = (L, 0) if L == 176 else (L // 2 - 2, L // 4 + 1)
    shape, pad = (n, 7, 5), (p, 0, 0)
    M, _ = dipole.pad_size(shape, pad)
    assert M[0] == L
    chi = np.asfortranarray(rng.standard_normal(shape).astype(dt))
    vsz, bdir = (0.9, 1.0, 1.1), (0.4, -0.1, 0.91)
    got = dipole.dipole_conv(chi, vsz, bdir, pad)
    up = fastfft.padfastfft(chi.astype(np.float64), pad)
    D = fastfft.dipole_kernel(up.shape, vsz, bdir)
    ref = np.real(np.fft.ifftn(D * np.fft.fftn(up)))[fastfft.unpad_slices(up.shape, shape, pad)]
    tol = 1e-12 if dt == np.float64 else 1e-5
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol


def test_dipole_conv_pad_sizes_match_padfastfft(cuda):
    for shape, pad in (((24, 22, 20), (12, 11, 10)), ((7, 9, 11), (3, 0, 5)), ((176, 256, 5), (88, 128, 0))):
        M, off = dipole.pad_size(shape, pad)
        up = fastfft.padfastfft(np.ones(shape), pad)
        assert M == up.shape
        sl = fastfft.unpad_slices(up.shape, shape, pad)
        assert off == tuple(s.start for s in sl)


def test_harness_field_on_gpu_equals_host(cuda):
    spec = phantom.PhantomSpec(shape=(32, 30, 28), n_ellipsoids=6, semi_axes=(2, 6), mask="sphere",
                               mask_param=(12.0,), noise=0.0, seed=5, bdir=(0.1, 0.0, 0.99))
    ph = phantom.make(spec)
    got = dipole.dipole_conv(ph["chi"], spec.vsz, spec.bdir, [s // 2 for s in spec.shape])
    assert np.linalg.norm(got - ph["field"]) / np.linalg.norm(ph["field"]) < 1e-12


@pytest.mark.parametrize("dt", ["float64", "float32"])
def test_dipole_conv_device_form_on_side_stream(cuda, dt):
    """qsm_dipole_conv_dev is asynchronous on its stream; twice in a row on
    a side stream (cached plans, work space reused uncleared) it matches the
    host form bitwise."""
    import torch

    rng = np.random.default_rng(9)
    shape, vsz, bdir, pad = (26, 20, 18), (0.9, 1.0, 1.1), (0.1, 0.2, 0.95), (13, 10, 9)
    chi = np.asfortranarray(rng.standard_normal(shape).astype(dt))
    want = dipole.dipole_conv(chi, vsz, bdir, pad)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        c = torch.from_numpy(np.ascontiguousarray(chi.reshape(-1, order="F"))).to("cuda")
        c = c.view(shape[::-1]).permute(2, 1, 0)
        out = torch.empty_like(c)
        M, _ = dipole.pad_size(shape, pad)
        work = torch.full((int(np.prod(M)) * 2 * c.element_size(),), 255, dtype=torch.uint8,
                          device="cuda")
    for _ in range(2):
        dipole.dipole_conv_dev(c, vsz, bdir, pad, out, work=work, stream=st)
        st.synchronize()
        got = out.permute(2, 1, 0).contiguous().cpu().numpy().reshape(-1)
        got = got.reshape(shape, order="F")
        assert got.dtype == want.dtype
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dt,n", [(np.complex128, 2048), (np.complex64, 4096), (np.complex128, 1960)])
def test_fft3_longest_lengths(cuda, dt, n):
    """The longest supported transform lengths (one pencil per block, more
    than 64 KB of LDS) against pocketfft."""
    rng = np.random.default_rng(12)
    x = (rng.standard_normal((n, 2, 3)) + 1j * rng.standard_normal((n, 2, 3))).astype(dt)
    tol = 1e-13 if dt == np.complex128 else 3e-6
    X = dipole.fft3(x)
    ref = np.fft.fftn(x.astype(np.complex128))
    assert np.linalg.norm(X - ref) / np.linalg.norm(ref) < tol
