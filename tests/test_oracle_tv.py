"""Pins of the tv.m / rts.m restatement (oracle/tv_admm.py) by what the mathematics
fixes, independent of the oracle's own construction:

* the u-subproblem (tv.m:196-203: an FFT-diagonalised solve) equals the
  minimum-norm solution of its normal equations (mu D'D + rho G'G) u =
  mu D' f + G'(rho v - l), assembled from explicit DFT and periodic
  difference matrices on a tiny grid (B0 along an axis, where D is even
  and the solve is real);
* the Laplacian symbol built as the reference builds it (padded, shifted,
  transformed stencil, tv.m:163-165) equals the closed form
  sum_a (2 - 2 cos(2 pi m_a / n_a)) / h_a^2 -- the form the GPU solve
  generates in-kernel;
* soft thresholding (tv.m:226-242) on hand-worked cases;
* dipoleKernel.m special values: D(0) = 0, -2/3 along B0, 1/3 across;
* a zero field gives x = 0 and stops at the first iteration;
* rts.m: the LSMR of step 1 (third_party/lsmr/lsmr.m) against scipy's
  implementation of the same published algorithm, and step 2's
  x-subproblem against its normal equations with M = mu [|D| > delta].
"""
import numpy as np
import pytest

from oracle import tv_admm as tva


def _dft(n):
    j = np.arange(n)
    return np.exp(-2j * np.pi * np.outer(j, j) / n)


def _fdiff(n, h):
    """Periodic forward difference (u[i+1] - u[i]) / h as a matrix."""
    A = -np.eye(n)
    A[np.arange(n), (np.arange(n) + 1) % n] += 1.0
    return A / h


def test_u_step_is_the_normal_equation_solution():
    rng = np.random.default_rng(3)
    sz, vsz = (4, 3, 5), (0.9, 1.1, 1.3)
    nx, ny, nz = sz
    N = nx * ny * nz
    mu, rho = 7.5e3, 9.0
    D = tva.dipole_kernel_k(sz, vsz, (0, 0, 1), np.float64)
    f = np.asfortranarray(rng.standard_normal(sz))
    v = [np.asfortranarray(rng.standard_normal(sz)) for _ in range(3)]
    lm = [np.asfortranarray(rng.standard_normal(sz)) for _ in range(3)]
    # oracle step
    F = (mu * np.conj(D)) * tva.fft3(f, np.float64)
    iA = tva.inv_a(mu * np.abs(D) ** 2, tva.laplace_symbol(sz, vsz), rho)
    u = tva.u_step(F, iA, v, lm, rho, vsz, np.float64)
    # dense: x fastest (Fortran vec) -> kron(z, y, x)
    W = np.kron(_dft(nz), np.kron(_dft(ny), _dft(nx)))
    Dop = np.real(np.linalg.solve(W, np.diag(D.reshape(-1, order="F")) @ W))
    Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
    Gx = np.kron(Iz, np.kron(Iy, _fdiff(nx, vsz[0])))
    Gy = np.kron(Iz, np.kron(_fdiff(ny, vsz[1]), Ix))
    Gz = np.kron(_fdiff(nz, vsz[2]), np.kron(Iy, Ix))
    G = np.vstack([Gx, Gy, Gz])
    A = mu * Dop.T @ Dop + rho * G.T @ G
    vec = lambda a: a.reshape(-1, order="F")
    rhs = mu * Dop.T @ vec(f) + G.T @ np.concatenate([rho * vec(v[c]) - vec(lm[c]) for c in range(3)])
    want = np.linalg.lstsq(A, rhs, rcond=1e-12)[0]
    assert A.shape == (N, N)
    np.testing.assert_allclose(vec(u), want, rtol=0, atol=1e-10 * np.abs(want).max())


@pytest.mark.parametrize("sz,vsz", [((6, 5, 7), (1.0, 1.0, 1.0)), ((8, 9, 4), (0.9, 1.2, 0.7))])
def test_laplace_symbol_closed_form(sz, vsz):
    L = tva.laplace_symbol(sz, vsz)
    m = np.meshgrid(*(np.arange(n) for n in sz), indexing="ij")
    want = sum((2 - 2 * np.cos(2 * np.pi * m[a] / sz[a])) / vsz[a] ** 2 for a in range(3))
    np.testing.assert_allclose(L, want, rtol=0, atol=1e-12)
    assert L[0, 0, 0] == pytest.approx(0.0, abs=1e-12) and (L.reshape(-1)[1:] > 0).all()


def test_shrink_closed_forms():
    T = np.float64
    r = np.array([3.0, -0.5, 0.2, -2.0, 0.0])
    z = np.zeros_like(r)
    a1, _, _ = tva.shrink(r, z, z, 0.5, 1, T)
    np.testing.assert_array_equal(a1, [2.5, -0.0, 0.0, -1.5, 0.0])
    # isotropic: |(3, 4, 0)| = 5 -> scaled by (5 - 1)/5; |(0.3, 0.4, 0)| = 0.5 <= 1 -> 0
    x, y, w = tva.shrink(np.array([3.0, 0.3, 0.0]), np.array([4.0, 0.4, 0.0]), np.zeros(3), 1.0, 2, T)
    np.testing.assert_allclose(x, [2.4, 0.0, 0.0])
    np.testing.assert_allclose(y, [3.2, 0.0, 0.0])
    np.testing.assert_array_equal(w, [0.0, 0.0, 0.0])


def test_dipole_kernel_special_values():
    sz, vsz = (8, 6, 10), (1.0, 1.0, 1.0)
    D = tva.dipole_kernel_k(sz, vsz, (0, 0, 1), np.float64)
    assert D[0, 0, 0] == 0.0
    np.testing.assert_allclose(D[0, 0, 1:], -2.0 / 3.0)   # k along B0
    np.testing.assert_allclose(D[1:, 1:, 0], 1.0 / 3.0)   # k across B0
    assert D.min() >= -2.0 / 3.0 - 1e-15 and D.max() <= 1.0 / 3.0 + 1e-15
    # B0 along x: the same kernel with the axes exchanged
    Dx = tva.dipole_kernel_k((10, 6, 8), vsz, (1, 0, 0), np.float64)
    np.testing.assert_allclose(Dx, np.transpose(D, (2, 1, 0)), atol=1e-15)


def test_zero_field_stops_at_once():
    sz = (8, 7, 6)
    x, tr = tva.tv(np.zeros(sz), np.ones(sz), (1, 1, 1))
    assert np.all(x == 0) and len(tr) == 1 and tr[0][0] == 0.0 and tr[0][2] == 0.0


def test_lsmr_matches_scipy():
    """The LSMR restatement (third_party/lsmr/lsmr.m) against scipy's
    implementation of the same published algorithm (Fong and Saunders),
    diagonal complex operator, a few iterations and to convergence."""
    from scipy.sparse.linalg import LinearOperator
    from scipy.sparse.linalg import lsmr as sp_lsmr

    rng = np.random.default_rng(5)
    n = 300
    d = rng.uniform(-2.0 / 3.0, 1.0 / 3.0, n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    A = LinearOperator((n, n), matvec=lambda x: d * x, rmatvec=lambda x: d * x, dtype=complex)
    for it in (1, 2, 4, 12):
        x, itn, _ = tva.lsmr(d, b, 1e-14, 1e-14, it)
        want = sp_lsmr(A, b, damp=0.0, atol=1e-14, btol=1e-14, conlim=1e8, maxiter=it)[0]
        assert itn == it
        np.testing.assert_allclose(x, want, rtol=1e-10, atol=1e-12 * np.abs(want).max())


def test_rts_step2_u_step_is_the_normal_equation_solution():
    """rts.m's x-subproblem: (M + rho G'G) x = M x1 + G'(rho y - p) with
    M = mu [|D| > delta] as a Fourier multiplier (B0 along an axis)."""
    rng = np.random.default_rng(8)
    sz, vsz = (4, 5, 3), (1.0, 0.9, 1.2)
    nx, ny, nz = sz
    mu, rho, delta = 1e5, 10.0, 0.15
    D = tva.dipole_kernel_k(sz, vsz, (0, 0, 1), np.float64)
    x1 = np.asfortranarray(rng.standard_normal(sz))
    y = [np.asfortranarray(rng.standard_normal(sz)) for _ in range(3)]
    p = [np.asfortranarray(rng.standard_normal(sz)) for _ in range(3)]
    Mk = mu * (np.abs(D) > delta).astype(np.float64)
    F = Mk * tva.fft3(x1, np.float64)
    iA = tva.inv_a(Mk, tva.laplace_symbol(sz, vsz), rho)
    x = tva.u_step(F, iA, y, p, rho, vsz, np.float64)
    W = np.kron(_dft(nz), np.kron(_dft(ny), _dft(nx)))
    vec = lambda a: a.reshape(-1, order="F")
    Mop = np.real(np.linalg.solve(W, np.diag(vec(Mk)) @ W))
    Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
    G = np.vstack([np.kron(Iz, np.kron(Iy, _fdiff(nx, vsz[0]))),
                   np.kron(Iz, np.kron(_fdiff(ny, vsz[1]), Ix)),
                   np.kron(_fdiff(nz, vsz[2]), np.kron(Iy, Ix))])
    A = Mop + rho * G.T @ G
    rhs = Mop @ vec(x1) + G.T @ np.concatenate([rho * vec(y[c]) - vec(p[c]) for c in range(3)])
    want = np.linalg.lstsq(A, rhs, rcond=1e-12)[0]
    np.testing.assert_allclose(vec(x), want, rtol=0, atol=1e-9 * np.abs(want).max())
