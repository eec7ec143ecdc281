"""Pins of the masked Poisson restatement (oracle/poisson.py):

* on a full box the discrete problem -Lap_h v = f with v = 0 just outside
  the volume is diagonalised by the type-I sine transform: the sparse
  solution equals the DST closed form (which fixes the stencil, its signs,
  the voxel-size scaling and the zero border);
* unwrapLaplacian of a wrapped phase that vanishes on the volume's outer
  layer, with the mask the volume's interior (lapw_mex.c fills interior
  voxels only) and steps below pi, returns the unwrapped phase exactly;
* a masked problem: the solution is zero off the mask and the residual of
  the original stencil vanishes on it.
"""
import numpy as np
import pytest
from scipy.fft import dstn, idstn

from oracle import poisson as po


@pytest.mark.parametrize("shape,vsz", [((9, 7, 6), (1.0, 1.0, 1.0)), ((8, 10, 5), (0.9, 1.2, 1.1))])
def test_box_solution_is_the_dst_closed_form(shape, vsz):
    rng = np.random.default_rng(0)
    f = np.asfortranarray(rng.standard_normal(shape))
    v = po.solve(f, np.ones(shape, bool), vsz)
    lam = sum((2 - 2 * np.cos(np.pi * (np.arange(n) + 1) / (n + 1))).reshape(
        [-1 if a == ax else 1 for a in range(3)]) / vsz[ax] ** 2 for ax, n in enumerate(shape))
    want = idstn(dstn(f, type=1) / lam, type=1)
    np.testing.assert_allclose(v, want, rtol=0, atol=1e-12 * np.abs(want).max())


def sine_phase(shape, amp):
    """amp * prod sin(pi i / (n - 1)): zero on the outer layer."""
    g = np.meshgrid(*(np.sin(np.pi * np.arange(n) / (n - 1)) for n in shape), indexing="ij")
    return np.asfortranarray(amp * g[0] * g[1] * g[2])


def interior(shape):
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


@pytest.mark.parametrize("vsz", [(1.0, 1.0, 1.0), (0.8, 1.0, 1.3)])
def test_unwrap_recovers_a_wrapped_phase(vsz):
    shape = (20, 18, 16)
    phi = sine_phase(shape, 12.0)             # max step ~ 12 * pi / 15 < pi
    wrapped = np.angle(np.exp(1j * phi))
    assert np.abs(wrapped - phi).max() > 6    # really wrapped
    m = interior(shape)
    up, d2 = po.unwrap_laplacian(wrapped, m, vsz)
    np.testing.assert_allclose(up, phi, atol=1e-9)
    assert np.all(d2[~m] == 0)


def test_masked_solution_residual():
    rng = np.random.default_rng(1)
    shape = (14, 13, 12)
    c = [(n - 1) / 2 for n in shape]
    g = np.meshgrid(*(np.arange(n) - c[a] for a, n in enumerate(shape)), indexing="ij")
    m = (g[0] ** 2 + g[1] ** 2 + g[2] ** 2) <= 25
    f = rng.standard_normal(shape)
    v = po.solve(f, m, (1.0, 1.1, 0.9))
    assert np.all(v[~m] == 0)
    hx, hy, hz = 1.0, 1 / 1.21, 1 / 0.81
    P = np.pad(v, 1)
    lap = (hx * (P[2:, 1:-1, 1:-1] + P[:-2, 1:-1, 1:-1]) + hy * (P[1:-1, 2:, 1:-1] + P[1:-1, :-2, 1:-1])
           + hz * (P[1:-1, 1:-1, 2:] + P[1:-1, 1:-1, :-2]) - 2 * (hx + hy + hz) * v)
    np.testing.assert_allclose(-lap[m], f[m], atol=1e-10)


@pytest.mark.parametrize("mg", [{}, {"nlevels": 1}, {"mu": 2, "npre": 2, "npost": 2},
                                {"tol": 1e-3, "maxit": 3}])
def test_port_of_the_reference_mgpcg_converges_to_the_exact_solution(mg):
    """The numpy restatement of mgpcg_mex.c (used to time the reference's
    algorithm on the CPU) reaches the exact discrete solution, in few
    iterations: a working multigrid preconditioner."""
    rng = np.random.default_rng(2)
    shape = (26, 22, 20)
    c = [(n - 1) / 2 for n in shape]
    g = np.meshgrid(*(np.arange(n) - c[a] for a, n in enumerate(shape)), indexing="ij")
    m = ((g[0] ** 2 + g[1] ** 2 + g[2] ** 2) <= 81) & ~(((g[0] - 2) ** 2 + g[1] ** 2) < 4)
    f = rng.standard_normal(shape) * m
    v, it, res = po.mgpcg_port(f, m, (1.0, 1.2, 0.9), 1e-12, 100, mg)
    want = po.solve(f, m, (1.0, 1.2, 0.9))
    assert np.linalg.norm(v - want) <= 1e-9 * np.linalg.norm(want)
    assert it <= (40 if mg.get("nlevels") == 1 else 20), it
    assert np.all(np.diff(np.log(res)) < 0.5)


def test_port_unwrap_recovers_a_wrapped_phase():
    shape = (24, 22, 20)
    phi = sine_phase(shape, 10.0)
    up, _, it = po.unwrap_laplacian_port(np.angle(np.exp(1j * phi)), interior(shape), (1, 1, 1))
    np.testing.assert_allclose(up, phi, atol=1e-6)


def test_manufactured_solution_second_order():
    """testPoissonSolvers.m's model problem: u = sin(pX) sin(pY) sin(pZ),
    p = 2 pi, on the unit cube at h = 1/n with the interior active and
    f = 3 p^2 u; the exact discrete solution converges to u at second order
    (error ratios ~4 per halving of h)."""
    p = 2 * np.pi
    errs = []
    for n in (8, 16, 32):
        h = 1.0 / n
        x = np.arange(n + 1) * h
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        u = np.sin(p * X) * np.sin(p * Y) * np.sin(p * Z)
        G = np.zeros(u.shape, bool)
        G[1:-1, 1:-1, 1:-1] = True
        v = po.solve(3 * p * p * u * G, G, (h, h, h))
        errs.append(np.linalg.norm((v - u)[G]) / np.sqrt(G.sum()))
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(3.6 < r < 4.7 for r in ratios), (errs, ratios)
