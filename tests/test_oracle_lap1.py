"""Pins of the masked Laplacian restatement (oracle lap1; lap1_mex.c, called
by src/utils/fd/lap.m):

* a literal walk of lap1_mex.c's loops: indices in units of the axis stride
  with size_t wrap-around in the inside test ((i-1) < NXX, (j-nx) < NYY,
  (k-nxny) < NZZ) and in the one-sided forms' reach tests (i-3a < N), the
  func_pt table {o, b, f, c} indexed by 2 G[+a] + G[-a]; the restatement's
  per-axis rule must give the same bits on random masks;
* cubics: on a full mask with every extent >= 4, each voxel takes the central
  or the third-order one-sided form, both exact for cubics, so d2u is the
  exact Laplacian 6 a x + 6 b y + 6 c z + 2 (p + q + r) (integer arithmetic,
  exact in float64) -- a second-order one-sided form would miss it;
* second-order one-sided forms: a three-wide mask along x, where they are
  exact for quadratics but the third-order ones are out of reach;
* the first-order fallback of a two-wide slab on a linear ramp: forward
  gives the slope, backward the NEGATIVE slope (u[l-a] - u[l], the
  reference's sign), times 1/h^2;
* outside G the caller's values stay, an isolated interior voxel gets 0.
"""
import numpy as np
import pytest

from oracle import oracle as orc

M64 = (1 << 64) - 1


def lap1_walk(d2u, u, G, h):
    """lap1_mex.c (lapd, cd/fd/bd/od) walked literally in Python."""
    nx, ny, nz = u.shape
    U = u.reshape(-1, order="F")
    g = (np.asarray(G) != 0).reshape(-1, order="F").astype(int)
    du = d2u.reshape(-1, order="F").copy()
    nxny = nx * ny
    nxnynz = nxny * nz
    NX, NY, NZ = nx - 1, nx * (ny - 1), nxny * (nz - 1)
    NXX, NYY, NZZ = (nx - 2) & M64, (nx * (ny - 2)) & M64, (nxny * (nz - 2)) & M64
    hx, hy, hz = (1.0 / (v * v) for v in h)

    def cd(l, a, N, i):
        return U[l - a] - 2.0 * U[l] + U[l + a]

    def fd(l, a, N, i):
        if ((i + 3 * a) & M64) < N and g[l + 2 * a] and g[l + 3 * a]:
            return 2.0 * U[l] - 5.0 * U[l + a] + 4.0 * U[l + 2 * a] - U[l + 3 * a]
        if ((i + 2 * a) & M64) < N and g[l + 2 * a]:
            return U[l] - 2.0 * U[l + a] + U[l + 2 * a]
        return U[l + a] - U[l]

    def bd(l, a, N, i):
        if ((i - 3 * a) & M64) < N and g[l - 3 * a] and g[l - 2 * a]:
            return -U[l - 3 * a] + 4.0 * U[l - 2 * a] - 5.0 * U[l - a] + 2.0 * U[l]
        if ((i - 2 * a) & M64) < N and g[l - 2 * a]:
            return U[l - 2 * a] - 2.0 * U[l - a] + U[l]
        return U[l - a] - U[l]

    def od(l, a, N, i):
        return 0.0

    table = (od, bd, fd, cd)
    for k in range(0, nxnynz, nxny):
        for j in range(0, nxny, nx):
            l = j + k
            for i in range(nx):
                if g[l]:
                    idx = (2 * g[l + 1] + g[l - 1] if ((i - 1) & M64) < NXX
                           else (i == 0) * 2 + (i == NX))
                    du[l] = hx * table[idx](l, 1, nx, i)
                    idx = (2 * g[l + nx] + g[l - nx] if ((j - nx) & M64) < NYY
                           else (j == 0) * 2 + (j == NY))
                    du[l] += hy * table[idx](l, nx, nxny, j)
                    idx = (2 * g[l + nxny] + g[l - nxny] if ((k - nxny) & M64) < NZZ
                           else (k == 0) * 2 + (k == NZ))
                    du[l] += hz * table[idx](l, nxny, nxnynz, k)
                l += 1
    return du.reshape(u.shape, order="F")


@pytest.mark.parametrize("shape,frac", [((6, 5, 4), 0.8), ((2, 2, 2), 1.0), ((7, 3, 5), 0.6),
                                         ((9, 8, 7), 0.7), ((4, 9, 2), 0.5)])
def test_lap1_is_the_reference_loop_walk(shape, frac):
    rng = np.random.default_rng(sum(shape))
    u = np.asfortranarray(rng.standard_normal(shape))
    G = rng.random(shape) < frac
    h = (0.9, 1.1, 1.3)
    init = np.asfortranarray(rng.standard_normal(shape))
    np.testing.assert_array_equal(orc.lap1(u, G, h, init=init), lap1_walk(init, u, G, h))


def _poly(shape, cub, quad, lin=(1.0, -2.0, 3.0)):
    x, y, z = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    u = (cub[0] * x ** 3 + cub[1] * y ** 3 + cub[2] * z ** 3
         + quad[0] * x ** 2 + quad[1] * y ** 2 + quad[2] * z ** 2
         + lin[0] * x + lin[1] * y + lin[2] * z)
    lap = 6 * (cub[0] * x + cub[1] * y + cub[2] * z) + 2 * sum(quad)
    return np.asfortranarray(u), lap


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_cubics_on_a_full_mask(dt):
    shape = (7, 5, 4)
    u, lap = _poly(shape, (1.0, -1.0, 2.0), (3.0, -2.0, 1.0))
    d = orc.lap1(u.astype(dt), np.ones(shape, bool), (1, 1, 1))
    np.testing.assert_array_equal(d, lap.astype(dt))
    # with spacing: 1/h^2 scales each axis term (h = 2, 0.5, 1 are exact)
    u, _ = _poly(shape, (0.0, 0.0, 0.0), (3.0, -2.0, 1.0))
    d = orc.lap1(u, np.ones(shape, bool), (2.0, 0.5, 1.0))
    np.testing.assert_array_equal(d, np.full(shape, 2 * 3.0 / 4 + 2 * -2.0 * 4 + 2 * 1.0))


def test_second_order_one_sided_on_a_three_wide_mask():
    shape = (9, 6, 6)
    u, _ = _poly(shape, (0.0, 0.0, 0.0), (3.0, 0.0, 0.0), lin=(2.0, 0.0, 0.0))
    G = np.zeros(shape, bool)
    G[3:6] = True
    d = orc.lap1(u, G, (1, 1, 1), init=np.full(shape, 7.0))
    assert np.all(d[G] == 6.0) and np.all(d[~G] == 7.0)
    # a cubic along x: the three-wide mask's one-sided forms are second order,
    # exact for quadratics only -- they differ from the true 6x at the ends
    u, lap = _poly(shape, (1.0, 0.0, 0.0), (0.0, 0.0, 0.0), lin=(0.0, 0.0, 0.0))
    d = orc.lap1(u, G, (1, 1, 1))
    assert d[4, 2, 2] == lap[4, 2, 2]                # central: exact
    assert d[3, 2, 2] == u[3, 2, 2] - 2 * u[4, 2, 2] + u[5, 2, 2] != lap[3, 2, 2]
    assert d[5, 2, 2] == u[3, 2, 2] - 2 * u[4, 2, 2] + u[5, 2, 2] != lap[5, 2, 2]


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_first_order_fallback_sign(dt):
    shape = (8, 5, 5)
    u = np.asfortranarray(np.broadcast_to((3.0 * np.arange(8))[:, None, None], shape)).astype(dt)
    G = np.zeros(shape, bool)
    G[4:6] = True
    d = orc.lap1(u, G, (0.5, 1.0, 1.0))
    assert np.all(d[4] == 3.0 * 4) and np.all(d[5] == -3.0 * 4)
    # on the array's own faces the forms are one-sided whatever G holds
    G = np.zeros(shape, bool)
    G[0] = G[-1] = True
    d = orc.lap1(u, G, (1.0, 1.0, 1.0))
    assert np.all(d[0] == 3.0) and np.all(d[-1] == -3.0)


def test_isolated_voxel_and_untouched_outside():
    shape = (6, 6, 6)
    rng = np.random.default_rng(5)
    u = np.asfortranarray(rng.standard_normal(shape))
    G = np.zeros(shape, bool)
    G[2, 3, 3] = True
    d = orc.lap1(u, G, (1, 1, 1), init=np.full(shape, -4.0))
    assert d[2, 3, 3] == 0.0 and np.all(d[~G] == -4.0)
