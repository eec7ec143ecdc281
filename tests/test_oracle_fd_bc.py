"""Pins of the backward and central difference restatements (oracle
gradb, gradbm, gradb_adj, gradbm_adj, gradc, gradcm; src/utils/fd/mex):

* gradc against a literal walk of gradc_mex.c's loops in the reference's
  order (tiny grids, float64, every loop with its own bounds and step,
  including the k = 0 "i = nx-1" loop that steps j by 1 and the missing
  last corner): the restatement's per-voxel rule plus its one overwrite
  loop must give the same bits;
* linear ramps: every one-sided and central difference of a linear field is
  its slope (exactly, for integer slopes), so gradb, gradb_adj and gradc
  are the slopes everywhere they write, except where the reference's loops
  say otherwise (gradc's overwritten range, the unwritten corners);
* the masked forms on a full mask against the unmasked ones, and an
  isolated voxel (no neighbour in G: all differences 0).
"""
import numpy as np
import pytest

from oracle import oracle as orc


def gradc_loops(u, h):
    """gradc_mex.c (gradcd), its loops one by one in program order."""
    nx, ny, nz = u.shape
    U = u.reshape(-1, order="F")
    dx, dy, dz = (np.zeros(U.size) for _ in range(3))
    NX, NY, nxny = nx - 1, nx * (ny - 1), nx * ny
    NZ = nxny * (nz - 1)
    hx, hy, hz = (1.0 / v for v in h)
    hx2, hy2, hz2 = (0.5 / v for v in h)
    F = lambda l, s, c: c * (U[l + s] - U[l])          # noqa: E731
    B = lambda l, s, c: c * (U[l] - U[l - s])          # noqa: E731
    Cn = lambda l, s, c: c * (U[l + s] - U[l - s])     # noqa: E731

    def put(l, a, b, c):
        dx[l], dy[l], dz[l] = a, b, c

    # plane k = 0
    put(0, F(0, 1, hx), F(0, nx, hy), F(0, nxny, hz))
    for l in range(nx, NY, nx):
        put(l, F(l, 1, hx), Cn(l, nx, hy2), F(l, nxny, hz))
    for l in range(1, NX):
        put(l, Cn(l, 1, hx2), F(l, nx, hy), F(l, nxny, hz))
    for j in range(nx, NY, nx):
        for i in range(1, NX):
            l = i + j
            put(l, Cn(l, 1, hx2), Cn(l, nx, hy2), F(l, nxny, hz))
    put(NY, F(NY, 1, hx), B(NY, nx, hy), F(NY, nxny, hz))
    for i in range(1, NX):
        l = i + NY
        put(l, Cn(l, 1, hx2), B(l, nx, hy), F(l, nxny, hz))
    put(NX, B(NX, 1, hx), F(NX, nx, hy), F(NX, nxny, hz))
    for j in range(nx, NY):                           # the reference's ++j
        l = j + NX
        put(l, B(l, 1, hx), Cn(l, nx, hy2), F(l, nxny, hz))
    put(NX + NY, B(NX + NY, 1, hx), B(NX + NY, nx, hy), F(NX + NY, nxny, hz))
    # planes 1 .. nz-2
    for k in range(nxny, NZ, nxny):
        put(k, F(k, 1, hx), F(k, nx, hy), Cn(k, nxny, hz2))
        for i in range(1, nx):
            l = i + k
            put(l, Cn(l, 1, hx2), F(l, nx, hy), Cn(l, nxny, hz2))
        l = NX + k
        put(l, B(l, 1, hx), F(l, nx, hy), Cn(l, nxny, hz2))
        for j in range(nx, NY, nx):
            l = j + k
            put(l, F(l, 1, hx), Cn(l, nx, hy2), Cn(l, nxny, hz2))
            for i in range(1, NX):
                l = i + j + k
                put(l, Cn(l, 1, hx2), Cn(l, nx, hy2), Cn(l, nxny, hz2))
            l = NX + j + k
            put(l, B(l, 1, hx), Cn(l, nx, hy2), Cn(l, nxny, hz2))
        l = NY + k
        put(l, F(l, 1, hx), B(l, nx, hy), Cn(l, nxny, hz2))
        for i in range(1, NX):
            l = i + NY + k
            put(l, Cn(l, 1, hx2), B(l, nx, hy), Cn(l, nxny, hz2))
        l = NX + NY + k
        put(l, B(l, 1, hx), B(l, nx, hy), Cn(l, nxny, hz2))
    # plane k = nz-1
    put(NZ, F(NZ, 1, hx), F(NZ, nx, hy), B(NZ, nxny, hz))
    for j in range(nx, NY, nx):
        l = j + NZ
        put(l, F(l, 1, hx), Cn(l, nx, hy2), B(l, nxny, hz))
    for i in range(1, NX):
        l = i + NZ
        put(l, Cn(l, 1, hx2), F(l, nx, hy), B(l, nxny, hz))
    for j in range(nx, NY, nx):
        for i in range(1, NX):
            l = i + j + NZ
            put(l, Cn(l, 1, hx2), Cn(l, nx, hy2), B(l, nxny, hz))
    put(NY + NZ, F(NY + NZ, 1, hx), B(NY + NZ, nx, hy), B(NY + NZ, nxny, hz))
    for i in range(1, NX):
        l = i + NY + NZ
        put(l, Cn(l, 1, hx2), B(l, nx, hy), B(l, nxny, hz))
    put(NX + NZ, B(NX + NZ, 1, hx), F(NX + NZ, nx, hy), B(NX + NZ, nxny, hz))
    for j in range(nx, NY, nx):
        l = NX + j + NZ
        put(l, B(l, 1, hx), Cn(l, nx, hy2), B(l, nxny, hz))
    return tuple(a.reshape(u.shape, order="F") for a in (dx, dy, dz))


@pytest.mark.parametrize("shape", [(5, 4, 3), (2, 2, 2), (3, 5, 4), (7, 3, 2), (4, 6, 5)])
def test_gradc_is_the_reference_loop_order(shape):
    u = np.asfortranarray(np.random.default_rng(sum(shape)).standard_normal(shape))
    h = (0.9, 1.1, 1.3)
    for got, want in zip(orc.gradc(u, h), gradc_loops(u, h)):
        np.testing.assert_array_equal(got, want)


def _ramp(shape, a=3.0, b=-2.0, c=5.0):
    g = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    return np.asfortranarray(a * g[0] + b * g[1] + c * g[2])


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_ramps(dt):
    shape = (6, 5, 4)
    u = _ramp(shape).astype(dt)
    dx, dy, dz = orc.gradb(u, (1, 1, 1))
    assert np.all(dx == 3) and np.all(dy == -2) and np.all(dz == 5)
    du = orc.gradb_adj(*[_ramp(shape, *s).astype(dt) for s in ((1, 0, 0), (0, 1, 0), (0, 0, 1))],
                       (1, 1, 1), init=np.full(shape, 7, dt))
    assert du[-1, -1, -1] == 7                       # never written
    du[-1, -1, -1] = -3
    assert np.all(du == -3)
    dx, dy, dz = orc.gradc(u, (1, 1, 1), init=[np.full(shape, 9, dt)] * 3)
    nx, ny = shape[0], shape[1]
    flat = np.arange(u.size).reshape(shape, order="F")
    over = (flat >= (nx - 1) + nx) & (flat < (nx - 1) + nx * (ny - 1))   # plane k = 0 only
    corner = np.zeros(shape, bool)
    corner[-1, -1, -1] = True
    j = np.indices(shape)[1]
    last = over & (j == ny - 1)    # its central dy reaches into plane 1, row 0
    assert np.all(dz[~corner] == 5)
    assert np.all(dy[~corner & ~last] == -2)
    assert np.all(dy[last] == (5 - (-2) * (ny - 2)) / 2)
    assert np.all(dx[~corner & ~over] == 3)
    # in the overwritten range dx = u(l) - u(l-1): the slope, except across a
    # row start, where it is u(0, j) - u(nx-1, j-1)
    i = np.indices(shape)[0]
    assert np.all(dx[over & (i > 0)] == 3)
    assert np.all(dx[over & (i == 0)] == -2 - 3 * (nx - 1))
    assert dx[-1, -1, -1] == dy[-1, -1, -1] == dz[-1, -1, -1] == 9


def test_masked_forms_on_a_full_mask_and_an_isolated_voxel():
    shape = (5, 4, 6)
    rng = np.random.default_rng(3)
    u = np.asfortranarray(rng.standard_normal(shape))
    G = np.ones(shape, bool)
    h = (1.0, 0.8, 1.2)
    for a, b in zip(orc.gradbm(u, G, h), orc.gradb(u, h)):
        np.testing.assert_array_equal(a, b)
    x, y, z = (np.asfortranarray(rng.standard_normal(shape)) for _ in range(3))
    a = orc.gradbm_adj(x, y, z, G, h)
    b = orc.gradb_adj(x, y, z, h)
    a[-1, -1, -1] = b[-1, -1, -1] = 0                # gradb_adj never writes it
    np.testing.assert_array_equal(a, b)
    # gradcm on a full mask is the clean rule: gradc outside the range its
    # k = 0 loop rewrites and the corner it skips
    c = orc.gradcm(u, G, h)
    d = orc.gradc(u, h)
    nx, ny = shape[0], shape[1]
    flat = np.arange(u.size).reshape(shape, order="F")
    keep = ~((flat >= (nx - 1) + nx) & (flat < (nx - 1) + nx * (ny - 1)))
    keep[-1, -1, -1] = False
    for p, q in zip(c, d):
        np.testing.assert_array_equal(p[keep], q[keep])
    iso = np.zeros(shape, bool)
    iso[2, 2, 3] = True
    for fn in (orc.gradbm, orc.gradcm):
        for a in fn(u, iso, h, init=[np.full(shape, 4.0)] * 3):
            assert a[2, 2, 3] == 0 and np.all(a[~iso] == 4.0)
    du = orc.gradbm_adj(x, y, z, iso, h, init=np.full(shape, 4.0))
    assert du[2, 2, 3] == 0 and np.all(du[~iso] == 4.0)
