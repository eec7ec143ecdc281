"""Pins of the oracle's finite-difference restatements (gradfp, gradfp_adj,
gradf, gradfm, gradf_adj, gradfm_adj, lapw).

* gradfp / gradfp_adj: impulses and periodic wrap; exact transpose pair
  (dot-product test), which fixes every index and sign (F12).
* gradf / gradf_adj: ramps and quadratics, including the extrapolated last
  face and the forward first face (they are NOT a transpose pair).
* gradfm / gradfm_adj: reduce to the unmasked forms on a full mask (bitwise,
  same summation order) and follow the forward/backward/zero rule on an
  explicit small mask.
* lapw: quadratics give 2a+2b+2c inside and 0 on the faces; adding 2*pi to
  a voxel leaves it unchanged; agrees with the angle-difference form
  atan2(sin d, cos d) that src/unwrap/lapw.m:40-70 keeps as a comment.
"""
import numpy as np
import pytest

from oracle import oracle as orc


def F(a, dt=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dt))


SH = (7, 6, 5)
H = (0.9, 1.1, 1.3)


def _idx(shape):
    i = np.arange(shape[0])[:, None, None] * np.ones(shape)
    j = np.arange(shape[1])[None, :, None] * np.ones(shape)
    k = np.arange(shape[2])[None, None, :] * np.ones(shape)
    return i, j, k


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_gradfp_impulse_and_wrap(dt):
    u = np.zeros(SH)
    u[0, 2, 4] = 1.0
    dx, dy, dz = orc.gradfp(F(u, dt), H)
    # forward difference: d[l] = (u[l+1] - u[l]) / h, periodic
    assert dx[0, 2, 4] == pytest.approx(-1 / H[0], rel=1e-6)
    assert dx[-1, 2, 4] == pytest.approx(1 / H[0], rel=1e-6)  # wraps from i = nx-1
    assert dy[0, 1, 4] == pytest.approx(1 / H[1], rel=1e-6)
    assert dz[0, 2, 3] == pytest.approx(1 / H[2], rel=1e-6)
    assert dz[0, 2, 0] == 0.0  # u[.., 0] - u[.., 0]... no wrap target here
    assert np.count_nonzero(dx) == 2 and np.count_nonzero(dy) == 2 and np.count_nonzero(dz) == 2


@pytest.mark.parametrize("shape", [(7, 6, 5), (1, 4, 3), (5, 1, 1)])
def test_gradfp_adjoint_is_transpose(shape):
    rng = np.random.default_rng(0)
    u = F(rng.standard_normal(shape))
    x, y, z = (F(rng.standard_normal(shape)) for _ in range(3))
    dx, dy, dz = orc.gradfp(u, H)
    du = orc.gradfp_adj(x, y, z, H)
    lhs = (dx * x).sum() + (dy * y).sum() + (dz * z).sum()
    rhs = (u * du).sum()
    assert abs(lhs - rhs) <= 1e-13 * (abs(lhs) + 1)


def test_gradfp_adj_region_term_order_is_a_sum():
    """The reference sums the three axis terms in a region-dependent order;
    the value is that sum (checked against numpy's periodic divergence)."""
    rng = np.random.default_rng(1)
    x, y, z = (F(rng.standard_normal(SH)) for _ in range(3))
    du = orc.gradfp_adj(x, y, z, H)
    want = -((x - np.roll(x, 1, 0)) / H[0] + (y - np.roll(y, 1, 1)) / H[1] + (z - np.roll(z, 1, 2)) / H[2])
    np.testing.assert_allclose(du, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_gradf_ramp_and_last_face(dt):
    i, j, k = _idx(SH)
    dx, dy, dz = orc.gradf(F(2 * i - 3 * j + 0.5 * k, dt), H)
    tol = 1e-6 if dt == np.float32 else 1e-14
    np.testing.assert_allclose(dx, 2 / H[0], rtol=tol)
    np.testing.assert_allclose(dy, -3 / H[1], rtol=tol)
    np.testing.assert_allclose(dz, 0.5 / H[2], rtol=tol)
    # quadratic: forward difference inside, backward on the last face
    dx, _, _ = orc.gradf(F(i * i, dt), H)
    np.testing.assert_allclose(dx[:-1], (2 * i[:-1] + 1) / H[0], rtol=tol)
    np.testing.assert_allclose(dx[-1], (2 * i[-1] - 1) / H[0], rtol=tol)


def test_gradf_adj_first_face_forward():
    i, j, k = _idx(SH)
    x = F(i * i)
    z0 = np.zeros(SH)
    du = orc.gradf_adj(x, F(z0), F(z0), H)
    # -(backward difference)/h inside, -(forward difference)/h on the first face
    np.testing.assert_allclose(du[1:], -(2 * i[1:] - 1) / H[0], rtol=1e-14)
    np.testing.assert_allclose(du[0], -(1.0) / H[0], rtol=1e-14)


def test_gradf_pair_is_not_a_transpose():
    rng = np.random.default_rng(2)
    u = F(rng.standard_normal(SH))
    x, y, z = (F(rng.standard_normal(SH)) for _ in range(3))
    dx, dy, dz = orc.gradf(u, H)
    du = orc.gradf_adj(x, y, z, H)
    lhs = (dx * x).sum() + (dy * y).sum() + (dz * z).sum()
    assert abs(lhs - (u * du).sum()) > 1e-6  # F12: replicate boundaries


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_masked_forms_reduce_to_unmasked_on_full_mask(dt):
    rng = np.random.default_rng(3)
    u = F(rng.standard_normal(SH), dt)
    G = np.ones(SH, dtype=np.uint8)
    for a, b in zip(orc.gradfm(u, G, H), orc.gradf(u, H)):
        np.testing.assert_array_equal(a, b)
    x, y, z = (F(rng.standard_normal(SH), dt) for _ in range(3))
    np.testing.assert_array_equal(orc.gradfm_adj(x, y, z, G, H), orc.gradf_adj(x, y, z, H))


def test_gradfm_rule_on_a_small_mask():
    u = F(np.arange(5.0) ** 2).reshape(5, 1, 1, order="F")
    u = F(u)
    G = np.array([0, 1, 1, 0, 1], dtype=np.uint8).reshape(5, 1, 1)
    init = [F(np.full((5, 1, 1), 7.0)) for _ in range(3)]
    dx, dy, dz = orc.gradfm(u, F(G, np.uint8), (1, 1, 1), init=init)
    # i=0 outside G: untouched; i=1: forward (i=2 in G): 4-1=3; i=2: +1 not in
    # G, -1 in G: backward 4-1=3; i=3 outside; i=4: no neighbour in G: 0
    np.testing.assert_array_equal(dx.ravel(), [7.0, 3.0, 3.0, 7.0, 0.0])
    np.testing.assert_array_equal(dy.ravel(), [7.0, 0.0, 0.0, 7.0, 0.0])


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_lapw_quadratic_interior_and_faces(dt):
    shape = (8, 7, 6)
    h = (0.9, 1.0, 1.2)
    i, j, k = _idx(shape)
    x, y, z = i * h[0], j * h[1], k * h[2]
    a, b, c = 0.01, -0.02, 0.015
    u = F(a * x * x + b * y * y + c * z * z, dt)
    d = orc.lapw(u, h)
    it = np.zeros(shape, bool)
    it[1:-1, 1:-1, 1:-1] = True
    tol = 2e-4 if dt == np.float32 else 1e-10
    np.testing.assert_allclose(d[it], 2 * (a + b + c), rtol=tol)
    assert np.all(d[~it] == 0)


def test_lapw_wrap_invariance_and_angle_form():
    shape = (9, 8, 7)
    rng = np.random.default_rng(4)
    true = np.cumsum(rng.standard_normal(shape) * 0.8, axis=0)
    wrapped = np.angle(np.exp(1j * true))
    d0 = orc.lapw(F(wrapped), (1, 1, 1))
    w2 = wrapped.copy()
    w2[4, 4, 3] += 2 * np.pi
    d1 = orc.lapw(F(w2), (1, 1, 1))
    np.testing.assert_allclose(d1, d0, atol=1e-12)

    def ad(dlt):
        return np.arctan2(np.sin(dlt), np.cos(dlt))
    it = np.zeros(shape, bool)
    it[1:-1, 1:-1, 1:-1] = True
    want = np.zeros(shape)
    for ax in range(3):
        fw = ad(np.roll(wrapped, -1, ax) - wrapped)
        bw = ad(wrapped - np.roll(wrapped, 1, ax))
        want += fw - bw
    np.testing.assert_allclose(d0[it], want[it], atol=1e-12)
