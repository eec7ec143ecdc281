"""Pins of the oracle's primal-dual restatement (oracle/qsm_oracle.c).

The reference C-MEX cannot be built here (it needs MATLAB's mex.h) and its
repository holds no golden vectors for this path, so the restatement is
pinned by what the mathematics fixes, chosen so that plausible mistakes (a
dropped term, a wrong sign, index or coefficient, a transposed operand)
fail at least one check:

* step sizes: closed form and tau = sigma (pd_tgv_mex.c:272-278);
* init_bitmask: an independent transcription of the reference's 13-bit
  neighbour code (pd_tgv_mex.c:1066-1116) on random masks;
* closed forms of every term on polynomial fields (quadratics for the
  Laplacian / wave stencils, linear fields for gradients, symmetric
  gradient, divergences), including the anisotropic div2 scaling the
  reference uses (each w component scaled by its own h, pd_tgv_mex.c:934-959);
* adjointness: the dual's linear part is sigma*K and the primal's is
  -tau*K^T for isotropic voxels on masked domains (q paired with weights
  1,2,2,1,2,1), which fixes every sign and every forward/backward index;
* projection invariants (|p| <= alpha1, weighted |q| <= alpha0);
* the TV dual's projection onto the lambda-ball as the Euclidean projection
  of a known vector (chi_ = 0 leaves p at its input), with lambda < 1 so
  that a threshold on |p| or on lambda instead of lambda^2 fails;
* convergence norms on data whose ratios are exact (1, 1/4, 3/4, 0, NaN for
  0/0), the flag's two strict thresholds on both sides, and the single
  path's double accumulation and mixed float/double difference term;
* whole loop: zero data stays zero, tol = 0 never stops, and the iterates
  converge to the constraint Delta(eta) - W(chi) + f = 0 on the interior
  mask (the saddle point of the primal-dual problem).
"""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.filterwarnings("ignore::RuntimeWarning")


def F(a, dt=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dt))


def zeros(shape, n=None, dt=np.float64):
    return np.zeros(shape if n is None else shape + (n,), dtype=dt, order="F")


def coords(shape, h):
    i = np.arange(shape[0])[:, None, None] * h[0]
    j = np.arange(shape[1])[None, :, None] * h[1]
    k = np.arange(shape[2])[None, None, :] * h[2]
    return (np.broadcast_to(i, shape), np.broadcast_to(j, shape), np.broadcast_to(k, shape))


def interior(shape):
    m = np.zeros(shape, dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


# ---------------------------------------------------------------------------
# step sizes and masks
# ---------------------------------------------------------------------------

def test_step_sizes_closed_form():
    tau, sigma = orc.step_sizes((1.0, 1.0, 1.0))
    # gnrm = 4*3 = 12, wnrm = 6*(1+1+2)/3 = 8, K2 = 144 + 64 = 208
    assert tau == pytest.approx(1 / np.sqrt(208.0), rel=1e-15)
    assert sigma == pytest.approx(tau, rel=1e-15)
    h = (0.9, 0.9, 1.2)
    tau, sigma = orc.step_sizes(h)
    g = 4 * sum(1 / x ** 2 for x in h)
    w = 6 * (1 / h[0] ** 2 + 1 / h[1] ** 2 + 2 / h[2] ** 2) / 3
    assert tau ** 2 * (g * g + w * w) == pytest.approx(1.0, rel=1e-14)
    assert sigma == pytest.approx(tau, rel=1e-14)


def _bitmask_reference_style(m):
    """The reference's init_bitmask, transcribed bit by bit from its
    neighbour-code description (o = 1..12 bit positions, INSIDE = bits
    4 | 8 | 12, OUTSIDE = bits 1 | 5 | 9)."""
    nx, ny, nz = m.shape
    mi = np.zeros(m.shape, dtype=np.uint8)
    inside = (1 << 4) | (1 << 8) | (1 << 12)
    outside = (1 << 1) | (1 << 5) | (1 << 9)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                mb = 0
                for o, idx, n in ((1, i, nx), (5, j, ny), (9, k, nz)):
                    nb = [i, j, k]
                    fw = idx < n - 1
                    if fw:
                        nb[[1, 5, 9].index(o)] += 1
                        fw = bool(m[tuple(nb)])
                    nb = [i, j, k]
                    bw = idx > 0
                    if bw:
                        nb[[1, 5, 9].index(o)] -= 1
                        bw = bool(m[tuple(nb)])
                    mb |= (1 << (o + 2)) if fw else 0
                    mb |= (1 << (o + 1)) if bw else 0
                    mb |= 1 << (o + 2 * fw + bw)
                mb = 0 if (mb & outside) == outside else mb
                mi[i, j, k] = (mb & inside) == inside
    return mi


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_init_mask_matches_bitfield_formulation(seed):
    rng = np.random.default_rng(seed)
    m = F(rng.random((7, 6, 5)) < 0.75, np.uint8)
    np.testing.assert_array_equal(orc.init_mask(m), _bitmask_reference_style(m))


def test_init_mask_ignores_own_voxel():
    """A hole whose six neighbours are in m is still dual-active (F7)."""
    m = np.ones((5, 5, 5), dtype=np.uint8)
    m[2, 2, 2] = 0
    mi = orc.init_mask(F(m, np.uint8))
    assert mi[2, 2, 2] == 1
    assert mi[1, 2, 2] == 0  # its -x neighbour now misses a neighbour
    assert mi[0].sum() == 0 and mi[-1].sum() == 0


# ---------------------------------------------------------------------------
# closed forms of the dual terms
# ---------------------------------------------------------------------------

SHAPE = (9, 8, 7)
HANI = (0.9, 1.1, 1.3)


def _dual(cb=None, eb=None, wb=None, f=None, nu=None, p=None, q=None, h=HANI,
          a0=1e30, a1=1e30, dt=np.float64):
    sh = SHAPE
    nu = zeros(sh, dt=dt) if nu is None else F(nu, dt)
    p = zeros(sh, 3, dt) if p is None else F(p, dt)
    q = zeros(sh, 6, dt) if q is None else F(q, dt)
    cb = zeros(sh, dt=dt) if cb is None else F(cb, dt)
    eb = zeros(sh, dt=dt) if eb is None else F(eb, dt)
    wb = zeros(sh, 3, dt) if wb is None else F(wb, dt)
    f = zeros(sh, dt=dt) if f is None else F(f, dt)
    mi = F(interior(sh), np.uint8)
    tau, sigma = orc.step_sizes(h)
    orc.tgv_dual(nu, p, q, cb, eb, wb, f, mi, h, a0, a1, sigma)
    return nu, p, q, sigma


def test_dual_laplacian_of_quadratic():
    x, y, z = coords(SHAPE, HANI)
    for e, lap in ((x * x, 2.0), (y * y, 2.0), (z * z, 2.0), (3 * x * x - y * y + 0.5 * z * z, 6 - 2 + 1)):
        nu, _, _, s = _dual(eb=e)
        it = interior(SHAPE)
        np.testing.assert_allclose(nu[it], s * lap, rtol=1e-10)
        assert np.all(nu[~it] == 0)


def test_dual_wave_operator_signs():
    """wave = -(1/3)d_xx - (1/3)d_yy + (2/3)d_zz enters nu with a minus sign
    (pd_tgv_mex.c:380-383, :581-585): B0 along z."""
    x, y, z = coords(SHAPE, HANI)
    it = interior(SHAPE)
    for c, wave in ((x * x, -2 / 3), (y * y, -2 / 3), (z * z, 4 / 3)):
        nu, _, _, s = _dual(cb=c)
        np.testing.assert_allclose(nu[it], -s * wave, rtol=1e-10)


def test_dual_f_term_and_accumulation():
    rng = np.random.default_rng(3)
    f = rng.standard_normal(SHAPE)
    nu0 = rng.standard_normal(SHAPE)
    nu, _, _, s = _dual(f=f, nu=nu0)
    it = interior(SHAPE)
    np.testing.assert_allclose(nu[it], nu0[it] + f[it], rtol=1e-15)  # f is pre-scaled by sigma
    np.testing.assert_array_equal(nu[~it], nu0[~it])


def test_dual_p_forward_gradient_minus_w():
    x, y, z = coords(SHAPE, HANI)
    c = 2.0 * x - 3.0 * y + 0.5 * z
    wb = zeros(SHAPE, 3)
    wb[..., 0], wb[..., 1], wb[..., 2] = 0.25, -1.0, 0.75
    _, p, _, s = _dual(cb=c, wb=wb)
    it = interior(SHAPE)
    np.testing.assert_allclose(p[..., 0][it], s * (2.0 - 0.25), rtol=1e-10)
    np.testing.assert_allclose(p[..., 1][it], s * (-3.0 + 1.0), rtol=1e-10)
    np.testing.assert_allclose(p[..., 2][it], s * (0.5 - 0.75), rtol=1e-10)


def test_dual_p_uses_forward_difference():
    """An impulse in chi_ at l gives p1 = -sigma/hx at l and +sigma/hx at l-1."""
    c = zeros(SHAPE)
    c[4, 4, 3] = 1.0
    _, p, _, s = _dual(cb=c)
    assert p[4, 4, 3, 0] == pytest.approx(-s / HANI[0], rel=1e-14)
    assert p[3, 4, 3, 0] == pytest.approx(s / HANI[0], rel=1e-14)
    assert p[5, 4, 3, 0] == 0.0
    assert p[4, 3, 3, 1] == pytest.approx(s / HANI[1], rel=1e-14)
    assert p[4, 4, 2, 2] == pytest.approx(s / HANI[2], rel=1e-14)


def test_dual_symmetric_gradient():
    """q = sigma*E(w_) with E = [[dx w1, (dx w2 + dy w1)/2, (dx w3 + dz w1)/2],
    [., dy w2, (dy w3 + dz w2)/2], [., ., dz w3]], forward differences."""
    x, y, z = coords(SHAPE, HANI)
    wb = zeros(SHAPE, 3)
    wb[..., 0] = 1.0 * x + 2.0 * y + 3.0 * z
    wb[..., 1] = 4.0 * x + 5.0 * y + 6.0 * z
    wb[..., 2] = 7.0 * x + 8.0 * y + 9.0 * z
    _, _, q, s = _dual(wb=wb)
    it = interior(SHAPE)
    want = [1.0, (4.0 + 2.0) / 2, (7.0 + 3.0) / 2, 5.0, (8.0 + 6.0) / 2, 9.0]
    for c in range(6):
        np.testing.assert_allclose(q[..., c][it], s * want[c], rtol=1e-10, err_msg=f"q{c + 1}")


def test_dual_projections():
    rng = np.random.default_rng(4)
    cb = rng.standard_normal(SHAPE)
    wb = rng.standard_normal(SHAPE + (3,))
    a1, a0 = 0.05, 0.08
    nu, p, q, s = _dual(cb=cb, wb=wb, a0=a0, a1=a1)
    np_, _, q_free, _ = _dual(cb=cb, wb=wb)
    _, p_free, _, _ = _dual(cb=cb, wb=wb)
    pn = np.sqrt((p ** 2).sum(-1))
    qn = np.sqrt(q[..., 0] ** 2 + q[..., 3] ** 2 + q[..., 5] ** 2
                 + 2 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 4] ** 2))
    assert pn.max() <= a1 * (1 + 1e-14)
    assert qn.max() <= a0 * (1 + 1e-14)
    # projected vectors are scaled copies of the unprojected ones
    pf = np.sqrt((p_free ** 2).sum(-1))
    big = pf > a1
    assert big.sum() > 10
    np.testing.assert_allclose(p[big], p_free[big] * (a1 / pf[big])[:, None], rtol=1e-13)
    small = (pf <= a1) & interior(SHAPE)
    np.testing.assert_array_equal(p[small], p_free[small])


# ---------------------------------------------------------------------------
# closed forms of the primal terms
# ---------------------------------------------------------------------------

def _primal(nu=None, p=None, q=None, chi=None, eta=None, w=None, h=HANI, dt=np.float64):
    sh = SHAPE
    chi = zeros(sh, dt=dt) if chi is None else F(chi, dt)
    eta = zeros(sh, dt=dt) if eta is None else F(eta, dt)
    w = zeros(sh, 3, dt) if w is None else F(w, dt)
    cb, eb, wb = zeros(sh, dt=dt), zeros(sh, dt=dt), zeros(sh, 3, dt)
    nu = zeros(sh, dt=dt) if nu is None else F(nu, dt)
    p = zeros(sh, 3, dt) if p is None else F(p, dt)
    q = zeros(sh, 6, dt) if q is None else F(q, dt)
    m = F(interior(sh), np.uint8)
    tau, sigma = orc.step_sizes(h)
    orc.tgv_primal(chi, eta, w, cb, eb, wb, nu, p, q, m, h, tau)
    return chi, eta, w, cb, eb, wb, tau


def test_primal_eta_prox_divides_by_one_plus_tau():
    x, y, z = coords(SHAPE, HANI)
    rng = np.random.default_rng(5)
    eta0 = rng.standard_normal(SHAPE)
    nu = 0.3 * x * x - 0.2 * y * y + 0.1 * z * z
    chi, eta, w, cb, eb, wb, tau = _primal(nu=nu, eta=eta0)
    it = interior(SHAPE)
    lap = 2 * (0.3 - 0.2 + 0.1)
    np.testing.assert_allclose(eta[it], (eta0[it] - tau * lap) / (1 + tau), rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(eb[it], 2 * eta[it] - eta0[it], rtol=1e-10, atol=1e-14)


def test_primal_chi_divergence_and_wave():
    x, y, z = coords(SHAPE, HANI)
    p = zeros(SHAPE, 3)
    p[..., 0], p[..., 1], p[..., 2] = 1.5 * x, -2.0 * y, 0.5 * z
    nu = z * z
    chi, eta, w, cb, eb, wb, tau = _primal(nu=nu, p=p)
    it = interior(SHAPE)
    # chi += tau*(div^- p + wave(nu)), wave(z^2) = (2/3)*2
    np.testing.assert_allclose(chi[it], tau * ((1.5 - 2.0 + 0.5) + 4 / 3), rtol=1e-10)
    np.testing.assert_allclose(cb[it], 2 * chi[it], rtol=1e-14)


def test_primal_divergence_is_backward():
    p = zeros(SHAPE, 3)
    p[4, 4, 3, 0] = 1.0
    chi, *_, tau = _primal(p=p)
    assert chi[4, 4, 3] == pytest.approx(tau / HANI[0], rel=1e-14)
    assert chi[5, 4, 3] == pytest.approx(-tau / HANI[0], rel=1e-14)
    assert chi[3, 4, 3] == 0.0


def test_primal_w_update_with_reference_scaling():
    """w_c += tau*p_c + (tau/h_c)*(sum of backward differences of row c of q
    in index units): each component is scaled by its own h (pd_tgv_mex.c:
    934-959), not by the difference direction's."""
    i = np.arange(SHAPE[0])[:, None, None] * np.ones(SHAPE)
    j = np.arange(SHAPE[1])[None, :, None] * np.ones(SHAPE)
    k = np.arange(SHAPE[2])[None, None, :] * np.ones(SHAPE)
    q = zeros(SHAPE, 6)
    # index-unit slopes: q1 in x, q2 in x and y, q3 in z, q4 in y, q5 in y and z, q6 in z
    q[..., 0] = 1.0 * i
    q[..., 1] = 2.0 * i + 3.0 * j
    q[..., 2] = 4.0 * k
    q[..., 3] = 5.0 * j
    q[..., 4] = 6.0 * j + 7.0 * k
    q[..., 5] = 8.0 * k
    p = zeros(SHAPE, 3)
    p[...] = (0.1, 0.2, 0.3)
    chi, eta, w, cb, eb, wb, tau = _primal(p=p, q=q)
    it = interior(SHAPE)
    hx, hy, hz = HANI
    want1 = tau * 0.1 + (tau / hx) * (1.0 + 3.0 + 4.0)   # dx q1 + dy q2 + dz q3
    want2 = tau * 0.2 + (tau / hy) * (2.0 + 5.0 + 7.0)   # dx q2 + dy q4 + dz q5
    want3 = tau * 0.3 + (tau / hz) * (0.0 + 6.0 + 8.0)   # dx q3 + dy q5 + dz q6
    np.testing.assert_allclose(w[..., 0][it], want1, rtol=1e-12)
    np.testing.assert_allclose(w[..., 1][it], want2, rtol=1e-12)
    np.testing.assert_allclose(w[..., 2][it], want3, rtol=1e-12)


# ---------------------------------------------------------------------------
# adjointness of the dual and primal linear parts
# ---------------------------------------------------------------------------

QW = np.array([1.0, 2.0, 2.0, 1.0, 2.0, 1.0])


def _smooth_mask(shape):
    c = [(n - 1) / 2 for n in shape]
    x = (np.arange(shape[0]) - c[0])[:, None, None] / (0.45 * shape[0])
    y = (np.arange(shape[1]) - c[1])[None, :, None] / (0.45 * shape[1])
    z = (np.arange(shape[2]) - c[2])[None, None, :] / (0.45 * shape[2])
    # kept off the grid faces: the primal sweep never updates face voxels,
    # so a face voxel in m would carry x without K^T y
    return F(((x * x + y * y + z * z) <= 1.0) & interior(shape), np.uint8)


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.7, 0.7, 0.7)])
@pytest.mark.parametrize("seed", [0, 1])
def test_dual_primal_adjoint(h, seed):
    shape = (12, 11, 10)
    rng = np.random.default_rng(seed)
    m = _smooth_mask(shape)
    mi = orc.init_mask(m)
    assert np.all(m[mi.astype(bool)] == 1)  # no holes: mi inside m
    mb, mib = m.astype(bool), mi.astype(bool)
    # x = (chi_, eta_, w_) on m
    cb = F(rng.standard_normal(shape) * mb)
    eb = F(rng.standard_normal(shape) * mb)
    wb = F(rng.standard_normal(shape + (3,)) * mb[..., None])
    # y = (nu, p, q) on mi
    nu = F(rng.standard_normal(shape) * mib)
    p = F(rng.standard_normal(shape + (3,)) * mib[..., None])
    q = F(rng.standard_normal(shape + (6,)) * mib[..., None])
    tau, sigma = orc.step_sizes(h)
    # K x via the dual from zero duals, f = 0, no projection
    knu, kp, kq = zeros(shape), zeros(shape, 3), zeros(shape, 6)
    orc.tgv_dual(knu, kp, kq, cb, eb, wb, zeros(shape), mi, h, 1e300, 1e300, sigma)
    knu, kp, kq = knu / sigma, kp / sigma, kq / sigma
    # K^T y via the primal from zero primals: increments = -tau K^T y (eta
    # before its 1/(1+tau) prox scaling)
    chi, eta, w = zeros(shape), zeros(shape), zeros(shape, 3)
    orc.tgv_primal(chi, eta, w, zeros(shape), zeros(shape), zeros(shape, 3), nu, p, q, m, h, tau)
    kt_chi = -chi / tau
    kt_eta = -eta * (1 + tau) / tau
    kt_w = -w / tau
    lhs = (knu * nu).sum() + (kp * p).sum() + (kq * q * QW).sum()
    rhs = (kt_chi * cb).sum() + (kt_eta * eb).sum() + (kt_w * wb).sum()
    scale = np.sqrt((knu ** 2).sum() * (nu ** 2).sum())
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)


def test_tv_dual_primal_adjoint():
    shape, h = (11, 12, 9), (1.0, 1.0, 1.0)
    rng = np.random.default_rng(7)
    m = _smooth_mask(shape)
    mi = orc.init_mask(m)
    mb, mib = m.astype(bool), mi.astype(bool)
    cb = F(rng.standard_normal(shape) * mb)
    eb = F(rng.standard_normal(shape) * mb)
    nu = F(rng.standard_normal(shape) * mib)
    p = F(rng.standard_normal(shape + (3,)) * mib[..., None])
    tau, sigma = orc.step_sizes(h)
    knu, kp = zeros(shape), zeros(shape, 3)
    orc.tv_dual(knu, kp, cb, eb, zeros(shape), mi, h, 1e300, sigma)
    chi, eta = zeros(shape), zeros(shape)
    orc.tv_primal(chi, eta, zeros(shape), zeros(shape), nu, p, m, h, tau)
    lhs = (knu / sigma * nu).sum() + (kp / sigma * p).sum()
    rhs = (-chi / tau * cb).sum() + (-eta * (1 + tau) / tau * eb).sum()
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


# ---------------------------------------------------------------------------
# TV dual: projection onto the lambda-ball (pd_tv_mex.c:395-402, 498-506)
# ---------------------------------------------------------------------------

def _tv_dual_run(p0, lam, dt, cb=None):
    sh = p0.shape[:3]
    nu = zeros(sh, dt=dt)
    p = F(p0, dt)
    cb = zeros(sh, dt=dt) if cb is None else F(cb, dt)
    mi = F(interior(sh), np.uint8)
    tau, sigma = orc.step_sizes(HANI)
    orc.tv_dual(nu, p, cb, zeros(sh, dt=dt), zeros(sh, dt=dt), mi, HANI, lam, sigma)
    return nu, p


@pytest.mark.parametrize("dt,rt", [(np.float64, 1e-15), (np.float32, 3e-7)])
def test_tv_dual_projection_onto_lambda_ball(dt, rt):
    """With chi_ = eta_ = f = 0 the dual step leaves p at its input, so the
    result must be the Euclidean projection of p0 onto {|p| <= lam}:
    p0 where |p0| <= lam, lam * p0 / |p0| elsewhere.  lam < 1 separates the
    right threshold (|p|^2 > lam^2) from |p|^2 > lam or |p| > lam^2, and the
    scaled copy fixes the factor lam / sqrt(|p|^2) (not lam / |p|^2)."""
    lam = 0.05
    rng = np.random.default_rng(21)
    sh = SHAPE
    d = rng.standard_normal(sh + (3,))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    # magnitudes across [0.3 lam, 3 lam], away from the rounding band at lam
    r = lam * rng.uniform(0.3, 3.0, sh)
    r[np.abs(r / lam - 1) < 1e-4] = 0.5 * lam
    p0 = (d * r[..., None]).astype(dt)
    _, p = _tv_dual_run(p0, lam, dt)
    it = interior(sh)
    p0d = p0.astype(np.float64)
    n0 = np.linalg.norm(p0d, axis=-1)
    inside = it & (n0 < lam)
    outside = it & (n0 > lam)
    assert inside.sum() > 20 and outside.sum() > 20
    # untouched below the threshold, bit for bit
    np.testing.assert_array_equal(p[inside], p0[inside])
    # projected onto the sphere of radius lam along p0 above it
    want = p0d[outside] * (lam / n0[outside])[:, None]
    np.testing.assert_allclose(p[outside].astype(np.float64), want, rtol=4 * rt, atol=0)
    pn = np.linalg.norm(p.astype(np.float64), axis=-1)
    assert pn[it].max() <= lam * (1 + 4 * rt)
    np.testing.assert_allclose(pn[outside], lam, rtol=4 * rt)
    # voxels outside the dual mask are not visited
    np.testing.assert_array_equal(p[~it], p0[~it])


@pytest.mark.parametrize("dt,rt", [(np.float64, 1e-13), (np.float32, 1e-5)])
def test_tv_dual_projection_after_gradient_step(dt, rt):
    """A finite lam on a random chi_: every |p| <= lam, projected vectors are
    the unprojected update scaled by lam / |update|, the rest and nu are
    those of the run without projection (lam = 1e300)."""
    lam = 0.05
    rng = np.random.default_rng(22)
    cb = 0.3 * rng.standard_normal(SHAPE)
    p0 = 0.01 * rng.standard_normal(SHAPE + (3,))
    nu_f, p_free = _tv_dual_run(p0, 1e300, dt, cb)
    nu, p = _tv_dual_run(p0, lam, dt, cb)
    np.testing.assert_array_equal(nu, nu_f)
    it = interior(SHAPE)
    pf = np.linalg.norm(p_free.astype(np.float64), axis=-1)
    big = it & (pf > lam * (1 + 1e-4))
    small = it & (pf < lam * (1 - 1e-4))
    assert big.sum() > 20 and small.sum() > 20
    np.testing.assert_array_equal(p[small], p_free[small])
    np.testing.assert_allclose(p[big].astype(np.float64),
                               p_free[big].astype(np.float64) * (lam / pf[big])[:, None],
                               rtol=rt)
    assert np.linalg.norm(p.astype(np.float64), axis=-1)[it].max() <= lam * (1 + rt)


# ---------------------------------------------------------------------------
# convergence norms and flag (pd_tgv_mex.c:980-1045, pd_tv_mex.c:730-795)
# ---------------------------------------------------------------------------

def _conv(chi, chi0, eta, eta0, tol, dt=np.float64):
    a = [np.ascontiguousarray(x, dtype=dt) for x in (chi, chi0, eta, eta0)]
    return orc.convergence(*a, tol)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_convergence_closed_forms(dt):
    """n1 = ||chi - chi0|| / ||chi||, n2 = ||eta - eta0|| / ||eta||, on
    values where the ratios are exact: chi0 = 0 gives 1, chi0 = 3 chi / 4
    gives 1/4 (Sterbenz: the difference is exact), eta0 = eta gives 0.  A
    swapped numerator / denominator gives 4 instead of 1/4; swapped n1 / n2
    is caught by the distinct values."""
    rng = np.random.default_rng(23)
    # short significands (11 bits) so that 3 chi / 4 is exact in either class
    chi = rng.integers(512, 2048, 1000) / 1024.0 * np.where(rng.random(1000) < 0.5, -1, 1)
    eta = rng.integers(512, 2048, 1000) / 1024.0
    chi = chi.astype(dt)
    eta = eta.astype(dt)
    n1, n2, flag = _conv(chi, np.zeros_like(chi), eta, eta, 0.0, dt)
    assert n1 == 1.0 and n2 == 0.0 and flag == 0
    n1, n2, _ = _conv(chi, chi * dt(0.75), eta, np.zeros_like(eta), 0.0, dt)
    assert n1 == 0.25 and n2 == 1.0
    # 0/0: both NaN, never a stop (pd_tgv_mex.c:1006-1010)
    z = np.zeros(10, dt)
    n1, n2, flag = _conv(z, z, z, z, 1.0, dt)
    assert np.isnan(n1) and np.isnan(n2) and flag == 0


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_convergence_flag_thresholds(dt):
    """flag = (n1 < tol) and (n2 < 2 tol), strict, checked on both sides of
    each threshold with n1 = 1/4 and n2 = 3/4."""
    rng = np.random.default_rng(24)
    # short significands: 3 chi / 4, chi / 4 and 3 eta / 4 are exact in either class
    chi = (rng.integers(512, 2048, 500) / 1024.0).astype(dt)
    eta = (4 * rng.integers(1, 1 << 12, 500)).astype(dt)
    c0 = chi * dt(0.75)
    e0 = eta * dt(0.25)
    n1, n2, _ = _conv(chi, c0, eta, e0, 0.0, dt)
    assert n1 == 0.25 and n2 == 0.75
    # n1 < tol and n2 < 2 tol
    assert _conv(chi, c0, eta, e0, 0.40, dt)[2] == 1     # 0.25 < 0.40, 0.75 < 0.80
    assert _conv(chi, c0, eta, e0, 0.376, dt)[2] == 1    # 0.75 < 0.752
    assert _conv(chi, c0, eta, e0, 0.374, dt)[2] == 0    # 0.75 >= 0.748
    # the n1 threshold alone (eta unchanged: n2 = 0)
    assert _conv(chi, c0, eta, eta, 0.2501, dt)[2] == 1
    assert _conv(chi, c0, eta, eta, 0.2499, dt)[2] == 0
    assert _conv(chi, c0, eta, eta, 0.25, dt)[2] == 0    # strict: n1 == tol is no stop
    # the n2 threshold is strict too: n2 = 3/4 at 2 tol = 3/4
    assert _conv(chi, chi, eta, e0, 0.375, dt)[2] == 0


def test_convergence_random_against_direct_sums():
    """Random data: the norms equal the square roots of the plain sums (any
    summation order agrees to rounding)."""
    rng = np.random.default_rng(25)
    chi, chi0, eta, eta0 = rng.standard_normal((4, 4096))
    n1, n2, _ = _conv(chi, chi0, eta, eta0, 0.0)
    assert n1 == pytest.approx(np.sqrt(((chi - chi0) ** 2).sum() / (chi ** 2).sum()), rel=1e-13)
    assert n2 == pytest.approx(np.sqrt(((eta - eta0) ** 2).sum() / (eta ** 2).sum()), rel=1e-13)


def test_convergence_single_path_accumulates_in_double():
    """The single path sums in double and forms each difference term as
    float(chi - chi0) * (double(chi) - chi0) (pd_tgv_mex.c:1031-1036).

    * d1 = 1 + 2^-26 must not round to 1 (a float accumulator would):
      chi = (1, 2^-13), chi0 = (1, 0) gives n1 = sqrt(2^-26 / (1 + 2^-26)).
    * chi = 1, chi0 = 2^-30: float(1 - 2^-30) = 1 while the double
      difference is exact, so n1^2 = 1 - 2^-30 (all-double would give
      (1 - 2^-30)^2, all-float 1)."""
    f32 = np.float32
    one = np.ones(1, f32)
    n1, _, _ = _conv(np.array([1.0, 2.0 ** -13], f32), np.array([1.0, 0.0], f32),
                     np.array([1.0, 1.0], f32), np.array([1.0, 1.0], f32), 0.0, f32)
    assert n1 == np.sqrt(2.0 ** -26 / (1.0 + 2.0 ** -26))
    assert n1 != 2.0 ** -13
    n1, _, _ = _conv(one, np.array([2.0 ** -30], f32), one, one, 0.0, f32)
    assert n1 == np.sqrt(1.0 - 2.0 ** -30)
    assert n1 != 1.0 and n1 != 1.0 - 2.0 ** -30


# ---------------------------------------------------------------------------
# whole loop
# ---------------------------------------------------------------------------

def test_zero_data_stays_zero():
    m = _smooth_mask((10, 10, 10))
    for dt in (np.float64, np.float32):
        chi, it, nr = orc.pd_tgv(zeros((10, 10, 10), dt=dt), m, (1, 1, 1), 1e-3, 2.0, 0.0, 30)
        assert it == 30 and np.all(chi == 0)
        chi, it, nr = orc.pd_tv(zeros((10, 10, 10), dt=dt), m, (1, 1, 1), 1e-3, 0.0, 30)
        assert np.all(chi == 0)
        assert np.isnan(nr[0])  # 0/0 norms never stop the loop (F8)


def test_tol_zero_runs_every_iteration_and_tol_stops_on_multiple_of_ten():
    rng = np.random.default_rng(8)
    m = _smooth_mask((12, 12, 12))
    f = F(rng.standard_normal((12, 12, 12)) * 1e-2)
    _, it, _ = orc.pd_tgv(f, m, (1, 1, 1), 1e-3, 2.0, 0.0, 57)
    assert it == 57
    _, it, nr = orc.pd_tgv(f, m, (1, 1, 1), 1e-3, 2.0, 1e-1, 500)
    assert it % 10 == 0 and it < 500
    assert nr[0] < 1e-1 and nr[1] < 2e-1


def test_maxiter_zero_rejected():
    with pytest.raises(ValueError):
        orc.pd_tgv(zeros((5, 5, 5)), np.ones((5, 5, 5)), (1, 1, 1), 1e-3, 2.0, 0.0, 0)


def _run_steps(f, m, h, lam, alpha, iters, tgv=True):
    sh = f.shape
    mi = orc.init_mask(m)
    tau, sigma = orc.step_sizes(h)
    fs = F(sigma * f)
    chi, eta, w = zeros(sh), zeros(sh), zeros(sh, 3)
    cb, eb, wb = zeros(sh), zeros(sh), zeros(sh, 3)
    nu, p, q = zeros(sh), zeros(sh, 3), zeros(sh, 6)
    for _ in range(iters):
        if tgv:
            orc.tgv_dual(nu, p, q, cb, eb, wb, fs, mi, h, alpha * lam, lam, sigma)
            orc.tgv_primal(chi, eta, w, cb, eb, wb, nu, p, q, m, h, tau)
        else:
            orc.tv_dual(nu, p, cb, eb, fs, mi, h, lam, sigma)
            orc.tv_primal(chi, eta, cb, eb, nu, p, m, h, tau)
    return chi, eta


def _constraint_residual(chi, eta, f, m, h):
    """|Delta eta - W chi + f| on the interior mask, relative to |f|."""
    mi = orc.init_mask(m).astype(bool)
    def d2(u, ax):
        return (np.roll(u, -1, ax) - 2 * u + np.roll(u, 1, ax)) / h[ax] ** 2
    lap = d2(eta, 0) + d2(eta, 1) + d2(eta, 2)
    wave = -d2(chi, 0) / 3 - d2(chi, 1) / 3 + 2 * d2(chi, 2) / 3
    r = (lap - wave + f)[mi]
    return np.linalg.norm(r) / np.linalg.norm(f[mi])


def test_loop_converges_to_the_constraint():
    """The dual variable nu is the multiplier of Delta(eta) - W(chi) + f = 0;
    the iterates approach that constraint (a sign error in any coupling
    term makes the residual grow instead)."""
    shape, h = (14, 14, 14), (1.0, 1.0, 1.0)
    rng = np.random.default_rng(9)
    m = _smooth_mask(shape)
    f = F(rng.standard_normal(shape) * 1e-2 * m)
    chi, eta = _run_steps(f, m, h, 1e-4, 2.0, 3000)
    r = _constraint_residual(chi, eta, f, m, h)
    chi0, eta0 = _run_steps(f, m, h, 1e-4, 2.0, 300)
    r0 = _constraint_residual(chi0, eta0, f, m, h)
    assert r < 0.05 and r < 0.5 * r0, (r0, r)


def test_step_loop_equals_whole_solve():
    """The building-block sweeps, driven like mx_pdd, reproduce the whole
    solve bit for bit (f scaling by sigma, dual then primal)."""
    shape, h = (10, 11, 12), (0.9, 0.9, 1.2)
    rng = np.random.default_rng(10)
    m = _smooth_mask(shape)
    f = F(rng.standard_normal(shape) * 1e-2)
    chi, _ = _run_steps(f, m, h, 2e-3, 2.0, 25)
    ref, it, _ = orc.pd_tgv(f, m, h, 2e-3, 2.0, 0.0, 25)
    np.testing.assert_array_equal(chi, ref)
    chi, _ = _run_steps(f, m, h, 2e-3, 0.0, 25, tgv=False)
    ref, it, _ = orc.pd_tv(f, m, h, 2e-3, 0.0, 25)
    np.testing.assert_array_equal(chi, ref)


def test_single_path_tracks_double_path():
    shape, h = (12, 12, 12), (1.0, 1.0, 1.0)
    rng = np.random.default_rng(11)
    m = _smooth_mask(shape)
    f = F(rng.standard_normal(shape) * 1e-2)
    c64, _, _ = orc.pd_tgv(f, m, h, 1e-3, 2.0, 0.0, 200)
    c32, _, _ = orc.pd_tgv(f.astype(np.float32, order="F"), m, h, 1e-3, 2.0, 0.0, 200)
    assert np.linalg.norm(c32 - c64) / np.linalg.norm(c64) < 1e-4
