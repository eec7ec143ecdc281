"""mgpcg.m (masked Poisson solve, multigrid-preconditioned CG) and
unwrapLaplacian.m on the GPU against the oracle (oracle/poisson.py: the
same discrete system solved exactly by a sparse direct solver).

mgpcg returns the discrete solution to its stopping rule ||r|| <= tol ||f||,
so the bar is that rule: the true relative residual of the GPU solution is
within a small factor of tol, and its distance to the exact solution is
within the conditioning of the problem (FP64, tol 1e-12: relative L2
<= 1e-8; FP32, tol 1e-5: <= 1e-3).  A working multigrid preconditioner
converges in few iterations, nearly independent of the size; that count is
pinned too (it fails for a wrong coarse-level stencil or transfer).
Unwrapping a wrapped phase whose steps stay below pi must give the phase
itself; d2phas is lapw's output, bitwise the oracle's."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as orc
from oracle import poisson as po
from qsm_amd import QsmError
from qsm_amd import unwrap as uw

pytestmark = pytest.mark.gpu


def _ball(shape, r, hole=False, offset=(0, 0, 0)):
    c = [(n - 1) / 2 + o for n, o in zip(shape, offset)]
    g = np.meshgrid(*(np.arange(n) - c[a] for a, n in enumerate(shape)), indexing="ij")
    m = (g[0] ** 2 + g[1] ** 2 + g[2] ** 2) <= r * r
    if hole:
        m &= ((g[0] - 2) ** 2 + g[1] ** 2 + (g[2] + 1) ** 2) > 9
    return np.asfortranarray(m)


MASKS = {
    "ball": lambda: ((40, 36, 30), _ball((40, 36, 30), 13)),
    "ball_hole": lambda: ((33, 29, 27), _ball((33, 29, 27), 12, hole=True)),
    # touches the x faces: the solver's own zero border is the boundary there
    "touching": lambda: ((24, 30, 26), _ball((24, 30, 26), 13, offset=(-2, 0, 0))),
    "box": lambda: ((17, 16, 15), np.ones((17, 16, 15), bool)),
}


def _true_relres(v, f, m, vsz):
    A, idx = po.neg_laplacian(m, vsz)
    fv = np.asarray(f, np.float64).reshape(-1, order="F")
    vv = np.asarray(v, np.float64).reshape(-1, order="F")
    return np.linalg.norm(A @ vv[idx] - fv[idx]) / np.linalg.norm(fv)


@pytest.mark.parametrize("vsz", [(1.0, 1.0, 1.0), (0.8, 1.0, 1.4)])
@pytest.mark.parametrize("name", list(MASKS))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_mgpcg_matches_exact_solve(cuda, name, vsz, dt):
    shape, m = MASKS[name]()
    rng = np.random.default_rng(sum(map(ord, name)))
    f = np.asfortranarray((rng.standard_normal(shape) * m).astype(dt))
    want = po.solve(f, m, vsz)
    tol = 1e-12 if dt == np.float64 else 1e-5
    v, (it, res) = uw.mgpcg(f, m, vsz, tol, 200, return_info=True)
    assert v.dtype == dt and v.shape == shape
    assert np.all(v[~m] == 0)
    assert 1 <= it < 200 and res[-1] <= tol * np.linalg.norm(f.astype(np.float64)) * 1.0001
    if dt == np.float64:
        assert _true_relres(v, f, m, vsz) < 10 * tol
        assert rel_l2(v, want) < 1e-8
        assert it <= 25, it          # multigrid-preconditioned: few iterations
    else:
        assert _true_relres(v, f, m, vsz) < 1e-4
        assert rel_l2(v, want) < 1e-3


@pytest.mark.parametrize("mg", [{"nlevels": 1}, {"nlevels": 2, "mu": 2}, {"tol": 1e-2, "maxit": 4},
                                {"npre": 3, "npost": 0, "nboundary": 0}])
def test_mgpcg_mgopts(cuda, mg):
    """One level (the coarse CG alone is the preconditioner), a W-cycle,
    tolerance-stopped cycles and unbalanced smoothing all converge to the
    same solution."""
    shape, m = MASKS["ball_hole"]()
    f = np.asfortranarray(np.random.default_rng(3).standard_normal(shape) * m)
    v, (it, _) = uw.mgpcg(f, m, (1, 1, 1), 1e-12, 300, mg, return_info=True)
    assert it < 300
    assert rel_l2(v, po.solve(f, m, (1, 1, 1))) < 1e-8


def test_mgpcg_defaults(cuda):
    """Default tol sqrt(eps) and maxit numel(f)."""
    shape, m = MASKS["ball"]()
    f = np.asfortranarray(np.random.default_rng(4).standard_normal(shape) * m)
    v, (it, res) = uw.mgpcg(f, m, (1, 1, 1), return_info=True)
    assert rel_l2(v, po.solve(f, m, (1, 1, 1))) < 1e-5
    assert res[-1] <= np.sqrt(np.finfo(np.float64).eps) * np.linalg.norm(f)
    assert it < 20


def test_mgpcg_off_mask_rhs(cuda):
    """A right-hand side nonzero off the mask: the solution is that of
    f .* mask, but ||f|| and ||r|| count the whole array (mgpcg_mex.c:
    mx_pcg), so the rule never stops it: it runs to maxit, or stops at the
    breakdown once the residual on the mask is exactly zero, with a finite
    result."""
    shape, m = MASKS["ball"]()
    f = np.asfortranarray(np.random.default_rng(4).standard_normal(shape))
    v, (it, res) = uw.mgpcg(f, m, (1, 1, 1), 1e-10, 40, return_info=True)
    assert np.all(np.isfinite(v))
    # iterations past convergence run at rounding level
    assert rel_l2(v, po.solve(f * m, m, (1, 1, 1))) < 1e-6
    assert res[-1] > 1e-10 * np.linalg.norm(f)
    assert np.all(v[~m] == 0)


def test_mgpcg_zero_and_empty(cuda):
    shape, m = MASKS["box"]()
    assert np.all(uw.mgpcg(np.zeros(shape), m, (1, 1, 1)) == 0)
    f = np.ones(shape)
    assert np.all(uw.mgpcg(f, np.zeros(shape, bool), (1, 1, 1), maxit=5) == 0)


def _wrapped(shape, amp, dt):
    phi = np.asfortranarray(
        amp * np.prod(np.meshgrid(*(np.sin(np.pi * np.arange(n) / (n - 1)) for n in shape),
                                  indexing="ij"), axis=0))
    return phi, np.asfortranarray(np.angle(np.exp(1j * phi)).astype(dt))


@pytest.mark.parametrize("vsz", [(1.0, 1.0, 1.0), (0.8, 1.0, 1.3)])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_unwrap_recovers_phase(cuda, dt, vsz):
    shape = (34, 30, 26)
    phi, w = _wrapped(shape, 12.0, dt)
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    up, d2 = uw.unwrap_laplacian(w, m, vsz)
    assert up.dtype == dt and d2.dtype == dt
    np.testing.assert_array_equal(d2, orc.lapw(w, vsz))
    atol = 1e-6 if dt == np.float64 else 2e-3
    np.testing.assert_allclose(up, phi, atol=atol)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_unwrap_matches_oracle_on_a_ball_multi_echo(cuda, dt):
    """A ball mask inside a larger volume (the crop to its box grown by one
    voxel), two echoes: per echo the oracle's exact solve; each echo equals
    a single-echo call bitwise."""
    shape = (44, 40, 36)
    m = _ball(shape, 13, hole=True, offset=(2, -1, 0))
    phi1, w1 = _wrapped(shape, 9.0, dt)
    rng = np.random.default_rng(5)
    w2 = np.asfortranarray(np.angle(np.exp(1j * (3 * phi1 + 0.3 * rng.standard_normal(shape))))
                           .astype(dt))
    p = np.stack([w1, w2], axis=3)
    up, d2 = uw.unwrap_laplacian(p, m, (1.0, 1.1, 0.9))
    want_u, want_d = po.unwrap_laplacian(p.astype(np.float64), m, (1.0, 1.1, 0.9))
    for t in range(2):
        lim = 1e-6 if dt == np.float64 else 2e-3
        assert rel_l2(up[..., t], want_u[..., t]) < lim
        one_u, one_d = uw.unwrap_laplacian(p[..., t], m, (1.0, 1.1, 0.9))
        np.testing.assert_array_equal(up[..., t], one_u)
        np.testing.assert_array_equal(d2[..., t], one_d)
    if dt == np.float64:
        np.testing.assert_array_equal(d2, want_d)
    assert np.all(up[~np.broadcast_to(m[..., None], up.shape)] == 0)


def test_unwrap_empty_mask_and_checks(cuda):
    shape = (12, 11, 10)
    up, d2 = uw.unwrap_laplacian(np.ones(shape), np.zeros(shape), (1, 1, 1))
    assert np.all(up == 0) and np.all(d2 == 0)
    with pytest.raises(ValueError):
        uw.unwrap_laplacian(np.full(shape, np.nan), np.ones(shape), (1, 1, 1))
    with pytest.raises(ValueError):
        uw.unwrap_laplacian(np.ones(shape), np.full(shape, 2), (1, 1, 1))


@pytest.mark.parametrize("dt", ["float64", "float32"])
def test_unwrap_device_form_on_side_stream(cuda, dt):
    """qsm_unwrap_laplacian_dev on a torch side stream with a dirty work
    space, twice: bitwise the host form (the mask spans the volume's
    interior, so the host form's crop, its box grown by one voxel, is the
    whole volume)."""
    import torch

    shape = (30, 28, 26)
    _, w = _wrapped(shape, 10.0, dt)
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = ~_ball((28, 26, 24), 5)
    want_u, want_d = uw.unwrap_laplacian(w, m, (1, 1, 1))

    def dev(a):
        t = torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to("cuda")
        return t.view(shape[::-1]).permute(2, 1, 0)

    def host(t):
        return t.permute(2, 1, 0).contiguous().cpu().numpy().reshape(-1).reshape(shape, order="F")

    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        tu, tm = dev(w), dev(m.astype(np.uint8))
        up, d2 = torch.empty_like(tu), torch.empty_like(tu)
        work = torch.full((uw.work_bytes(tu.dtype, shape),), 255, dtype=torch.uint8, device="cuda")
    for _ in range(2):
        it = uw.unwrap_laplacian_dev(tu, tm, (1, 1, 1), up, d2, work=work, stream=st)
        st.synchronize()
        assert it >= 1
        np.testing.assert_array_equal(host(up), want_u)
        np.testing.assert_array_equal(host(d2), want_d)


def test_unwrap_rejects_bad_vsz(cuda):
    with pytest.raises(ValueError):
        uw.unwrap_laplacian(np.zeros((8, 8, 8)), np.ones((8, 8, 8)), (1, 0, 1))
    with pytest.raises(QsmError):
        uw.mgpcg(np.zeros((8, 8, 8)), np.ones((8, 8, 8)), (1, 1, 1), mgopts={"nlevels": 20})


def test_mgpcg_manufactured_solution_second_order(cuda):
    """The reference's model problem (testPoissonSolvers.m: modelProblem):
    u = sin(pX) sin(pY) sin(pZ) on the unit cube sampled at h = 1/n, the
    interior active, f = 3 p^2 u.  The discrete solution converges to u at
    second order in h: each halving of h divides the error by ~4 (a wrong
    voxel-size scaling or stencil would not)."""
    p = 2 * np.pi
    errs = []
    for n in (16, 32, 64):
        h = 1.0 / n
        x = np.arange(n + 1) * h
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        u = np.sin(p * X) * np.sin(p * Y) * np.sin(p * Z)
        G = np.zeros(u.shape, bool)
        G[1:-1, 1:-1, 1:-1] = True
        f = np.asfortranarray(3 * p * p * u * G)
        v = uw.mgpcg(f, G, (h, h, h), 1e-12, 100)
        errs.append(np.linalg.norm((v - u)[G]) / np.sqrt(G.sum()))
    ratios = [errs[i] / errs[i + 1] for i in range(2)]
    assert all(3.6 < r < 4.4 for r in ratios), (errs, ratios)
