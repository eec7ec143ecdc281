"""qsm_pd_solve_static, the allocation-free solve: chi bitwise equal to the
drop-in entry points qsm_pd_tgv / qsm_pd_tv on two calls in a row (the
second with a smaller volume, so the static work space is dirty from the
first), a tolerance stop, refusal of volumes beyond the work space, and the
call captured into a graph and replayed with other work in between."""
import numpy as np
import pytest

from helpers import brainish_mask, from_dev, to_dev
from qsm_amd import QsmError, _lib, pd
from test_gpu_pd import _phantom_input

pytestmark = pytest.mark.gpu


def _dropin(kind, d2f, m, vsz, lam, alpha, tol, iters):
    if kind == _lib.TGV:
        return pd.pd_tgv_mex(d2f, m, vsz, lam, alpha, tol, iters, 0)
    return pd.pd_tv_mex(d2f, m, vsz, lam, tol, iters, 0)


def _static(kind, d2f, m, vsz, lam, alpha, tol, iters, stream=None):
    import torch

    fd, md = to_dev(d2f), to_dev(m)
    chi = torch.full_like(fd, float("nan"))
    pd.solve_static(kind, fd, md, d2f.shape, vsz, lam, alpha, tol, iters, chi, stream)
    torch.cuda.synchronize()
    return from_dev(chi, d2f.shape)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("kind", [_lib.TGV, _lib.TV])
def test_two_calls_equal_dropin(cuda, dt, kind):
    """Call 1 at 40x36x30, call 2 at 22x20x18 (smaller, on the work space
    call 1 left behind): both bitwise the drop-in's chi."""
    print(f"[no_alloc] two calls kind={kind} dt={np.dtype(dt).name}", flush=True)
    for i, (shape, vsz, iters) in enumerate((((40, 36, 30), (0.9, 0.9, 1.2), 47),
                                             ((22, 20, 18), (1.0, 1.0, 1.0), 33))):
        d2f = _phantom_input(shape, vsz, 60 + i, dt)
        m = brainish_mask(shape, seed=61 + i, frac=0.85)
        want = _dropin(kind, d2f, m, vsz, 2e-3, 2.0, 0.0, iters)
        got = _static(kind, d2f, m, vsz, 2e-3, 2.0, 0.0, iters)
        assert np.abs(want).max() > 0
        np.testing.assert_array_equal(got, want, err_msg=f"call {i + 1} {shape}")


def test_tolerance_stop_equals_dropin(cuda):
    shape, vsz = (24, 22, 20), (1.0, 1.0, 1.0)
    d2f = _phantom_input(shape, vsz, 70, np.float64)
    m = brainish_mask(shape, seed=71, frac=0.9)
    print("[no_alloc] tolerance stop", flush=True)
    want = _dropin(_lib.TGV, d2f, m, vsz, 1e-3, 2.0, 5e-3, 2000)
    got = _static(_lib.TGV, d2f, m, vsz, 1e-3, 2.0, 5e-3, 2000)
    np.testing.assert_array_equal(got, want)


def test_refuses_volume_beyond_work_space(cuda):
    import torch

    shape = (80, 80, 80)  # FP64 TGV: about 131 MB of state
    n = int(np.prod(shape))
    d = torch.zeros(n, dtype=torch.float64, device="cuda")
    m = torch.ones(n, dtype=torch.uint8, device="cuda")
    c = torch.zeros_like(d)
    with pytest.raises(QsmError, match="static work space") as e:
        pd.solve_static(_lib.TGV, d, m, shape, (1, 1, 1), 1e-3, 2.0, 0.0, 10, c)
    assert e.value.code == _lib.QSM_EARG
    # a smaller volume of the same kind still runs afterwards
    shape = (20, 20, 20)
    d2f = _phantom_input(shape, (1, 1, 1), 72, np.float64)
    mk = brainish_mask(shape, seed=73, frac=0.9)
    np.testing.assert_array_equal(_static(_lib.TGV, d2f, mk, (1, 1, 1), 1e-3, 2.0, 0.0, 12),
                                  _dropin(_lib.TGV, d2f, mk, (1, 1, 1), 1e-3, 2.0, 0.0, 12))


def test_captured_in_a_graph_and_replayed(cuda):
    """The call issues kernels only, so it can be captured; every replay,
    with a different eager solve (which overwrites the work space) and
    eager work on chi in between, gives the drop-in's chi."""
    import torch

    shape, vsz, iters = (30, 28, 24), (1.0, 1.0, 1.0), 40
    d2f = _phantom_input(shape, vsz, 74, np.float32)
    m = brainish_mask(shape, seed=75, frac=0.85)
    want = _dropin(_lib.TV, d2f, m, vsz, 2e-3, 0.0, 0.0, iters)
    other = _phantom_input((26, 25, 23), vsz, 76, np.float32)
    om = brainish_mask((26, 25, 23), seed=77, frac=0.85)
    other_want = _dropin(_lib.TGV, other, om, vsz, 2e-3, 2.0, 0.0, 20)
    fd, md = to_dev(d2f), to_dev(m)
    chi = torch.zeros_like(fd)
    side = torch.cuda.Stream()
    # first use on this device outside the capture (work-space lookup)
    with torch.cuda.stream(side):
        pd.solve_static(_lib.TV, fd, md, shape, vsz, 2e-3, 0.0, 0.0, iters, chi, side)
    side.synchronize()
    np.testing.assert_array_equal(from_dev(chi, shape), want)
    print("[no_alloc] capturing", flush=True)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        st = torch.cuda.current_stream()
        pd.solve_static(_lib.TV, fd, md, shape, vsz, 2e-3, 0.0, 0.0, iters, chi, st)
    torch.cuda.synchronize()
    for r in range(3):
        print(f"[no_alloc] replay {r}", flush=True)
        chi.fill_(float("nan"))
        got_other = _static(_lib.TGV, other, om, vsz, 2e-3, 2.0, 0.0, 20)
        np.testing.assert_array_equal(got_other, other_want)
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(from_dev(chi, shape), want, err_msg=f"replay {r}")
