"""The drop-in on the benchmark configurations themselves (synthetic
phantoms of configs.py): FP64 chi bitwise equal to the oracle's at C1 (64^3
TV, its full 500 iterations) and C2 (224x224x128 TGV, 0.9x0.9x1.2 mm, B0
tilted 15 degrees, ellipsoidal mask; a bounded 60 iterations so that the
CPU oracle finishes in seconds), and the FP32 path within 1e-4 relative L2
of the FP64 chi."""
import numpy as np
import pytest

from oracle import oracle as orc
from qsm_amd import configs, fd, pd, phantom

pytestmark = pytest.mark.gpu


def _input(name):
    cfg = configs.CONFIGS[name]
    ph = phantom.make(cfg["spec"], field="gpu")
    d2f = np.asfortranarray(-fd.lapw(ph["field"], cfg["spec"].vsz))
    return cfg, d2f, ph["mask"].astype(np.uint8, order="F")


def _rel(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))


def test_c1_tv_full_solve(cuda):
    cfg, d2f, m = _input("c1")
    vsz, lam, iters = cfg["spec"].vsz, cfg["lam"], cfg["iters"]
    ref, it, _ = orc.pd_tv(d2f, m, vsz, lam, 0.0, iters)
    got = pd.pd_tv_mex(d2f, m, vsz, lam, 0.0, iters, 0)
    np.testing.assert_array_equal(got, ref)
    got32 = pd.pd_tv_mex(d2f.astype(np.float32, order="F"), m, vsz, lam, 0.0, iters, 0)
    assert _rel(got32, ref) <= 1e-4


def test_c2_tgv_dropin(cuda):
    cfg, d2f, m = _input("c2")
    vsz, lam, alpha, iters = cfg["spec"].vsz, cfg["lam"], cfg["alpha"], 60
    ref, it, _ = orc.pd_tgv(d2f, m, vsz, lam, alpha, 0.0, iters)
    got = pd.pd_tgv_mex(d2f, m, vsz, lam, alpha, 0.0, iters, 0)
    np.testing.assert_array_equal(got, ref)
    got32 = pd.pd_tgv_mex(d2f.astype(np.float32, order="F"), m, vsz, lam, alpha, 0.0, iters, 0)
    assert _rel(got32, ref) <= 1e-4
