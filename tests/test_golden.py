"""The committed golden vectors (tests/golden/*.npz, written by
tests/golden/make_golden.py) against the oracle and the numpy mirrors:
the oracle must keep reproducing them bitwise."""
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc
from qsm_amd import fastfft

G = Path(__file__).resolve().parent / "golden"
H = (0.9, 1.0, 1.1)


@pytest.fixture(scope="module")
def pd():
    return np.load(G / "pd.npz")


@pytest.fixture(scope="module")
def fdg():
    return np.load(G / "fd.npz")


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype
    np.testing.assert_array_equal(np.ascontiguousarray(a).view(np.uint8),
                                  np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize("dt,tag", [(np.float64, "f64"), (np.float32, "f32")])
def test_oracle_reproduces_pd_vectors(pd, dt, tag):
    f = np.asfortranarray(pd["f"].astype(dt))
    chi, it, nr = orc.pd_tgv(f, pd["mask"], H, 2e-3, 2.0, 0.0, 40)
    _same(chi, pd[f"tgv_{tag}_chi"])
    # the norms are OpenMP reductions (order varies with the thread count)
    assert it == 40
    np.testing.assert_allclose(np.array(nr), pd[f"tgv_{tag}_norms"], rtol=1e-12)
    chi, it, nr = orc.pd_tv(f, pd["mask"], H, 3e-3, 0.0, 40)
    _same(chi, pd[f"tv_{tag}_chi"])


def test_oracle_reproduces_tolerance_stop(pd):
    chi, it, nr = orc.pd_tgv(pd["f"], pd["mask"], H, 2e-3, 2.0, 2e-2, 400)
    assert it == int(pd["tgv_tol_iters"]) and it % 10 == 0 and it < 400
    _same(chi, pd["tgv_tol_chi"])


@pytest.mark.parametrize("dt,tag", [(np.float64, "f64"), (np.float32, "f32")])
def test_oracle_reproduces_fd_vectors(fdg, dt, tag):
    U, X, Y, Z, P = (np.asfortranarray(fdg[k].astype(dt)) for k in ("u", "x", "y", "z", "phase"))
    Gm = fdg["G"]
    for name, res in (("gradfp", orc.gradfp(U, H)), ("gradf", orc.gradf(U, H)),
                      ("gradfm", orc.gradfm(U, Gm, H))):
        for c, a in zip("xyz", res):
            _same(a, fdg[f"{name}_{tag}_d{c}"])
    _same(orc.gradfp_adj(X, Y, Z, H), fdg[f"gradfp_adj_{tag}"])
    _same(orc.gradf_adj(X, Y, Z, H), fdg[f"gradf_adj_{tag}"])
    _same(orc.gradfm_adj(X, Y, Z, Gm, H), fdg[f"gradfm_adj_{tag}"])
    _same(orc.lapw(P, H), fdg[f"lapw_{tag}"])


def test_numpy_mirrors_reproduce_fft_vectors():
    d = np.load(G / "fft.npz")
    np.testing.assert_allclose(np.fft.fftn(d["fft_x"]), d["fft_X"], rtol=0, atol=1e-12)
    pad = tuple(int(p) for p in d["dip_pad"])
    up = fastfft.padfastfft(d["dip_chi"], pad)
    D = fastfft.dipole_kernel(up.shape, tuple(d["dip_vsz"]), tuple(d["dip_bdir"]))
    _same(D, d["dip_D"])
