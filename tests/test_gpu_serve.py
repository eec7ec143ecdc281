"""Serving: requests of mixed sizes arrive while others run; each result is
bitwise the same as the request solved alone (the oracle's solve)."""
import numpy as np
import pytest

from helpers import brainish_mask
from oracle import oracle as orc
from qsm_amd import _lib
from qsm_amd.batch import Batch
from qsm_amd.serve import Request, Server, poisson_arrivals
from test_gpu_pd import _phantom_input

pytestmark = pytest.mark.gpu


def test_serving_results_equal_solo_solves(cuda):
    rng = np.random.default_rng(5)
    shapes = [(28, 24, 20), (36, 30, 18), (22, 22, 22), (40, 20, 16), (30, 30, 24), (26, 34, 21)]
    arrivals = poisson_arrivals(50.0, len(shapes), rng)
    reqs = []
    for i, (shape, t) in enumerate(zip(shapes, arrivals)):
        kind = _lib.TGV if i % 2 == 0 else _lib.TV
        vsz = (1.0, 1.0, 1.0) if i % 3 else (0.9, 0.9, 1.2)
        d2f = _phantom_input(shape, vsz, 40 + i, np.float32)
        m = brainish_mask(shape, seed=50 + i, frac=0.9)
        iters = int(rng.integers(30, 90))
        reqs.append(Request(i, float(t), kind, d2f, m, vsz, 2e-3, 2.0, iters, kz=8))
    # a budget that holds about three of these at once forces queueing
    with Batch(np.float32, max_volumes=8) as probe:
        one = probe.volume_bytes(_lib.TGV, (40, 30, 24), 8, 100)
    with Batch(np.float32, hbm_budget=3 * one, max_volumes=8) as b:
        stats = Server(b, chunk=10).run(reqs)
    assert stats.n == len(reqs) and stats.p99 >= stats.p50 > 0
    for r in reqs:
        if r.kind == _lib.TGV:
            ref, it, _ = orc.pd_tgv(r.d2f, r.mask, r.vsz, r.lam, r.alpha, 0.0, r.maxiter)
        else:
            ref, it, _ = orc.pd_tv(r.d2f, r.mask, r.vsz, r.lam, 0.0, r.maxiter)
        assert r.iters == it == r.maxiter
        np.testing.assert_array_equal(r.chi, ref, err_msg=f"request {r.rid}")


def test_serving_with_live_voxel_cap(cuda):
    """With the live work capped at two volumes' worth, requests still match
    the oracle bitwise, at most two run at once, and they finish in arrival
    order (equal budgets)."""
    rng = np.random.default_rng(11)
    shape = (26, 24, 20)
    reqs = []
    for i, t in enumerate(poisson_arrivals(200.0, 5, rng)):
        kind = _lib.TGV if i % 2 else _lib.TV
        d2f = _phantom_input(shape, (1.0, 1.0, 1.0), 70 + i, np.float32)
        m = brainish_mask(shape, seed=80 + i, frac=0.9)
        reqs.append(Request(i, float(t), kind, d2f, m, (1.0, 1.0, 1.0), 2e-3, 2.0, 40, kz=8))
    with Batch(np.float32, max_volumes=8) as b:
        stats = Server(b, chunk=10, max_live_voxels=2 * int(np.prod(shape))).run(reqs)
    assert stats.peak_live <= 2
    done = [r.done for r in reqs]
    assert done == sorted(done)
    for r in reqs:
        if r.kind == _lib.TGV:
            ref = orc.pd_tgv(r.d2f, r.mask, r.vsz, r.lam, r.alpha, 0.0, r.maxiter)[0]
        else:
            ref = orc.pd_tv(r.d2f, r.mask, r.vsz, r.lam, 0.0, r.maxiter)[0]
        np.testing.assert_array_equal(r.chi, ref, err_msg=f"request {r.rid}")


def test_serving_from_nifti_files(cuda, tmp_path):
    """Payloads as NIfTI files (loadNii.m): phase and mask written with
    saveNii, requests built from the files (d2f = -lapw(phase) on the device,
    the voxel size from the header), served, and chi written back
    (saveNii.m); the chi read from the file is bitwise the oracle's solve of
    the oracle's -lapw of the same phase."""
    from qsm_amd.nifti import loadNii, saveNii
    from qsm_amd.serve import request_from_nii, save_chi

    rng = np.random.default_rng(9)
    reqs, want = [], []
    for i, (shape, vsz) in enumerate((((26, 22, 18), (1.0, 1.0, 1.0)), ((24, 28, 20), (0.9, 1.1, 1.2)))):
        g = np.meshgrid(*(np.linspace(-1, 1, n) for n in shape), indexing="ij")
        phase = np.asfortranarray((0.8 * np.sin(3 * g[0]) * g[1] + 0.3 * g[2] ** 2
                                   + 0.01 * rng.standard_normal(shape)).astype(np.float32))
        m = brainish_mask(shape, seed=70 + i, frac=0.9)
        saveNii(str(tmp_path / f"phase{i}.nii.gz"), phase, vsz)
        saveNii(str(tmp_path / f"mask{i}.nii"), m.astype(np.uint8), vsz)
        r = request_from_nii(i, 0.01 * i, str(tmp_path / f"phase{i}.nii.gz"),
                             str(tmp_path / f"mask{i}.nii"), _lib.TGV, 2e-3, 2.0, 40)
        vs = tuple(float(np.float32(v)) for v in vsz)
        assert r.vsz == vs and r.d2f.dtype == np.float32
        d2f = -orc.lapw(phase, vs)
        np.testing.assert_array_equal(r.d2f, d2f)
        want.append(orc.pd_tgv(d2f, m.astype(np.uint8), vs, 2e-3, 2.0, 0.0, 40)[0])
        reqs.append(r)
    with Batch(np.float32, max_volumes=4) as b:
        Server(b, chunk=10).run(reqs)
    for r, w in zip(reqs, want):
        fn = str(tmp_path / f"chi{r.rid}.nii")
        save_chi(r, fn)
        got = loadNii(fn)
        np.testing.assert_array_equal(got["img"], w)
        np.testing.assert_allclose(got["hdr"]["dime"]["pixdim"][1:4], r.vsz, rtol=1e-6)
