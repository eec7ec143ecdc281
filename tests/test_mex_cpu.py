"""The MEX drop-in wrappers' argument handling (no GPU needed: every case
here is rejected before any library call).  Usage strings are the
reference's (pd_tgv_mex.c:107-112, pd_tv_mex.c:103-108, gradfp_mex.c:16-19,
gradfp_adj_mex.c:18-21, gradfm_mex.c:18-21, gradfm_adj_mex.c:18-21,
lapw_mex.c:24-27, mgpcg_mex.c:78-82, gradb_mex.c:16-19, gradbm_mex.c:18-21,
gradb_adj_mex.c:18-21, gradbm_adj_mex.c:18-21, gradc_mex.c:16-19,
gradcm_mex.c:18-21)."""
import numpy as np
import pytest

from mex_driver import LIB, MexError, call

pytestmark = pytest.mark.skipif(not LIB.exists(), reason="MEX CI library not built")

USAGE = {
    "pd_tgv_mex": ("Usage: chi = pd_tgv_mex(d2f, mask, vsz, lam, alpha, tol, maxiter, verbose);", 8),
    "pd_tv_mex": ("Usage: chi = pd_tv_mex(d2f, mask, vsz, lam, tol, maxiter, verbose);", 7),
    "gradfp_mex": ("Usage: gradfp_mex(dx, dy, dz, u, h);", 5),
    "gradfp_adj_mex": ("Usage: gradfp_adj_mex(du, x, y, z, h);", 5),
    "gradf_mex": ("Usage: gradf_mex(dx, dy, dz, u, h);", 5),
    "gradf_adj_mex": ("Usage: gradf_adj_mex(du, x, y, z, h);", 5),
    "gradfm_mex": ("Usage: gradfm_mex(dx, dy, dz, u, G, h);", 6),
    "gradfm_adj_mex": ("Usage: gradfm_adj_mex(du, x, y, z, G, h);", 6),
    "lapw_mex": ("Usage: lapw_mex(d2u, u, h);", 3),
    "lap1_mex": ("Usage: lap_mex(d2u, u, G, h);", 4),
    "gradb_mex": ("Usage: gradb_mex(dx, dy, dz, u, h);", 5),
    "gradbm_mex": ("Usage: gradbm_mex(dx, dy, dz, u, G, h);", 6),
    "gradb_adj_mex": ("Usage: gradb_adj_mex(du, x, y, z, h);", 5),
    "gradbm_adj_mex": ("Usage: gradbm_adj_mex(du, x, y, z, G, h);", 6),
    "gradc_mex": ("Usage: gradc_mex(dx, dy, dz, u, h);", 5),
    "gradcm_mex": ("Usage: gradcm_mex(dx, dy, dz, u, G, h);", 6),
    "mgpcg_mex": ("Usage: v = mgpcg_mex(v, f, mask, vsz, tolcg, maxitcg, tolmg, maxitmg, mu, "
                  "npre, npost, nboundary, nlevels, verbose)", 14),
}


@pytest.mark.parametrize("name", sorted(USAGE))
def test_wrong_argument_count_gives_reference_usage(name):
    msg, nrhs = USAGE[name]
    args = [np.zeros((3, 3, 3))] * (nrhs - 1)
    with pytest.raises(MexError) as e:
        call(name, 1, *args)
    assert str(e.value) == msg
    with pytest.raises(MexError) as e:
        call(name, 2, *([np.zeros((3, 3, 3))] * nrhs))
    assert str(e.value) == msg


def test_pd_rejects_non_double_voxel_size():
    with pytest.raises(MexError, match="voxel size"):
        call("pd_tgv_mex", 1, np.zeros((4, 4, 4)), np.ones((4, 4, 4), bool),
             np.array([1, 1, 1], np.float32), 1e-3, 2.0, 0.0, 10, 0)


def test_pd_rejects_bad_mask_class_and_size():
    with pytest.raises(MexError, match="mask"):
        call("pd_tv_mex", 1, np.zeros((4, 4, 4)), np.ones((4, 4, 4)), np.ones(3),
             1e-3, 0.0, 10, 0)
    with pytest.raises(MexError, match="mask size"):
        call("pd_tv_mex", 1, np.zeros((4, 4, 4)), np.ones((4, 4, 3), bool), np.ones(3),
             1e-3, 0.0, 10, 0)


def test_fd_rejects_mixed_classes():
    z = np.zeros((4, 4, 4))
    with pytest.raises(MexError, match="same class"):
        call("gradfp_mex", 0, z, z, z.astype(np.float32), z, np.ones(3))
    with pytest.raises(MexError, match="single or double"):
        call("lapw_mex", 0, z.astype(np.int32), z.astype(np.int32), np.ones(3))


def test_mgpcg_mex_rejects_too_many_levels_and_bad_classes():
    z = np.zeros((6, 6, 6))
    args = [z, z, np.ones(z.shape, bool), np.ones(3), 1e-6, 10, -1.0, 1, 1, 1, 1, 2]
    with pytest.raises(MexError, match="MAX_DEPTH"):
        call("mgpcg_mex", 1, *args, 17, 0)
    with pytest.raises(MexError, match="same class"):
        call("mgpcg_mex", 1, z, z.astype(np.float32), *args[2:], 0, 0)
    with pytest.raises(MexError, match="mask"):
        call("mgpcg_mex", 1, z, z, np.ones(z.shape), *args[3:], 0, 0)
