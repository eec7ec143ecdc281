"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares; host-only entry points agree with the oracle."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from oracle import oracle as orc
from qsm_amd import _lib

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(qsm_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "qsm_pd_tgv" in names and "qsm_gradfp" in names and "qsm_lapw" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    if not _lib.LIB_PATH.exists():
        pytest.fail(f"{_lib.LIB_PATH} not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_library_loads_without_gpu():
    L = _lib.lib()
    assert L.qsm_version() >= 10000
    assert isinstance(L.qsm_last_error(), bytes)


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (0.9, 0.9, 1.2), (0.5, 2.0, 1.0)])
def test_step_sizes_match_oracle(h):
    t, s = C.c_double(), C.c_double()
    _lib.lib().qsm_step_sizes(_lib.h3(h), C.byref(t), C.byref(s))
    ot, os_ = orc.step_sizes(h)
    assert t.value == ot and s.value == os_


def test_footprint_counts_state_words():
    sz = _lib.sz3((256, 256, 176))
    N = 256 * 256 * 176
    f32 = _lib.lib().qsm_pd_footprint(_lib.TGV, _lib.F32, sz)
    # two banks of 15 words, f, one mask byte, small control/partials
    assert 31 * 4 * N + N <= f32 <= 31 * 4 * N + N + 64 * 2 ** 20
    f64tv = _lib.lib().qsm_pd_footprint(_lib.TV, _lib.F64, sz)
    assert 13 * 8 * N + N <= f64tv <= 13 * 8 * N + N + 64 * 2 ** 20


def test_tv_opts_defaults_are_tv_m():
    """qsm_tv_opts_default fills tv.m:38-86's defaults; the Python mirror
    keeps them for missing or empty fields and takes given ones."""
    from qsm_amd import tv as gtv

    o = gtv.make_opts(None)
    assert (o.mu, o.tv, o.rtol, o.atol, o.maxit, o.verbose) == (7.5e3, 1, 1e-1, 1e-4, 20, 1)
    assert (o.rho, o.alpha, o.rtau, o.stau, o.gamma) == (9.0, 1.5, 0.7, 1e2, 3.0)
    o = gtv.make_opts({"mu": 1e3, "tv": 2, "maxit": [], "rho": None})
    assert (o.mu, o.tv, o.maxit, o.rho) == (1e3, 2, 20, 9.0)


@pytest.mark.parametrize("bad,msg", [({"tv": 3}, "opts.tv"), ({"maxit": 0}, "maxit"),
                                     ({"rho": -1.0}, "rho")])
def test_tv_argument_checks_without_gpu(bad, msg):
    """Invalid opts are refused before any device work (QSM_EARG)."""
    import numpy as np

    from qsm_amd import QsmError
    from qsm_amd import tv as gtv

    f = np.zeros((8, 8, 8))
    with pytest.raises(QsmError, match=msg) as e:
        gtv.tv(f, np.ones(f.shape), (1, 1, 1), None, dict(bad, verbose=0))
    assert e.value.code == _lib.QSM_EARG


def test_rts_opts_defaults_are_rts_m():
    from qsm_amd import rts as grts

    o = grts.make_opts(None)
    assert (o.tolls, o.mu, o.delta, o.tv, o.rtol, o.atol, o.maxit) == (4, 1e5, 0.15, 1, 1e-1, 1e-2, 20)
    assert (o.verbose, o.rho, o.alpha, o.rtau, o.stau, o.gamma) == (1, 10.0, 1.5, 0.7, 1e2, 3.0)


def test_mg_opts_defaults_are_mgpcg_m():
    """qsm_mg_opts_default fills mgpcg.m:71-101's defaults (nlevels 0: the
    size rule max(1, floor(log2(min(size))) - 3), applied by the library)."""
    from qsm_amd import unwrap as uw

    o = uw.make_mgopts(None)
    assert (o.tol, o.maxit, o.mu, o.npre, o.npost, o.nboundary, o.nlevels) == (-1.0, 1, 1, 1, 1, 2, 0)
    o = uw.make_mgopts({"npre": 2, "nlevels": 3, "tol": []})
    assert (o.npre, o.nlevels, o.tol) == (2, 3, -1.0)


@pytest.mark.parametrize("kw,msg", [({"mgopts": {"nlevels": 17}}, "nlevels"),
                                    ({"mgopts": {"mu": 0}}, "mgopts"),
                                    ({"maxit": 0}, "maxit")])
def test_mgpcg_argument_checks_without_gpu(kw, msg):
    """More levels than the library holds, invalid mgopts and maxit are
    refused before any device work."""
    import numpy as np

    from qsm_amd import QsmError
    from qsm_amd import unwrap as uw

    f = np.zeros((16, 16, 16))
    with pytest.raises(QsmError, match=msg) as e:
        uw.mgpcg(f, np.ones(f.shape), (1, 1, 1), **kw)
    assert e.value.code == _lib.QSM_EARG


def test_mgpcg_work_bytes_cover_the_hierarchy():
    """Work space: per level v, f, r and two mask bytes, and on the levels
    that smooth (all but the coarsest) the band's index lists (n + 1 ints);
    PCG x, p; coarse CG scratch; band seed; the lists' per-row counts
    (2 ints per level-0 row); reduction partials."""
    n = [(34, 30, 26), (18, 16, 14), (10, 9, 8)]      # 32x28x24 padded, 3 levels
    vox = [a * b * c for a, b, c in n]
    w = _lib.lib().qsm_mgpcg_work_bytes(_lib.F64, (C.c_size_t * 3)(32, 28, 24), 3)
    lo = (sum(3 * 8 * v + 2 * v for v in vox) + sum(4 * (v + 1) for v in vox[:-1])
          + 2 * 8 * vox[0] + 2 * 8 * vox[-1] + vox[0] + 2 * 4 * n[0][1] * n[0][2])
    assert lo <= w <= lo + 64 * 1024


def test_sharp_work_bytes_only_hold_what_each_call_uses():
    """SHARP holds no reused spectra; V-SHARP holds the mask's (single
    precision) and, for one echo, f's as well."""
    from qsm_amd import bgremove as bg

    szc, vsz, r = (60, 70, 50), (1.0, 1.0, 1.0), 9.0
    for dt in ("float32", "float64"):
        s = bg.work_bytes(dt, szc, vsz, r)
        v1 = bg.vsharp_work_bytes(dt, szc, vsz, r, 1)
        v2 = bg.vsharp_work_bytes(dt, szc, vsz, r, 2)
        assert s < v2 < v1
        # f's spectrum (class of f) is at least the single-precision mask's
        assert v1 - v2 >= v2 - s


@pytest.mark.parametrize("rmax,vsz", [(float("nan"), (1, 1, 1)), (-3.0, (1, 1, 1)), (0.0, (1, 1, 1)),
                                      (float("inf"), (1, 1, 1)), (9.0, (1, 0, 1)),
                                      (9.0, (1, float("nan"), 1)), (9.0, (1, -1, 1))])
def test_sharp_work_bytes_refuse_bad_radius_or_voxel_size(rmax, vsz):
    """Bad radii or voxel sizes give 0 bytes from the library (never an
    undefined padding) and an error from the Python wrappers."""
    from qsm_amd import bgremove as bg

    sz = (C.c_size_t * 3)(20, 20, 20)
    h = (C.c_double * 3)(*[float(x) for x in vsz])
    assert _lib.lib().qsm_sharp_work_bytes(_lib.F64, sz, h, rmax) == 0
    assert _lib.lib().qsm_vsharp_work_bytes(_lib.F32, sz, h, rmax, 1) == 0
    with pytest.raises(ValueError):
        bg.work_bytes("float64", (20, 20, 20), vsz, rmax)
    with pytest.raises(ValueError):
        bg.vsharp_work_bytes("float32", (20, 20, 20), vsz, rmax, 2)


def test_sharp_work_bytes_refuse_empty_crop_and_no_echoes():
    from qsm_amd import bgremove  # noqa: F401  (registers the signatures)

    sz0 = (C.c_size_t * 3)(20, 0, 20)
    sz = (C.c_size_t * 3)(20, 20, 20)
    h = (C.c_double * 3)(1.0, 1.0, 1.0)
    assert _lib.lib().qsm_sharp_work_bytes(_lib.F64, sz0, h, 9.0) == 0
    assert _lib.lib().qsm_vsharp_work_bytes(_lib.F64, sz, h, 9.0, 0) == 0
    assert _lib.lib().qsm_sharp_work_bytes(7, sz, h, 9.0) == 0


def test_sstgv_refuses_bad_arguments_before_device_work():
    """qsm_sstgv checks its arguments before it looks for a device: a bad
    class, a null pointer, no echo, maxit 0 and a non-positive voxel size
    are QSM_EARG (with the reason as the last error)."""
    import numpy as np

    lib = _lib.lib()
    f = np.zeros((4, 4, 4), np.float64, order="F")
    m = np.ones((4, 4, 4), np.uint8, order="F")
    x = np.zeros_like(f)
    sz = (C.c_size_t * 3)(4, 4, 4)
    good = (C.c_double * 3)(1.0, 1.0, 1.0)
    bad = (C.c_double * 3)(1.0, 0.0, 1.0)
    noprint = _lib.PRINT_FN()

    def call(dtype=_lib.F64, fp=f.ctypes.data, ne=1, vsz=good, maxit=10):
        return lib.qsm_sstgv(dtype, fp, m.ctypes.data, sz, ne, vsz, 1e-3, 2.0, 0.0, maxit, 0,
                             x.ctypes.data, noprint, None)

    for kw, why in ((dict(dtype=7), "dtype"), (dict(fp=None), "null"), (dict(ne=0), "echo"),
                    (dict(maxit=0), "maxit"), (dict(vsz=bad), "voxel size")):
        assert call(**kw) == _lib.QSM_EARG, kw
        assert why in lib.qsm_last_error().decode(), (kw, lib.qsm_last_error())
