"""NIfTI-1 IO (loadNii.m / saveNii.m and the toolbox routines they call,
restated in qsm_amd/nifti.py).  MATLAB is not available, so the toolbox
cannot run here: parity is pinned by what fixes its behaviour

* the NIfTI-1 standard's header layout, decoded here independently with
  struct at the standard's byte offsets (sizeof_hdr 348, vox_offset 352,
  the sform save_nii_hdr.m writes when none is set);
* geometry: for an sform / qform that is a signed permutation, the loaded
  image is the stored one resampled onto +x, +y, +z world axes; checked by
  mapping every stored voxel to world coordinates (closed form);
* MATLAB's class conversion (round half away from zero, saturation) and its
  trailing-singleton rule; byte order; intensity scaling; spatial units.
"""
import gzip
import struct

import numpy as np
import pytest

from qsm_amd import nifti as nf


def _hdr_fields(raw):
    """Decode a NIfTI-1 header with the standard's offsets."""
    e = "<" if struct.unpack_from("<i", raw, 0)[0] == 348 else ">"
    u = lambda fmt, off: struct.unpack_from(e + fmt, raw, off)  # noqa: E731
    return {
        "sizeof_hdr": u("i", 0)[0], "regular": raw[38:39], "dim": u("8h", 40),
        "datatype": u("h", 70)[0], "bitpix": u("h", 72)[0], "pixdim": u("8f", 76),
        "vox_offset": u("f", 108)[0], "scl_slope": u("f", 112)[0], "xyzt_units": raw[123],
        "glmax": u("i", 140)[0], "glmin": u("i", 144)[0], "aux_file": raw[228:252],
        "qform_code": u("h", 252)[0], "sform_code": u("h", 254)[0],
        "srow": np.array([u("4f", 280), u("4f", 296), u("4f", 312)]), "magic": raw[344:348],
    }


def test_savenii_writes_the_standard_layout(tmp_path):
    img = np.asfortranarray(np.arange(4 * 3 * 2, dtype=np.float64).reshape((4, 3, 2), order="F") - 5.5)
    fn = tmp_path / "sub" / "chi.nii"
    nf.saveNii(str(fn), img, (0.9, 1.0, 1.25))
    raw = fn.read_bytes()
    h = _hdr_fields(raw)
    assert h["sizeof_hdr"] == 348 and h["magic"] == b"n+1\x00" and h["regular"] == b"r"
    assert h["dim"] == (3, 4, 3, 2, 1, 1, 1, 1)
    assert (h["datatype"], h["bitpix"]) == (64, 64)
    np.testing.assert_allclose(h["pixdim"][:4], (0.0, 0.9, 1.0, 1.25), rtol=1e-7)
    assert h["vox_offset"] == 352.0 and h["scl_slope"] == 0.0
    assert (h["glmax"], h["glmin"]) == (18, -6)           # round(17.5), round(-5.5)
    assert h["aux_file"].rstrip(b"\x00") == b"none"
    assert (h["qform_code"], h["sform_code"]) == (0, 1)
    # originator 0: offsets (1 - 0) * pixdim (save_nii_hdr.m)
    np.testing.assert_allclose(h["srow"], [[0.9, 0, 0, 0.9], [0, 1.0, 0, 1.0], [0, 0, 1.25, 1.25]],
                               rtol=1e-7)
    assert raw[348:352] == b"\x00" * 4 and len(raw) == 352 + img.size * 8
    np.testing.assert_array_equal(np.frombuffer(raw, "<f8", offset=352), img.reshape(-1, order="F"))


CLASSES = [np.uint8, np.int16, np.int32, np.float32, np.float64, np.int8, np.uint16, np.uint32,
           np.complex64, np.complex128]


@pytest.mark.parametrize("gz", [False, True])
@pytest.mark.parametrize("cls", CLASSES)
def test_round_trip_every_class(tmp_path, cls, gz):
    rng = np.random.default_rng(1)
    shape = (5, 4, 3, 2)
    x = rng.uniform(0, 100, shape)
    if np.dtype(cls).kind == "c":
        x = x + 1j * rng.uniform(-50, 50, shape)
    img = np.asfortranarray(np.round(x).astype(cls) if np.dtype(cls).kind in "iu" else x.astype(cls))
    fn = str(tmp_path / ("a.nii.gz" if gz else "a.nii"))
    nf.saveNii(fn, img, (1.1, 0.7, 2.0))
    nii = nf.loadNii(fn)
    assert nii["img"].dtype == img.dtype and nii["img"].shape == shape
    np.testing.assert_array_equal(nii["img"], img)
    np.testing.assert_allclose(nii["hdr"]["dime"]["pixdim"][1:4], (1.1, 0.7, 2.0), rtol=1e-6)
    assert nii["hdr"]["dime"]["datatype"] == nf._CLASS_DT[cls]
    assert nii["hdr"]["hist"]["rot_orient"] == [] and nii["filetype"] == 2
    if gz:
        with gzip.open(fn, "rb") as fh:
            assert fh.read(4) == struct.pack("<i", 348)


def test_make_nii_matlab_conversions():
    """Explicit datatype: MATLAB rounding (half away from zero) and
    saturation; trailing singleton dimensions dropped from dim."""
    nii = nf.make_nii(np.array([[1.5, -1.5], [40000.0, -2.5]]), (1, 1, 1), None, 4)
    np.testing.assert_array_equal(nii["img"], np.array([[2, -2], [32767, -3]], dtype=np.int16))
    nii = nf.make_nii(np.zeros((6, 5, 1)))
    assert list(nii["hdr"]["dime"]["dim"][:4]) == [2, 6, 5, 1]
    with pytest.raises(ValueError, match="not supported"):
        nf.make_nii(np.zeros((2, 2), dtype=np.int64))


def _write(fn, img, hdr_edit, big_endian=False):
    """A dataset with header fields edited before writing (the image bytes
    as stored, no reorientation)."""
    nii = nf.make_nii(img, (1, 1, 1))
    hdr_edit(nii["hdr"])
    nf.save_nii(nii, fn)
    if big_endian:
        raw = bytearray(open(fn, "rb").read())
        out = bytearray(raw)
        for fields in (nf._HK, nf._DIME, nf._HIST):
            for _, off, typ, n in fields:
                if typ in ("S", "u1"):
                    continue
                sz = np.dtype(typ).itemsize
                for k in range(n):
                    o = off + k * sz
                    out[o:o + sz] = raw[o:o + sz][::-1]
        data = np.frombuffer(bytes(raw[352:]), "<" + img.dtype.str[1:])
        out[352:] = data.astype(">" + img.dtype.str[1:]).tobytes()
        open(fn, "wb").write(bytes(out))


def _world_resample(img, R, T):
    """Closed form of the reorientation: every stored voxel v lands at world
    w = R v + T; the loaded image is indexed along +x, +y, +z with the
    voxel size |R| column sums.  Returns that image."""
    shape = img.shape[:3]
    idx = np.stack([g.reshape(-1, order="F")
                    for g in np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")], -1)
    w = idx @ R.T + T
    vs = np.abs(R).sum(axis=0)
    perm_vs = np.abs(R).sum(axis=1)         # voxel size along each world axis
    n = np.rint((w - w.min(axis=0)) / perm_vs).astype(int)
    out_shape = tuple(int(x) + 1 for x in n.max(axis=0))
    out = np.zeros(out_shape + img.shape[3:], img.dtype)
    flat = img.reshape((-1,) + img.shape[3:], order="F")
    out[n[:, 0], n[:, 1], n[:, 2]] = flat
    assert np.allclose(sorted(vs), sorted(perm_vs))
    return out


SFORMS = [
    np.diag([-2.0, 2.0, 3.0]),                                  # x reversed (radiological)
    np.array([[0, 1.5, 0], [2.0, 0, 0], [0, 0, -1.0]]),        # x <- j, y <- i, z reversed
    np.array([[0, 0, -1.0], [-1.0, 0, 0], [0, 1.2, 0]]),        # a cyclic permutation
]


@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("k", range(len(SFORMS)))
def test_sform_reorientation_is_the_world_resampling(tmp_path, k, big):
    R = SFORMS[k]
    T = np.array([10.0, -4.0, 7.0])
    img = np.asfortranarray(np.random.default_rng(k).standard_normal((6, 5, 4, 2)).astype(np.float32))

    def edit(h):
        h["hist"]["sform_code"] = 2.0
        h["hist"]["srow_x"] = np.r_[R[0], T[0]]
        h["hist"]["srow_y"] = np.r_[R[1], T[1]]
        h["hist"]["srow_z"] = np.r_[R[2], T[2]]

    fn = str(tmp_path / "o.nii")
    _write(fn, img, edit, big_endian=big)
    nii = nf.loadNii(fn)
    assert nii["machine"] == ("ieee-be" if big else "ieee-le")
    want = _world_resample(img, R, T)
    np.testing.assert_array_equal(nii["img"], want)
    np.testing.assert_allclose(nii["hdr"]["dime"]["pixdim"][1:4], np.abs(R).sum(axis=1), rtol=1e-6)
    assert list(nii["hdr"]["dime"]["dim"][1:4]) == list(want.shape[:3])
    assert nii["hdr"]["hist"]["sform_code"] == 0


@pytest.mark.parametrize("qfac", [1.0, -1.0])
def test_qform_quaternion(tmp_path, qfac):
    """(b, c, d) = (0, 0, 1): 180 degrees about z, x and y reversed; qfac = -1
    reverses z as well (the NIfTI-1 qform, xform_nii.m change_hdr)."""
    img = np.asfortranarray(np.arange(4 * 3 * 5, dtype=np.int16).reshape((4, 3, 5), order="F"))

    def edit(h):
        h["hist"]["qform_code"] = 1.0
        h["hist"]["quatern_d"] = 1.0
        h["dime"]["pixdim"] = np.array([qfac, 1.5, 1.0, 2.0, 1, 1, 1, 1])

    fn = str(tmp_path / "q.nii")
    _write(fn, img, edit)
    nii = nf.loadNii(fn)
    R = np.diag([-1.5, -1.0, 2.0 * qfac])
    np.testing.assert_array_equal(nii["img"], _world_resample(img, R, np.zeros(3)))
    np.testing.assert_allclose(nii["hdr"]["dime"]["pixdim"][:4], (abs(qfac), 1.5, 1.0, 2.0))


def test_tolerated_and_refused_shear(tmp_path):
    img = np.zeros((3, 3, 3), np.float32)
    img[0, 1, 2] = 1

    def edit(shear):
        def f(h):
            h["hist"]["sform_code"] = 1.0
            h["hist"]["srow_x"] = np.array([-1.0, shear, 0, 0])
            h["hist"]["srow_y"] = np.array([0, 1.0, 0, 0])
            h["hist"]["srow_z"] = np.array([0, 0, 1.0, 0])
        return f

    fn = str(tmp_path / "s.nii")
    _write(fn, img, edit(0.05))                    # below 10 % of the smallest axis
    nii = nf.loadNii(fn)
    assert nii["img"][2, 1, 2] == 1                 # x reversed
    assert "old_affine" in nii["hdr"]["hist"]
    _write(fn, img, edit(0.5))
    with pytest.raises(ValueError, match="Non-orthogonal"):
        nf.loadNii(fn)


@pytest.mark.parametrize("dt,slope,inter,want", [(np.int16, 0.5, 10.0, np.float32),
                                                 (np.float64, 2.0, -1.0, np.float64),
                                                 (np.uint8, 1.0, 0.0, np.uint8)])
def test_intensity_scaling(tmp_path, dt, slope, inter, want):
    img = np.asfortranarray(np.arange(24).reshape((4, 3, 2), order="F").astype(dt))

    def edit(h):
        h["dime"]["scl_slope"] = slope
        h["dime"]["scl_inter"] = inter

    fn = str(tmp_path / "c.nii")
    _write(fn, img, edit)
    nii = nf.loadNii(fn)
    assert nii["img"].dtype == want
    np.testing.assert_allclose(nii["img"], slope * img.astype(np.float64) + inter, rtol=1e-6)


@pytest.mark.parametrize("units,scale", [(1, 1e3), (3, 1e-3), (2, 1.0), (10, 1.0)])
def test_spatial_units(tmp_path, units, scale):
    def edit(h):
        h["dime"]["xyzt_units"] = float(units)

    fn = str(tmp_path / "u.nii")
    nii0 = nf.make_nii(np.zeros((2, 2, 2)), (0.5, 0.5, 1.0))
    nii0["hdr"]["dime"]["xyzt_units"] = float(units)
    nf.save_nii(nii0, fn)
    nii = nf.loadNii(fn)
    # saveNii's sform is diag(pixdim): the voxel size survives, then the unit
    np.testing.assert_allclose(nii["hdr"]["dime"]["pixdim"][1:4], np.array([0.5, 0.5, 1.0]) * scale)
    if scale != 1.0:
        assert int(nii["hdr"]["dime"]["xyzt_units"]) & 7 == 2


def test_loadnii_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        nf.loadNii(str(tmp_path / "none.nii"))
    bad = tmp_path / "x.img"
    bad.write_bytes(b"\x00" * 400)
    with pytest.raises(ValueError, match="extension"):
        nf.loadNii(str(bad))
    junk = tmp_path / "j.nii"
    junk.write_bytes(b"\x01" * 400)
    with pytest.raises(ValueError, match="corrupted"):
        nf.loadNii(str(junk))
    short = tmp_path / "t.nii"
    nf.saveNii(str(short), np.ones((4, 4, 4)))
    short.write_bytes(short.read_bytes()[:-8])
    with pytest.raises(ValueError, match="truncated"):
        nf.loadNii(str(short))
