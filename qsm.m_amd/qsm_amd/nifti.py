"""NIfTI-1 IO for request payloads: src/io/loadNii.m and saveNii.m with the
parts of the vendored NIfTI toolbox they call (third_party/NIfTI:
load_nii.m, load_nii_hdr.m, load_nii_img.m, xform_nii.m, make_nii.m,
save_nii.m, save_nii_hdr.m), restated for numpy.

Host-side code: reading and writing files is the CPU's work; the arrays it
returns are what the device entry points take (column-major, numpy arrays in
Fortran order, img[i, j, k] is MATLAB's img(i+1, j+1, k+1)).

A loaded dataset is a dict shaped like the toolbox's struct: nii["hdr"]
with "hk", "dime", "hist" (the NIfTI-1 header fields by their toolbox
names), nii["img"], nii["filetype"] (2: single .nii file, 1: .hdr/.img
pair magic, 0: Analyze), nii["fileprefix"], nii["machine"] ("ieee-le" or
"ieee-be") and nii["original"]["hdr"] (the header before xform_nii).

Scope: single-file datasets (.nii, .nii.gz), which is what loadNii accepts;
the whole image is read (loadNii passes no volume selection); extensions are
neither read nor written (load_nii does not read them, make_nii makes
none).  A .hdr/.img pair or an Analyze file is refused.

The header transform (sform/qform choice, tolerance zeroing, quaternion
matrix, old/new affine) and one error message follow xform_nii.m closely,
because those rules define what loadNii returns.  That toolbox carries this
notice, retained here:

    Copyright (c) 2014, Jimmy Shen
    All rights reserved.

    Redistribution and use in source and binary forms, with or without
    modification, are permitted provided that the following conditions are
    met:

        * Redistributions of source code must retain the above copyright
          notice, this list of conditions and the following disclaimer.
        * Redistributions in binary form must reproduce the above copyright
          notice, this list of conditions and the following disclaimer in
          the documentation and/or other materials provided with the
          distribution

    THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS
    IS" AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED TO,
    THE IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR
    PURPOSE ARE DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT OWNER OR
    CONTRIBUTORS BE LIABLE FOR ANY DIRECT, INDIRECT, INCIDENTAL, SPECIAL,
    EXEMPLARY, OR CONSEQUENTIAL DAMAGES (INCLUDING, BUT NOT LIMITED TO,
    PROCUREMENT OF SUBSTITUTE GOODS OR SERVICES; LOSS OF USE, DATA, OR
    PROFITS; OR BUSINESS INTERRUPTION) HOWEVER CAUSED AND ON ANY THEORY OF
    LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY, OR TORT (INCLUDING
    NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY OUT OF THE USE OF THIS
    SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
"""
from __future__ import annotations

import copy
import gzip
import math
import os

import numpy as np

# datatype -> (bitpix, numpy scalar type of one stored component, components)
# (load_nii_img.m:156-197, save_nii.m:141-180)
_DTYPES = {
    2: (8, np.uint8, 1), 4: (16, np.int16, 1), 8: (32, np.int32, 1),
    16: (32, np.float32, 1), 32: (64, np.float32, 2), 64: (64, np.float64, 1),
    128: (24, np.uint8, 3), 256: (8, np.int8, 1), 511: (96, np.float32, 3),
    512: (16, np.uint16, 1), 768: (32, np.uint32, 1), 1024: (64, np.int64, 1),
    1280: (64, np.uint64, 1), 1792: (128, np.float64, 2),
}

# make_nii.m:103-135: class of the image -> datatype
_CLASS_DT = {np.uint8: 2, np.int16: 4, np.int32: 8, np.float32: 16, np.float64: 64,
             np.int8: 256, np.uint16: 512, np.uint32: 768, np.complex64: 32,
             np.complex128: 1792}

# datatype -> class make_nii.m:170-200 casts the image to
_DT_CLASS = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 32: np.complex64,
             64: np.float64, 128: np.uint8, 256: np.int8, 512: np.uint16,
             768: np.uint32, 1792: np.complex128}

# The NIfTI-1 header: (name, offset, numpy type, count); strings are
# ("S", length).  Field order and sizes are the standard's (load_nii_hdr.m:
# 178-269 reads them in this order).
_HK = [("sizeof_hdr", 0, "i4", 1), ("data_type", 4, "S", 10), ("db_name", 14, "S", 18),
       ("extents", 32, "i4", 1), ("session_error", 36, "i2", 1), ("regular", 38, "S", 1),
       ("dim_info", 39, "u1", 1)]
_DIME = [("dim", 40, "i2", 8), ("intent_p1", 56, "f4", 1), ("intent_p2", 60, "f4", 1),
         ("intent_p3", 64, "f4", 1), ("intent_code", 68, "i2", 1), ("datatype", 70, "i2", 1),
         ("bitpix", 72, "i2", 1), ("slice_start", 74, "i2", 1), ("pixdim", 76, "f4", 8),
         ("vox_offset", 108, "f4", 1), ("scl_slope", 112, "f4", 1), ("scl_inter", 116, "f4", 1),
         ("slice_end", 120, "i2", 1), ("slice_code", 122, "u1", 1), ("xyzt_units", 123, "u1", 1),
         ("cal_max", 124, "f4", 1), ("cal_min", 128, "f4", 1), ("slice_duration", 132, "f4", 1),
         ("toffset", 136, "f4", 1), ("glmax", 140, "i4", 1), ("glmin", 144, "i4", 1)]
_HIST = [("descrip", 148, "S", 80), ("aux_file", 228, "S", 24), ("qform_code", 252, "i2", 1),
         ("sform_code", 254, "i2", 1), ("quatern_b", 256, "f4", 1), ("quatern_c", 260, "f4", 1),
         ("quatern_d", 264, "f4", 1), ("qoffset_x", 268, "f4", 1), ("qoffset_y", 272, "f4", 1),
         ("qoffset_z", 276, "f4", 1), ("srow_x", 280, "f4", 4), ("srow_y", 296, "f4", 4),
         ("srow_z", 312, "f4", 4), ("intent_name", 328, "S", 16), ("magic", 344, "S", 4)]
HDR_SIZE = 348


def _deblank(b: bytes) -> str:
    """MATLAB deblank of a char array read byte by byte: trailing white space
    and NUL characters removed."""
    return b.decode("latin-1").rstrip(" \t\n\v\f\r\x00")


def _unpack(buf: bytes, fields, end: str) -> dict:
    out = {}
    for name, off, typ, n in fields:
        if typ == "S":
            out[name] = _deblank(buf[off:off + n])
        else:
            a = np.frombuffer(buf, dtype=np.dtype(end + typ), count=n, offset=off)
            v = a.astype(np.float64)      # fread(..., 'int16') etc. return doubles
            out[name] = v if n > 1 else float(v[0])
    return out


def _parse_header(buf: bytes):
    """load_nii_hdr.m:108-176 and read_header: little-endian first, the other
    byte order when sizeof_hdr is not 348 there.  Returns (hdr, filetype,
    machine)."""
    if len(buf) < HDR_SIZE:
        raise ValueError("file is shorter than a NIfTI-1 header")
    machine, end = "ieee-le", "<"
    if int(np.frombuffer(buf, "<i4", 1, 0)[0]) != HDR_SIZE:
        machine, end = "ieee-be", ">"
        if int(np.frombuffer(buf, ">i4", 1, 0)[0]) != HDR_SIZE:
            raise ValueError("file is corrupted (sizeof_hdr is not 348 in either byte order)")
    hdr = {"hk": _unpack(buf, _HK, end), "dime": _unpack(buf, _DIME, end),
           "hist": _unpack(buf, _HIST, end)}
    # the Analyze 7.5 originator, read where that format keeps it (5 int16
    # at byte 253, overlapping NIfTI's qform/sform fields)
    hdr["hist"]["originator"] = np.frombuffer(buf, end + "i2", 5, 253).astype(np.float64)
    magic = hdr["hist"]["magic"]
    if magic not in ("n+1", "ni1"):
        hdr["hist"]["qform_code"] = 0.0
        hdr["hist"]["sform_code"] = 0.0
    filetype = 2 if magic == "n+1" else 1 if magic == "ni1" else 0
    return hdr, filetype, machine


def _matlab_shape(shape):
    """A MATLAB array's size: trailing singleton dimensions beyond the
    second are dropped."""
    s = list(shape)
    while len(s) > 2 and s[-1] == 1:
        s.pop()
    return tuple(s)


def _read_image(hdr, filetype, buf: bytes, end: str):
    """load_nii_img.m: read_image for the whole image (no volume selection)."""
    dime = hdr["dime"]
    dt = int(dime["datatype"])
    if dt not in _DTYPES:
        raise ValueError(f"This datatype is not supported: {dt}")
    bitpix, typ, ncomp = _DTYPES[dt]
    dime["bitpix"] = float(bitpix)
    dim = np.where(dime["dim"] < 1, 1.0, dime["dim"])
    dime["dim"] = dim
    d = [int(x) for x in dim[1:8]]
    n = int(np.prod(d)) * ncomp
    off = int(dime["vox_offset"]) if filetype == 2 else 0
    need = off + n * np.dtype(typ).itemsize
    if len(buf) < need:
        raise ValueError(f"image data truncated: {len(buf)} bytes, {need} needed")
    raw = np.frombuffer(buf, dtype=np.dtype(typ).newbyteorder(end), count=n, offset=off)
    img = raw.astype(np.dtype(typ).newbyteorder("="))
    if dt in (32, 1792):
        img = img[0::2] + 1j * img[1::2]
        img = img.astype(np.complex64 if dt == 32 else np.complex128)
    if img.size:
        if np.iscomplexobj(img):
            k = int(np.argmax(np.abs(img)))
            kmin = int(np.argmin(np.abs(img)))
            dime["glmax"], dime["glmin"] = complex(img[k]), complex(img[kmin])
        else:
            dime["glmax"], dime["glmin"] = float(img.max()), float(img.min())
    if dt == 128:
        img = img.reshape([3] + d, order="F")
        img = np.moveaxis(img, 0, 3)
    elif dt == 511:
        x = img.astype(np.float64)
        img = ((x - x.min()) / (x.max() - x.min())).astype(np.float32)
        img = np.moveaxis(img.reshape([3] + d, order="F"), 0, 3)
    else:
        img = img.reshape(d, order="F")
    return np.asfortranarray(img.reshape(_matlab_shape(img.shape), order="F")), hdr


def _get_orient(inv_r):
    """xform_nii.m: get_orient.  Per row: the column of its one nonzero,
    signed by the row sum; 1..3 = +x, +y, +z, 4..6 = the reverse."""
    orient = []
    for i in range(3):
        nz = np.flatnonzero(inv_r[i])
        code = int(nz[0] + 1) * int(np.sign(inv_r[i].sum()))
        orient.append({1: 1, 2: 2, 3: 3, -1: 4, -2: 5, -3: 6}[code])
    return orient


def _is_scaled_permutation(r):
    """The toolbox's test `det(R) ~= 0 & isequal(R(find(R)), sum(R)')`: one
    nonzero per column and nonsingular."""
    nz = r.flatten(order="F")
    nz = nz[nz != 0]
    return np.linalg.det(r) != 0 and nz.size == 3 and np.array_equal(nz, r.sum(axis=0))


def _non_orthogonal():
    return ValueError("Non-orthogonal rotation or shearing found inside the affine matrix "
                      "in this NIfTI file (xform_nii.m); reslice it first")


def _change_hdr(hdr, tolerance, preferred):
    """xform_nii.m: change_hdr."""
    hist, dime = hdr["hist"], hdr["dime"]
    orient = [1, 2, 3]
    use = None
    if preferred == "S":
        if hist["sform_code"] == 0:
            raise ValueError("User requires sform, sform not set in header")
        use = "s"
    if preferred == "Q":
        if hist["qform_code"] == 0:
            raise ValueError("User requires qform, qform not set in header")
        use = "q"
    if preferred == "s":
        use = "s" if hist["sform_code"] > 0 else "q" if hist["qform_code"] > 0 else None
    if preferred == "q":
        use = "q" if hist["qform_code"] > 0 else "s" if hist["sform_code"] > 0 else None

    def tolerate(r):
        rs = np.sort(np.abs(r.flatten(order="F")))
        r = r.copy()
        r[np.abs(r) < tolerance * rs[-3:].min()] = 0.0
        return r

    if use == "s":
        r = np.array([hist["srow_x"][:3], hist["srow_y"][:3], hist["srow_z"][:3]], dtype=np.float64)
        t = np.array([hist["srow_x"][3], hist["srow_y"][3], hist["srow_z"][3]], dtype=np.float64)
        if not _is_scaled_permutation(r):
            hist["old_affine"] = np.block([[r, t[:, None]], [np.zeros((1, 3)), np.ones((1, 1))]])
            r = tolerate(r)
            hist["new_affine"] = np.block([[r, t[:, None]], [np.zeros((1, 3)), np.ones((1, 1))]])
            if not _is_scaled_permutation(r):
                raise _non_orthogonal()
    elif use == "q":
        b, c, d = hist["quatern_b"], hist["quatern_c"], hist["quatern_d"]
        s = 1.0 - (b * b + c * c + d * d)
        if s < 0:
            if abs(s) < 1e-5:
                a = 0.0
            else:
                raise ValueError("Incorrect quaternion values in this NIFTI data.")
        else:
            a = math.sqrt(s)
        qfac = dime["pixdim"][0]
        if qfac == 0:
            qfac = 1.0
        scale = np.diag([dime["pixdim"][1], dime["pixdim"][2], qfac * dime["pixdim"][3]])
        r = np.array([[a * a + b * b - c * c - d * d, 2 * b * c - 2 * a * d, 2 * b * d + 2 * a * c],
                      [2 * b * c + 2 * a * d, a * a + c * c - b * b - d * d, 2 * c * d - 2 * a * b],
                      [2 * b * d - 2 * a * c, 2 * c * d + 2 * a * b, a * a + d * d - c * c - b * b]])
        t = np.array([hist["qoffset_x"], hist["qoffset_y"], hist["qoffset_z"]], dtype=np.float64)
        if not _is_scaled_permutation(r):
            hist["old_affine"] = np.block([[r @ scale, t[:, None]], [np.zeros((1, 3)), np.ones((1, 1))]])
            r = tolerate(r) @ scale
            hist["new_affine"] = np.block([[r, t[:, None]], [np.zeros((1, 3)), np.ones((1, 1))]])
            if not _is_scaled_permutation(r):
                raise _non_orthogonal()
        else:
            r = r @ scale
    if use is not None:
        inv_r = np.linalg.inv(r)
        pix = dime["pixdim"].copy()
        pix[1:4] = np.abs(r.sum(axis=0))
        dime["pixdim"] = pix
        org = hist["originator"].copy()
        org[:3] = inv_r @ (-t) + 1.0
        hist["originator"] = org
        orient = _get_orient(inv_r)
        hist["qform_code"] = 0.0
        hist["sform_code"] = 0.0
    units = int(dime["xyzt_units"]) & 7
    space = 1e3 if units == 1 else 1e-3 if units == 3 else 1.0
    if space != 1.0:
        pix = dime["pixdim"].copy()
        pix[1:4] = pix[1:4] * space
        dime["pixdim"] = pix
        dime["xyzt_units"] = float((int(dime["xyzt_units"]) & ~0b111) | 0b010)
    dime["pixdim"] = np.abs(dime["pixdim"])
    return hdr, orient


def xform_nii(nii, tolerance=0.1, preferred_form="s"):
    """xform_nii.m: intensity scaling, then the image and header turned to
    the orientation the affine (sform preferred) maps to +x, +y, +z."""
    nii["original"] = {"hdr": copy.deepcopy(nii["hdr"])}
    if tolerance <= 0:
        tolerance = np.finfo(np.float64).eps
    dime = nii["hdr"]["dime"]
    slope, inter, dt = dime["scl_slope"], dime["scl_inter"], int(dime["datatype"])
    if slope != 0 and dt in (2, 4, 8, 16, 64, 256, 512, 768) and (slope != 1 or inter != 0):
        img = slope * nii["img"].astype(np.float64) + inter
        if dt == 64:
            dime["datatype"], dime["bitpix"] = 64.0, 64.0
        else:
            img = img.astype(np.float32)
            dime["datatype"], dime["bitpix"] = 16.0, 32.0
        nii["img"] = np.asfortranarray(img)
        dime["glmax"], dime["glmin"] = float(img.max()), float(img.min())
        dime["scl_slope"] = 0.0
    if dime["scl_slope"] != 0 and int(dime["datatype"]) in (32, 1792):
        img = dime["scl_slope"] * nii["img"].astype(np.complex128) + dime["scl_inter"]
        if int(dime["datatype"]) == 32:
            img = img.astype(np.complex64)
        nii["img"] = np.asfortranarray(img)
        k, kmin = int(np.argmax(np.abs(img))), int(np.argmin(np.abs(img)))
        dime["glmax"], dime["glmin"] = complex(img.flat[k]), complex(img.flat[kmin])
        dime["scl_slope"] = 0.0
    if nii["filetype"] == 0:
        raise ValueError("Analyze 7.5 files are not supported (no NIfTI magic)")
    hdr, orient = _change_hdr(nii["hdr"], tolerance, preferred_form)
    hist, dime = hdr["hist"], hdr["dime"]
    if orient != [1, 2, 3]:
        if int(dime["datatype"]) in (32, 1792, 128, 511):
            # xform_nii.m indexes these as 4-D colour/component planes
            raise ValueError("reorienting complex or RGB images is not supported")
        rot = [((o + 2) % 3) + 1 for o in orient]
        flip = [o - r for o, r in zip(orient, rot)]
        img = nii["img"]
        for i in range(3):
            if flip[i]:
                img = np.flip(img, axis=i)
        perm = sorted(range(3), key=lambda i: rot[i])     # [tmp rot] = sort(rot)
        old_dim = dime["dim"][1:4].copy()
        new_dim = old_dim[perm]
        dd = dime["dim"].copy()
        dd[1:4] = new_dim
        dime["dim"] = dd
        pix = dime["pixdim"].copy()
        pix[1:4] = pix[1:4][perm]
        dime["pixdim"] = pix
        org = hist["originator"].copy()
        tmp = org[:3][perm]
        flipped = [flip[p] for p in perm]
        for i in range(3):
            if flipped[i] and tmp[i] != 0:
                tmp[i] = new_dim[i] - tmp[i] + 1
        org[:3] = tmp
        hist["originator"] = org
        hist["rot_orient"] = [p + 1 for p in perm]
        hist["flip_orient"] = flipped
        nd = img.ndim
        img = np.transpose(img, perm + list(range(3, nd)))
        nii["img"] = np.asfortranarray(img)
    else:
        hist["rot_orient"] = []
        hist["flip_orient"] = []
    nii["hdr"] = hdr
    return nii


def _split_name(filename: str):
    """(prefix, gz) of a single-file dataset name; others are refused."""
    fn = os.fspath(filename)
    gz = fn.endswith(".gz")
    base = fn[:-3] if gz else fn
    if not base.endswith(".nii"):
        raise ValueError(f"wrong file extension {fn} (.nii or .nii.gz)")
    return base[:-4], gz


def load_nii(filename, tolerance=0.1, preferred_form="s"):
    """load_nii.m for a single-file dataset (.nii, .nii.gz)."""
    prefix, gz = _split_name(filename)
    with (gzip.open(filename, "rb") if gz else open(filename, "rb")) as fh:
        buf = fh.read()
    hdr, filetype, machine = _parse_header(buf)
    if filetype != 2:
        raise ValueError(f"{filename}: not a single-file NIfTI-1 dataset (magic "
                         f"{hdr['hist']['magic']!r})")
    img, hdr = _read_image(hdr, filetype, buf, "<" if machine == "ieee-le" else ">")
    nii = {"hdr": hdr, "img": img, "filetype": filetype, "fileprefix": prefix, "machine": machine}
    return xform_nii(nii, tolerance, preferred_form)


def loadNii(filename):  # noqa: N802  (the reference's name)
    """loadNii.m: the dataset of a .nii or .nii.gz file (the .gz is read in
    memory rather than unpacked next to it)."""
    fn = os.fspath(filename)
    if not os.path.isfile(fn):
        raise FileNotFoundError(f"file {fn} does not exist")
    ext = os.path.splitext(fn)[1].lower()
    if ext not in (".gz", ".nii"):
        raise ValueError(f"wrong file extension {fn}")
    return load_nii(fn)


def _matlab_round(x):
    """MATLAB round: half away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def _matlab_cast(img, cls):
    """MATLAB's class conversion: to integers with rounding half away from
    zero and saturation (NaN -> 0); to floats as is."""
    cls = np.dtype(cls)
    if img.dtype == cls:
        return img
    if cls.kind in "iu":
        info = np.iinfo(cls)
        x = np.nan_to_num(_matlab_round(np.real(img).astype(np.float64)), nan=0.0)
        return np.clip(x, info.min, info.max).astype(cls)
    if cls.kind == "c":
        return img.astype(cls)
    return np.real(img).astype(cls) if np.iscomplexobj(img) else img.astype(cls)


def _extreme(img, which):
    """MATLAB max / min over all elements (complex: by magnitude), rounded
    as make_nii.m and save_nii.m round glmax / glmin; the real part is what
    an int32 header field keeps."""
    if img.size == 0:
        return 0.0
    if np.iscomplexobj(img):
        k = int(np.argmax(np.abs(img)) if which == "max" else np.argmin(np.abs(img)))
        v = np.real(img.flat[k])
    else:
        v = img.max() if which == "max" else img.min()
    return float(_matlab_round(np.float64(v)))


def make_nii(img, voxel_size=None, origin=None, datatype=None, descrip=""):
    """make_nii.m: a dataset struct for an array (up to 7-D)."""
    img = np.asarray(img)
    if img.ndim > 7:
        raise ValueError("NIfTI only allows a maximum of 7 Dimension matrix.")
    shape = _matlab_shape(img.shape if img.ndim >= 2 else (img.size, 1))
    dims = np.ones(8)
    dims[0] = len(shape)
    dims[1:1 + len(shape)] = shape
    cls = img.dtype.type
    if datatype is None:
        if cls not in _CLASS_DT:
            raise ValueError("Datatype is not supported by make_nii.")
        datatype = _CLASS_DT[cls]
    datatype = int(datatype)
    if datatype in (128, 511):
        dims = np.concatenate([dims[:4], dims[5:], [1.0]])
        dims[0] -= 1
    pixdim = np.array([0.0] + [1.0] * 7)
    if voxel_size is not None and np.size(voxel_size) > 0:
        pixdim[1:4] = np.broadcast_to(np.asarray(voxel_size, dtype=np.float64).reshape(-1), (3,))
    org = np.zeros(5)
    if origin is not None and np.size(origin) > 0:
        org[:3] = np.asarray(origin, dtype=np.float64).reshape(-1)[:3]
    if datatype not in _DTYPES or datatype in (1024, 1280):
        raise ValueError("Datatype is not supported by make_nii.")
    bitpix = _DTYPES[datatype][0]
    hdr = {
        "hk": {"sizeof_hdr": 348.0, "data_type": "", "db_name": "", "extents": 0.0,
               "session_error": 0.0, "regular": "r", "dim_info": 0.0},
        "dime": {"dim": dims, "intent_p1": 0.0, "intent_p2": 0.0, "intent_p3": 0.0,
                 "intent_code": 0.0, "datatype": float(datatype), "bitpix": float(bitpix),
                 "slice_start": 0.0, "pixdim": pixdim, "vox_offset": 0.0, "scl_slope": 0.0,
                 "scl_inter": 0.0, "slice_end": 0.0, "slice_code": 0.0, "xyzt_units": 0.0,
                 "cal_max": 0.0, "cal_min": 0.0, "slice_duration": 0.0, "toffset": 0.0,
                 "glmax": _extreme(img, "max"), "glmin": _extreme(img, "min")},
        "hist": {"descrip": descrip, "aux_file": "none", "qform_code": 0.0, "sform_code": 0.0,
                 "quatern_b": 0.0, "quatern_c": 0.0, "quatern_d": 0.0, "qoffset_x": 0.0,
                 "qoffset_y": 0.0, "qoffset_z": 0.0, "srow_x": np.zeros(4), "srow_y": np.zeros(4),
                 "srow_z": np.zeros(4), "intent_name": "", "magic": "", "originator": org},
    }
    if datatype == 511:
        x = img.astype(np.float64)
        out = ((x - x.min()) / (x.max() - x.min())).astype(np.float32)
        hdr["dime"]["glmax"], hdr["dime"]["glmin"] = float(out.max()), float(out.min())
    else:
        out = _matlab_cast(img, _DT_CLASS[datatype])
    return {"hdr": hdr, "img": np.asfortranarray(out)}


def _pack_string(s: str, n: int) -> bytes:
    b = s.encode("latin-1")[:n]
    return b + b"\x00" * (n - len(b))


def _header_bytes(hdr) -> bytes:
    """save_nii_hdr.m: write_header (little-endian, the native order of the
    machines it runs on)."""
    buf = bytearray(HDR_SIZE)
    for fields, part in ((_HK, hdr["hk"]), (_DIME, hdr["dime"]), (_HIST, hdr["hist"])):
        for name, off, typ, n in fields:
            v = part[name]
            if typ == "S":
                buf[off:off + n] = _pack_string(str(v), n)
            else:
                a = np.asarray(v, dtype=np.float64).reshape(-1)[:n]
                if typ != "f4":
                    a = _matlab_round(a)
                buf[off:off + n * np.dtype(typ).itemsize] = a.astype("<" + typ).tobytes()
    return bytes(buf)


def save_nii(nii, filename):
    """save_nii.m / write_nii for a single-file dataset (.nii, .nii.gz):
    the 348-byte header, four zero bytes (vox_offset 352), the image in
    column-major order in the datatype's precision."""
    prefix, gz = _split_name(filename)
    hdr = copy.deepcopy(nii["hdr"])
    img = np.asarray(nii["img"])
    dime, hist = hdr["dime"], hdr["hist"]
    dt = int(dime["datatype"])
    if dt not in _DTYPES:
        raise ValueError("This datatype is not supported")
    bitpix, typ, ncomp = _DTYPES[dt]
    dime["bitpix"] = float(bitpix)
    dime["glmax"] = _extreme(img, "max")
    dime["glmin"] = _extreme(img, "min")
    dime["vox_offset"] = 352.0
    hist["magic"] = "n+1"
    if hist["qform_code"] == 0 and hist["sform_code"] == 0:
        # save_nii_hdr.m:38-55: an sform from the voxel size and originator
        pix, org = dime["pixdim"], hist["originator"]
        hist["sform_code"] = 1.0
        hist["srow_x"] = np.array([pix[1], 0, 0, (1 - org[0]) * pix[1]], dtype=np.float64)
        hist["srow_y"] = np.array([0, pix[2], 0, (1 - org[1]) * pix[2]], dtype=np.float64)
        hist["srow_z"] = np.array([0, 0, pix[3], (1 - org[2]) * pix[3]], dtype=np.float64)
    if dt in (128, 511):
        if img.ndim < 4 or img.shape[3] != 3:
            raise ValueError("The NII structure does not appear to have 3 RGB color planes "
                             "in the 4th dimension")
        img = np.moveaxis(img, 3, 0)
    data = img.reshape(-1, order="F")
    if dt in (32, 1792):
        pair = np.empty(2 * data.size, dtype=np.float64)
        pair[0::2], pair[1::2] = np.real(data), np.imag(data)
        data = pair
    if np.dtype(typ).kind in "iu":
        data = _matlab_cast(np.real(data) if np.iscomplexobj(data) else data, typ)
    elif np.iscomplexobj(data):
        data = np.real(data).astype(typ)
    data = np.ascontiguousarray(data, dtype=np.dtype(typ).newbyteorder("<"))
    fn = os.fspath(filename)
    # gzip level 1: the stored bytes are the same whatever the level; a
    # higher one costs several times the time for a few per cent of size
    with (gzip.open(fn, "wb", compresslevel=1) if gz else open(fn, "wb")) as fh:
        fh.write(_header_bytes(hdr))
        fh.write(b"\x00" * 4)
        fh.write(memoryview(data).cast("B"))


def saveNii(filename, img, vsz=(1.0, 1.0, 1.0)):  # noqa: N802  (the reference's name)
    """saveNii.m: make_nii(img, vsz) written to filename (.nii or .nii.gz),
    creating its folder if needed; returns the dataset struct."""
    d = os.path.dirname(os.fspath(filename))
    if d:
        os.makedirs(d, exist_ok=True)
    nii = make_nii(img, vsz)
    save_nii(nii, filename)
    return nii
