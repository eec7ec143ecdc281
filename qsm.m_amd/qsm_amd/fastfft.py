"""Host-side size logic of the FFT path: nextfastfft, padfastfft,
unpadfastfft (src/utils/padfastfft.m, unpadfastfft.m,
third_party/nextfastfft.m), the k-space dipole kernel
(src/utils/kernels/dipoleKernel.m) and cropIndices (src/utils/cropIndices.m).

These are integer/size computations plus the dipole formula; the transforms
themselves run on the GPU (qsm_amd.dipole).
"""
from __future__ import annotations

import bisect
import math

import numpy as np

_MAXVAL = 2 ** 20
_TABLE: list[int] | None = None


def _table() -> list[int]:
    """Sorted 2,3,5,7-smooth numbers <= 2^20 (nextfastfft.m:61-91)."""
    global _TABLE
    if _TABLE is None:
        vals = []
        p2 = 1
        while p2 <= _MAXVAL:
            p3 = p2
            while p3 <= _MAXVAL:
                p5 = p3
                while p5 <= _MAXVAL:
                    p7 = p5
                    while p7 <= _MAXVAL:
                        vals.append(p7)
                        p7 *= 7
                    p5 *= 5
                p3 *= 3
            p2 *= 2
        _TABLE = sorted(vals)
    return _TABLE


def nextfastfft(n):
    """Smallest 7-smooth number >= n (nextfastfft.m:93-123).  Inputs above
    2^20 are divided by 2^k until in range, looked up, and multiplied back;
    works elementwise on arrays."""
    arr = np.asarray(n, dtype=np.float64)
    out = np.empty(arr.shape, dtype=np.int64)
    t = _table()
    for idx, v in np.ndenumerate(arr):
        red = 0
        if v > _MAXVAL:
            red = math.ceil(math.log2(v / _MAXVAL))
            v = v / 2 ** red
        k = bisect.bisect_left(t, v)  # first table entry >= v
        out[idx] = t[k] * 2 ** red
    return int(out) if out.ndim == 0 else out


def padfastfft(u, padsize=None, padval=0.0, direction="both"):
    """padfastfft.m:1-54 with a constant pad value.  Returns the padded array
    (zero-padded by default)."""
    u = np.asarray(u)
    sz = np.array(u.shape, dtype=np.int64)
    pads = np.zeros(len(sz), dtype=np.int64) if padsize is None else np.asarray(padsize, dtype=np.int64).reshape(-1)
    if len(pads) < len(sz):
        pads = np.concatenate([pads, np.zeros(len(sz) - len(pads), dtype=np.int64)])
    elif len(pads) > len(sz):
        sz = np.concatenate([sz, np.ones(len(pads) - len(sz), dtype=np.int64)])
        u = u.reshape(tuple(sz), order="F")
    if direction not in ("pre", "post", "both"):
        raise ValueError("direction must be 'pre', 'post' or 'both'")
    if direction != "both":
        pad = nextfastfft(sz + pads) - sz
        pad[pads == 0] = 0
        widths = [(int(p), 0) if direction == "pre" else (0, int(p)) for p in pad]
    else:
        pad = nextfastfft(sz + 2 * pads) - sz
        pad[pads == 0] = 0
        oddsz = sz & 1
        oddpad = pad & 1
        oddpre = oddpad & oddsz
        oddpost = oddpad & (1 - oddsz)
        half = pad // 2
        widths = [(int(h + a), int(h + b)) for h, a, b in zip(half, oddpre, oddpost)]
    return np.pad(u, widths, mode="constant", constant_values=padval)


def unpad_slices(shape, sz0, padsize=None, direction="both"):
    """Index slices unpadfastfft.m:38-80 selects (0-based)."""
    sz = list(shape)
    sz0 = [int(s) for s in np.asarray(sz0).reshape(-1)]
    pads = [0] * len(sz) if padsize is None else [int(p) for p in np.asarray(padsize).reshape(-1)]
    pads += [0] * (len(sz) - len(pads))
    if len(sz0) != len(sz):
        raise ValueError("dimension mismatch")
    if any(a < b for a, b in zip(sz, sz0)):
        raise ValueError("unpad size must be smaller or equal to current size")
    sl = []
    for s1, s0, p in zip(sz, sz0, pads):
        if direction == "pre":
            sl.append(slice(s1 - s0, s1))
        elif direction == "post":
            sl.append(slice(0, s1))
        else:
            pad = 0 if p == 0 else int(nextfastfft(s0 + 2 * p)) - s0
            oddpre = (pad & 1) & (s0 & 1)
            oddpost = (pad & 1) & (1 - (s0 & 1))
            h = pad // 2
            sl.append(slice(h + oddpre, s1 - (h + oddpost)))
    return tuple(sl)


def unpadfastfft(u, sz0, padsize=None, direction="both"):
    return np.asarray(u)[unpad_slices(np.shape(u), sz0, padsize, direction)]


def normalised_bdir(bdir):
    """dipoleKernel.m:35-38: bdir is normalised when its norm is not 1."""
    b = np.asarray(bdir if bdir is not None else (0.0, 0.0, 1.0), dtype=np.float64).reshape(3)
    n = float(np.linalg.norm(b))
    return b / n if n != 1 else b


def kgrid(n: int, h: float) -> np.ndarray:
    """ifftshift-ed k-space coordinates of one axis (dipoleKernel.m:74-87
    followed by the ifftshift of :64): index m holds m*dk for m < n-floor(n/2)
    and (m-n)*dk above, with dk = 1/(n*h)."""
    dk = 1.0 / (n * h)
    m = np.arange(n, dtype=np.int64)
    low = n // 2
    return np.where(m < n - low, m, m - n).astype(np.float64) * dk


def dipole_kernel(sz, vsz, bdir=(0.0, 0.0, 1.0), dtype=np.float64):
    """k-space dipole kernel D = 1/3 - (k.b)^2/|k|^2, DC at index 0, D(0)=0
    (dipoleKernel.m:56-66), column-major (nx, ny, nz)."""
    sz = [int(s) for s in sz]
    vsz = np.broadcast_to(np.asarray(vsz, dtype=np.float64), (3,))
    b = normalised_bdir(bdir)
    X = kgrid(sz[0], vsz[0])[:, None, None]
    Y = kgrid(sz[1], vsz[1])[None, :, None]
    Z = kgrid(sz[2], vsz[2])[None, None, :]
    kz = (b[0] * X + b[1] * Y) + b[2] * Z
    k2 = (X * X + Y * Y) + Z * Z
    with np.errstate(invalid="ignore", divide="ignore"):
        D = 1.0 / 3.0 - (kz * kz) / k2
    D[0, 0, 0] = 0.0
    return np.asfortranarray(D.astype(dtype))


def crop_indices(mask):
    """cropIndices.m:7-16: the index ranges (0-based, sorted) where the mask
    has any non-zero along each axis."""
    m = np.asarray(mask) != 0
    ix = np.nonzero(m.any(axis=(1, 2)))[0]
    iy = np.nonzero(m.any(axis=(0, 2)))[0]
    iz = np.nonzero(m.any(axis=(0, 1)))[0]
    return ix, iy, iz
