"""Seeded synthetic inputs for the inversion (the reference ships no phantom
or simulation code, so these are this project's own harness).

Pipeline per volume:
  1. chi: a sum of randomly rotated ellipsoids (overlaps add), centres
     uniform inside the mask's bounding box, semi-axes and values uniform in
     the configured ranges;
  2. field = real(ifftn(D * fftn(pad(chi)))) cropped back, with pad =
     zero padding to nextfastfft(2*dims) (padfastfft 'both' with padsize =
     dims/2) and D the k-space dipole kernel for (vsz, bdir);
  3. field += N(0, (c * std(field[mask]))^2);
  4. the solver input is -lapw(field) (computed by the caller on the GPU or
     by the oracle in tests).
All randomness comes from numpy.random.default_rng(seed).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field as dfield

import numpy as np

from .fastfft import dipole_kernel, padfastfft, unpad_slices


@dataclass
class PhantomSpec:
    shape: tuple
    vsz: tuple = (1.0, 1.0, 1.0)
    bdir: tuple = (0.0, 0.0, 1.0)
    n_ellipsoids: int = 12
    semi_axes: tuple = (4.0, 14.0)
    chi_range: tuple = (-0.1, 0.1)
    mask: str = "sphere"          # sphere | ellipsoid | box
    mask_param: tuple = (28.0,)   # sphere radius / ellipsoid semi-axes (fractions of dims)
    noise: float = 0.01
    seed: int = 1
    extra: dict = dfield(default_factory=dict)


def make_mask(shape, kind, param):
    nx, ny, nz = shape
    c = [(n - 1) / 2.0 for n in shape]
    x = (np.arange(nx) - c[0])[:, None, None]
    y = (np.arange(ny) - c[1])[None, :, None]
    z = (np.arange(nz) - c[2])[None, None, :]
    if kind == "box":
        return np.ones(shape, dtype=bool, order="F")
    if kind == "sphere":
        r = float(param[0])
        m = x * x + y * y + z * z <= r * r
    elif kind == "ellipsoid":
        a = [float(p) * n for p, n in zip(param, shape)]
        m = (x / a[0]) ** 2 + (y / a[1]) ** 2 + (z / a[2]) ** 2 <= 1.0
    else:
        raise ValueError(f"unknown mask kind {kind!r}")
    return np.asfortranarray(m)


def _rotation(rng) -> np.ndarray:
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def ellipsoids(shape, mask, n, semi_axes, chi_range, rng) -> np.ndarray:
    chi = np.zeros(shape, dtype=np.float64, order="F")
    idx = [np.nonzero(np.asarray(mask).any(axis=ax))[0] for ax in ((1, 2), (0, 2), (0, 1))]
    lo = [float(i.min()) if i.size else 0.0 for i in idx]
    hi = [float(i.max()) if i.size else float(s - 1) for i, s in zip(idx, shape)]
    for _ in range(n):
        c = np.array([rng.uniform(a, b) for a, b in zip(lo, hi)])
        ax = rng.uniform(semi_axes[0], semi_axes[1], size=3)
        R = _rotation(rng)
        v = rng.uniform(chi_range[0], chi_range[1])
        r = float(ax.max())
        sl = []
        for d in range(3):
            a = max(0, int(np.floor(c[d] - r)))
            b = min(shape[d], int(np.ceil(c[d] + r)) + 1)
            sl.append(slice(a, b))
        if any(s.start >= s.stop for s in sl):
            continue
        gx = np.arange(sl[0].start, sl[0].stop)[:, None, None] - c[0]
        gy = np.arange(sl[1].start, sl[1].stop)[None, :, None] - c[1]
        gz = np.arange(sl[2].start, sl[2].stop)[None, None, :] - c[2]
        # coordinates in the ellipsoid frame: R^T (x - c)
        u = [R[0, d] * gx + R[1, d] * gy + R[2, d] * gz for d in range(3)]
        inside = (u[0] / ax[0]) ** 2 + (u[1] / ax[1]) ** 2 + (u[2] / ax[2]) ** 2 <= 1.0
        chi[tuple(sl)] += np.where(inside, v, 0.0)
    return chi


def dipole_field(chi, vsz, bdir, workers=None) -> np.ndarray:
    """real(ifftn(D * fftn(padfastfft(chi, dims/2)))) unpadded (host FFT;
    the harness's own data generation, not the measured path)."""
    import scipy.fft as sfft

    shape = chi.shape
    pads = [s // 2 for s in shape]
    up = padfastfft(chi, pads)
    D = dipole_kernel(up.shape, vsz, bdir)
    w = workers or min(16, os.cpu_count() or 1)
    # full complex transforms and the real part, as the toolbox idiom
    # real(ifft3(D .* fft3(x))): on even-length axes D is not symmetric
    # under k -> -k at the Nyquist planes when B0 has two or more non-zero
    # components, so a real-to-complex transform (which assumes Hermitian
    # symmetry) would give a different field
    F = sfft.fftn(up, axes=(0, 1, 2), workers=w)
    F *= D
    f = sfft.ifftn(F, axes=(0, 1, 2), workers=w).real
    return np.asfortranarray(f[unpad_slices(up.shape, shape, pads)])


def dipole_field_gpu(chi, vsz, bdir) -> np.ndarray:
    """dipole_field through the library's GPU dipole convolution
    (qsm_dipole_conv, float64): same padding, kernel and crop."""
    from .dipole import dipole_conv

    return dipole_conv(np.asfortranarray(chi), vsz, bdir, [s // 2 for s in chi.shape])


def _host_part(spec: PhantomSpec):
    """The random draws of a spec, in the order make() takes them: mask,
    ellipsoids, then the unit-variance noise (rng.normal(0, s) is
    s * rng.standard_normal() bit for bit)."""
    rng = np.random.default_rng(spec.seed)
    mask = make_mask(spec.shape, spec.mask, spec.mask_param)
    chi = ellipsoids(spec.shape, mask, spec.n_ellipsoids, spec.semi_axes,
                     spec.chi_range, rng)
    z = rng.standard_normal(size=spec.shape)
    return chi, mask, z


def _finish(spec, chi, mask, z, field):
    if field == "gpu":
        fld = dipole_field_gpu(chi, spec.vsz, spec.bdir)
    elif field == "host":
        fld = dipole_field(chi, spec.vsz, spec.bdir)
    else:
        raise ValueError(f"field must be 'host' or 'gpu', not {field!r}")
    sd = float(np.std(fld[mask])) if mask.any() else 0.0
    fld = fld + (spec.noise * sd) * z
    return {"chi": chi, "field": np.asfortranarray(fld), "mask": mask}


def make(spec: PhantomSpec, field: str = "host"):
    """Returns dict(chi, field, mask) for the spec (field with noise).
    field="gpu" synthesises the field with the GPU dipole convolution."""
    if field not in ("gpu", "host"):
        raise ValueError(f"field must be 'host' or 'gpu', not {field!r}")
    return _finish(spec, *_host_part(spec), field)


def _to_input(spec, ph):
    from . import fd

    d2f = -fd.lapw(ph["field"], spec.vsz)
    return np.asfortranarray(d2f), ph["mask"].astype(np.uint8, order="F")


def solver_input(spec: PhantomSpec):
    """The primal-dual solver's input for a spec, as sstgv.m:79-87 forms it:
    (d2f = -lapw(field) in float64, mask as uint8), both column-major.  The
    field is synthesised with the GPU dipole convolution and lapw runs on
    the GPU."""
    return _to_input(spec, make(spec, field="gpu"))


def solver_inputs(specs, workers: int = 8, dtype=np.float64):
    """solver_input for many specs, equal to calling it on each (d2f cast
    to dtype): the random draws run in a thread pool, at most 2*workers
    specs ahead of the consumer, and the GPU steps (dipole convolution,
    lapw) one at a time in the calling thread, in the order of specs."""
    import collections
    import concurrent.futures as cf

    specs = list(specs)
    out = []
    workers = max(1, workers)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        q = collections.deque()
        nxt = 0
        while len(out) < len(specs):
            while nxt < len(specs) and len(q) < 2 * workers:
                q.append(ex.submit(_host_part, specs[nxt]))
                nxt += 1
            s = specs[len(out)]
            d2f, m = _to_input(s, _finish(s, *q.popleft().result(), "gpu"))
            out.append((np.asfortranarray(d2f.astype(dtype, copy=False)), m))
    return out
