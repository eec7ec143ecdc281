#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ (small, seeded).

The reference ships no test vectors for this path and cannot be built or run
here (MATLAB MEX), so the vectors come from the oracle (the C restatement of
the reference's loops, itself pinned by closed forms and invariants in
tests/test_oracle_pins.py and tests/test_oracle_fd.py) and, for the FFT
parts, from numpy's pocketfft.  They freeze those outputs so that a later
change to the oracle or to the GPU path shows up as a mismatch.

    python tests/golden/make_golden.py          # rewrites the .npz files
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
for p in (ROOT / "qsm.m_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

from oracle import oracle as orc  # noqa: E402
from qsm_amd import fastfft  # noqa: E402

PD_SHAPE = (12, 10, 9)
FD_SHAPE = (9, 8, 7)
H = (0.9, 1.0, 1.1)


def _mask(shape, seed):
    rng = np.random.default_rng(seed)
    c = [(n - 1) / 2 for n in shape]
    g = np.meshgrid(*(np.arange(n) - cc for n, cc in zip(shape, c)), indexing="ij")
    m = sum((x / (0.45 * n)) ** 2 for x, n in zip(g, shape)) <= 1.0
    m &= rng.random(shape) > 0.05
    return np.asfortranarray(m.astype(np.uint8))


def pd_cases():
    rng = np.random.default_rng(20)
    out = {}
    f = np.asfortranarray(rng.standard_normal(PD_SHAPE) * 1e-2)
    m = _mask(PD_SHAPE, 21)
    out["f"] = f
    out["mask"] = m
    for dt, tag in ((np.float64, "f64"), (np.float32, "f32")):
        fd_ = np.asfortranarray(f.astype(dt))
        chi, it, nr = orc.pd_tgv(fd_, m, H, 2e-3, 2.0, 0.0, 40)
        out[f"tgv_{tag}_chi"], out[f"tgv_{tag}_norms"] = chi, np.array(nr)
        chi, it, nr = orc.pd_tv(fd_, m, H, 3e-3, 0.0, 40)
        out[f"tv_{tag}_chi"], out[f"tv_{tag}_norms"] = chi, np.array(nr)
    # a tolerance stop (float64): iteration count and chi at the stop
    chi, it, nr = orc.pd_tgv(f, m, H, 2e-3, 2.0, 2e-2, 400)
    out["tgv_tol_chi"], out["tgv_tol_iters"], out["tgv_tol_norms"] = chi, np.array(it), np.array(nr)
    return out


def fd_cases():
    rng = np.random.default_rng(30)
    out = {}
    u = rng.standard_normal(FD_SHAPE)
    x, y, z = (rng.standard_normal(FD_SHAPE) for _ in range(3))
    G = (rng.random(FD_SHAPE) < 0.7).astype(np.uint8)
    ph = 4.0 * rng.standard_normal(FD_SHAPE)  # phase far outside [-pi, pi]
    out.update(u=u, x=x, y=y, z=z, G=G, phase=ph)
    for dt, tag in ((np.float64, "f64"), (np.float32, "f32")):
        U, X, Y, Z, P = (np.asfortranarray(a.astype(dt)) for a in (u, x, y, z, ph))
        for name, res in (("gradfp", orc.gradfp(U, H)), ("gradf", orc.gradf(U, H)),
                          ("gradfm", orc.gradfm(U, G, H))):
            for c, a in zip("xyz", res):
                out[f"{name}_{tag}_d{c}"] = a
        out[f"gradfp_adj_{tag}"] = orc.gradfp_adj(X, Y, Z, H)
        out[f"gradf_adj_{tag}"] = orc.gradf_adj(X, Y, Z, H)
        out[f"gradfm_adj_{tag}"] = orc.gradfm_adj(X, Y, Z, G, H)
        out[f"lapw_{tag}"] = orc.lapw(P, H)
    return out


def fft_cases():
    rng = np.random.default_rng(40)
    out = {}
    x = rng.standard_normal((14, 10, 9)) + 1j * rng.standard_normal((14, 10, 9))
    out["fft_x"] = x
    out["fft_X"] = np.fft.fftn(x)
    chi = rng.standard_normal((10, 12, 8))
    vsz, bdir, pad = (0.9, 1.0, 1.2), (0.1, -0.2, 0.97), (5, 6, 4)
    up = fastfft.padfastfft(chi, pad)
    D = fastfft.dipole_kernel(up.shape, vsz, bdir)
    out["dip_chi"] = chi
    out["dip_vsz"], out["dip_bdir"], out["dip_pad"] = np.array(vsz), np.array(bdir), np.array(pad)
    out["dip_D"] = D
    out["dip_field"] = np.real(np.fft.ifftn(D * np.fft.fftn(up)))[fastfft.unpad_slices(up.shape, chi.shape, pad)]
    return out


def main():
    for name, fn in (("pd", pd_cases), ("fd", fd_cases), ("fft", fft_cases)):
        data = fn()
        np.savez_compressed(HERE / f"{name}.npz", **data)
        print(name, sorted(data))


if __name__ == "__main__":
    main()
