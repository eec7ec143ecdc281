"""Shared helpers of the test suite (device staging, seeded states)."""
from __future__ import annotations

import numpy as np


def to_dev(a, device="cuda:0"):
    """Fortran-order flattening of a numpy array into a torch device tensor."""
    import torch

    flat = np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))
    return torch.from_numpy(flat).to(device)


def from_dev(t, shape):
    return np.asarray(t.detach().cpu().numpy()).reshape(shape, order="F")


def rand_field(rng, shape, dtype, scale=1.0):
    return np.asfortranarray((rng.standard_normal(shape) * scale).astype(dtype))


def brainish_mask(shape, seed=0, frac=0.8):
    """An ellipsoidal mask with a few holes punched in (exercises the
    interior-mask semantics: a voxel outside m whose six neighbours are in m
    is still updated by the dual)."""
    rng = np.random.default_rng(seed)
    c = [(n - 1) / 2 for n in shape]
    x = (np.arange(shape[0]) - c[0])[:, None, None] / (frac * shape[0] / 2)
    y = (np.arange(shape[1]) - c[1])[None, :, None] / (frac * shape[1] / 2)
    z = (np.arange(shape[2]) - c[2])[None, None, :] / (frac * shape[2] / 2)
    m = (x * x + y * y + z * z) <= 1.0
    idx = np.argwhere(m)
    if len(idx):
        for r in idx[rng.choice(len(idx), size=max(1, len(idx) // 50), replace=False)]:
            m[tuple(r)] = False
    return np.asfortranarray(m.astype(np.uint8))
