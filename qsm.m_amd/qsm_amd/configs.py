"""The benchmark workloads (seeded; sizes, phantoms, regularisation).

C1  TV, 64^3, 1 mm, B0 = z, 12 ellipsoids (semi-axes 4-14 vox), chi in
    [-0.1, 0.1] ppm, sphere mask r = 28, noise 0.01*std, lam = 1e-3, 500 it.
C2  TGV, 224x224x128, 0.9x0.9x1.2 mm, B0 tilted 15 deg about x
    (0, -sin 15, cos 15), 40 ellipsoids (semi-axes 6-30 vox), chi in
    [-0.15, 0.15], ellipsoidal mask (0.40, 0.45, 0.40)*dims, noise 0.02,
    alpha1 = 3e-3, alpha0 = 2*alpha1, 1000 it.
C3  256x256x176, 1 mm, B0 = z, 60 ellipsoids (semi-axes 6-32 vox), chi in
    [-0.2, 0.2], noise 0.01, 1000 it; TGV alpha1 = 1e-3, alpha0 = 2*alpha1
    (headline) and TV lam = 1e-3; full-box mask (roofline headline) or
    ellipsoidal brain mask.
C4  48 inversions, sizes uniform over {96^3, 128^3, 160x160x96,
    224x224x128, 256x256x176}, TV/TGV 50/50, each with its own seeded
    phantom, ellipsoidal mask and B0 within 20 deg of z; FP32 bulk.
C5  Poisson arrivals (20 req/s), TV/TGV 50/50, sizes 70 % 224x224x128,
    20 % 256x256x176, 10 % 320x320x192, each with its own seeded phantom,
    voxel size and an iteration budget uniform in [300, 1000]; FP32 bulk.
The primal-dual wave operator assumes B0 = z (pd_tgv_mex.c:380-383), so a
tilted B0 only changes the synthetic field (a model mismatch the reference
has too).
"""
from __future__ import annotations

import math

from .phantom import PhantomSpec

_S15 = math.sin(math.radians(15.0))
_C15 = math.cos(math.radians(15.0))

CONFIGS = {
    "c1": dict(kind="tv", lam=1e-3, alpha=0.0, iters=500,
               spec=PhantomSpec(shape=(64, 64, 64), vsz=(1.0, 1.0, 1.0), bdir=(0.0, 0.0, 1.0),
                                n_ellipsoids=12, semi_axes=(4.0, 14.0), chi_range=(-0.1, 0.1),
                                mask="sphere", mask_param=(28.0,), noise=0.01, seed=1)),
    "c2": dict(kind="tgv", lam=3e-3, alpha=2.0, iters=1000,
               spec=PhantomSpec(shape=(224, 224, 128), vsz=(0.9, 0.9, 1.2), bdir=(0.0, -_S15, _C15),
                                n_ellipsoids=40, semi_axes=(6.0, 30.0), chi_range=(-0.15, 0.15),
                                mask="ellipsoid", mask_param=(0.40, 0.45, 0.40), noise=0.02, seed=2)),
    "c3": dict(kind="tgv", lam=1e-3, alpha=2.0, iters=1000,
               spec=PhantomSpec(shape=(256, 256, 176), vsz=(1.0, 1.0, 1.0), bdir=(0.0, 0.0, 1.0),
                                n_ellipsoids=60, semi_axes=(6.0, 32.0), chi_range=(-0.2, 0.2),
                                mask="box", mask_param=(), noise=0.01, seed=3)),
    "c3_brain": dict(kind="tgv", lam=1e-3, alpha=2.0, iters=1000,
                     spec=PhantomSpec(shape=(256, 256, 176), vsz=(1.0, 1.0, 1.0), bdir=(0.0, 0.0, 1.0),
                                      n_ellipsoids=60, semi_axes=(6.0, 32.0), chi_range=(-0.2, 0.2),
                                      mask="ellipsoid", mask_param=(0.40, 0.45, 0.40), noise=0.01, seed=3)),
}


C4_SIZES = [(96, 96, 96), (128, 128, 128), (160, 160, 96), (224, 224, 128), (256, 256, 176)]
C5_SIZES = [(224, 224, 128), (256, 256, 176), (320, 320, 192)]
C5_PROBS = [0.7, 0.2, 0.1]
VOXEL_SIZES = [(1.0, 1.0, 1.0), (0.9, 0.9, 1.2), (0.8, 0.8, 1.0), (1.0, 1.0, 1.5)]


def _tilted(rng, max_deg):
    """A unit B0 direction at a uniform angle <= max_deg from z."""
    th = math.radians(rng.uniform(0.0, max_deg))
    ph = rng.uniform(0.0, 2.0 * math.pi)
    return (math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th))


def _spec(shape, vsz, bdir, seed):
    scale = min(shape) / 128.0
    return PhantomSpec(shape=shape, vsz=vsz, bdir=bdir, n_ellipsoids=30,
                       semi_axes=(4.0 * scale, 20.0 * scale), chi_range=(-0.15, 0.15),
                       mask="ellipsoid", mask_param=(0.42, 0.45, 0.42), noise=0.01,
                       seed=seed)


def c4_volumes(n=48, seed=4):
    """The C4 batch: list of dict(kind, lam, alpha, spec)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        shape = C4_SIZES[int(rng.integers(len(C4_SIZES)))]
        kind = "tgv" if rng.random() < 0.5 else "tv"
        out.append(dict(kind=kind, lam=1e-3, alpha=2.0 if kind == "tgv" else 0.0,
                        spec=_spec(shape, (1.0, 1.0, 1.0), _tilted(rng, 20.0), 1000 + i)))
    return out


def c5_requests(n=40, rate=20.0, seed=5):
    """The C5 request stream: list of dict(arrival, kind, lam, alpha,
    maxiter, spec), arrivals of a Poisson process of the given rate."""
    import numpy as np

    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    out = []
    for i in range(n):
        shape = C5_SIZES[int(rng.choice(len(C5_SIZES), p=C5_PROBS))]
        kind = "tgv" if rng.random() < 0.5 else "tv"
        vsz = VOXEL_SIZES[int(rng.integers(len(VOXEL_SIZES)))]
        out.append(dict(arrival=float(t[i]), kind=kind, lam=1e-3,
                        alpha=2.0 if kind == "tgv" else 0.0,
                        maxiter=int(rng.integers(300, 1001)),
                        spec=_spec(shape, vsz, _tilted(rng, 20.0), 5000 + i)))
    return out
