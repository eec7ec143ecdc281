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
