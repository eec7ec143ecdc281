"""Python binding of the oracle (oracle/lib/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  Arrays are
column-major (x fastest) like MATLAB; vector fields are component-major
(p has shape (nx, ny, nz, 3) in Fortran order).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "liboracle.so"

_sz = C.c_size_t
_p = C.c_void_p
_d = C.c_double
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        szp = C.POINTER(_sz)
        dp = C.POINTER(_d)
        L.orc_step_sizes.argtypes = [dp, dp, dp]
        L.orc_init_mask.argtypes = [_p, _p, szp]
        for t in "df":
            getattr(L, f"orc_tgv_dual_{t}").argtypes = [_p] * 8 + [szp, dp, _d, _d, _d]
            getattr(L, f"orc_tgv_primal_{t}").argtypes = [_p] * 10 + [szp, dp, _d]
            getattr(L, f"orc_tv_dual_{t}").argtypes = [_p] * 6 + [szp, dp, _d, _d]
            getattr(L, f"orc_tv_primal_{t}").argtypes = [_p] * 7 + [szp, dp, _d]
            f = getattr(L, f"orc_convergence_{t}")
            f.argtypes = [dp, dp, _p, _p, _p, _p, _d, _sz]
            f.restype = C.c_uint8
        L.orc_pd_tgv.argtypes = [C.c_int, _p, _p, szp, dp, _d, _d, _d, C.c_uint32,
                                 C.c_uint32, _p, C.POINTER(C.c_uint32), dp]
        L.orc_pd_tv.argtypes = [C.c_int, _p, _p, szp, dp, _d, _d, C.c_uint32,
                                C.c_uint32, _p, C.POINTER(C.c_uint32), dp]
        for n in ("gradfp", "gradf"):
            getattr(L, f"orc_{n}").argtypes = [C.c_int, _p, _p, _p, _p, szp, dp]
        for n in ("gradfp_adj", "gradf_adj"):
            getattr(L, f"orc_{n}").argtypes = [C.c_int, _p, _p, _p, _p, szp, dp]
        L.orc_gradfm.argtypes = [C.c_int, _p, _p, _p, _p, _p, szp, dp]
        for n in ("gradb", "gradc", "gradb_adj"):
            getattr(L, f"orc_{n}").argtypes = [C.c_int, _p, _p, _p, _p, szp, dp]
        L.orc_lap1.argtypes = [C.c_int, _p, _p, _p, szp, dp]
        for n in ("gradbm", "gradcm", "gradbm_adj"):
            getattr(L, f"orc_{n}").argtypes = [C.c_int, _p, _p, _p, _p, _p, szp, dp]
        L.orc_gradfm_adj.argtypes = [C.c_int, _p, _p, _p, _p, _p, szp, dp]
        L.orc_lapw.argtypes = [C.c_int, _p, _p, szp, dp]
        L.orc_last_loop_seconds.restype = _d
        L.orc_last_loop_seconds.argtypes = []
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _sz3(shape):
    return (_sz * 3)(*[int(s) for s in shape[:3]])


def _h3(h):
    h = np.asarray(h, dtype=np.float64).reshape(-1)
    if h.size == 1:
        h = np.repeat(h, 3)
    return (_d * 3)(*h.tolist())


def _F(a, dtype=None):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _code(dt):
    return 0 if np.dtype(dt) == np.float64 else 1


def _t(dt):
    return "d" if np.dtype(dt) == np.float64 else "f"


def _ptr(a):
    return a.ctypes.data


def step_sizes(h):
    t, s = _d(), _d()
    lib().orc_step_sizes(_h3(h), C.byref(t), C.byref(s))
    return t.value, s.value


def init_mask(m):
    m = _F(np.asarray(m) != 0, np.uint8)
    mi = np.zeros(m.shape, dtype=np.uint8, order="F")
    lib().orc_init_mask(_ptr(mi), _ptr(m), _sz3(m.shape))
    return mi


def pd_tgv(d2f, mask, h, lam, alpha, tol, maxiter, verbose=0):
    """Whole solve as pd_tgv_mex: returns (chi, iterations, (n1, n2))."""
    f = _F(d2f)
    m = _F(np.asarray(mask) != 0, np.uint8)
    chi = np.zeros(f.shape, dtype=f.dtype, order="F")
    it = C.c_uint32()
    nr = (_d * 2)()
    rc = lib().orc_pd_tgv(_code(f.dtype), _ptr(f), _ptr(m), _sz3(f.shape), _h3(h),
                          lam, alpha, tol, maxiter, verbose, _ptr(chi), C.byref(it), nr)
    if rc:
        raise ValueError("oracle rejected the arguments")
    return chi, it.value, (nr[0], nr[1])


def pd_tv(d2f, mask, h, lam, tol, maxiter, verbose=0):
    f = _F(d2f)
    m = _F(np.asarray(mask) != 0, np.uint8)
    chi = np.zeros(f.shape, dtype=f.dtype, order="F")
    it = C.c_uint32()
    nr = (_d * 2)()
    rc = lib().orc_pd_tv(_code(f.dtype), _ptr(f), _ptr(m), _sz3(f.shape), _h3(h),
                         lam, tol, maxiter, verbose, _ptr(chi), C.byref(it), nr)
    if rc:
        raise ValueError("oracle rejected the arguments")
    return chi, it.value, (nr[0], nr[1])


def last_loop_seconds() -> float:
    """Iteration-loop time of the last pd_tgv/pd_tv call."""
    return lib().orc_last_loop_seconds()


def set_threads(n: int) -> None:
    """OpenMP thread count of the restatement's loops."""
    lib().orc_set_threads(int(n))


# ---- single sweeps, in place on the given (Fortran-ordered) arrays ----------

def tgv_dual(nu, p, q, chi_, eta_, w_, f, mi, h, alpha0, alpha1, sigma):
    t = _t(nu.dtype)
    getattr(lib(), f"orc_tgv_dual_{t}")(_ptr(nu), _ptr(p), _ptr(q), _ptr(chi_), _ptr(eta_),
                                        _ptr(w_), _ptr(f), _ptr(mi), _sz3(nu.shape), _h3(h),
                                        alpha0, alpha1, sigma)


def tgv_primal(chi, eta, w, chi_, eta_, w_, nu, p, q, m, h, tau):
    t = _t(chi.dtype)
    getattr(lib(), f"orc_tgv_primal_{t}")(_ptr(chi), _ptr(eta), _ptr(w), _ptr(chi_), _ptr(eta_),
                                          _ptr(w_), _ptr(nu), _ptr(p), _ptr(q), _ptr(m),
                                          _sz3(chi.shape), _h3(h), tau)


def tv_dual(nu, p, chi_, eta_, f, mi, h, lam, sigma):
    t = _t(nu.dtype)
    getattr(lib(), f"orc_tv_dual_{t}")(_ptr(nu), _ptr(p), _ptr(chi_), _ptr(eta_), _ptr(f),
                                       _ptr(mi), _sz3(nu.shape), _h3(h), lam, sigma)


def tv_primal(chi, eta, chi_, eta_, nu, p, m, h, tau):
    t = _t(chi.dtype)
    getattr(lib(), f"orc_tv_primal_{t}")(_ptr(chi), _ptr(eta), _ptr(chi_), _ptr(eta_), _ptr(nu),
                                         _ptr(p), _ptr(m), _sz3(chi.shape), _h3(h), tau)


def convergence(chi, chi0, eta, eta0, tol):
    t = _t(chi.dtype)
    n1, n2 = _d(), _d()
    flag = getattr(lib(), f"orc_convergence_{t}")(C.byref(n1), C.byref(n2), _ptr(chi), _ptr(chi0),
                                                  _ptr(eta), _ptr(eta0), tol, chi.size)
    return n1.value, n2.value, int(flag)


# ---- finite differences (allocating) ----------------------------------------

def _z(u):
    return np.zeros(u.shape, dtype=u.dtype, order="F")


def gradfp(u, h):
    u = _F(u)
    d = [_z(u) for _ in range(3)]
    lib().orc_gradfp(_code(u.dtype), *map(_ptr, d), _ptr(u), _sz3(u.shape), _h3(h))
    return tuple(d)


def gradf(u, h):
    u = _F(u)
    d = [_z(u) for _ in range(3)]
    lib().orc_gradf(_code(u.dtype), *map(_ptr, d), _ptr(u), _sz3(u.shape), _h3(h))
    return tuple(d)


def gradfm(u, G, h, init=None):
    u = _F(u)
    g = _F(np.asarray(G) != 0, np.uint8)
    d = [_z(u) if init is None else _F(init[c]).copy(order="F") for c in range(3)]
    lib().orc_gradfm(_code(u.dtype), *map(_ptr, d), _ptr(u), _ptr(g), _sz3(u.shape), _h3(h))
    return tuple(d)


def gradfp_adj(x, y, z, h):
    x, y, z = _F(x), _F(y), _F(z)
    du = _z(x)
    lib().orc_gradfp_adj(_code(x.dtype), _ptr(du), _ptr(x), _ptr(y), _ptr(z), _sz3(x.shape), _h3(h))
    return du


def gradf_adj(x, y, z, h):
    x, y, z = _F(x), _F(y), _F(z)
    du = _z(x)
    lib().orc_gradf_adj(_code(x.dtype), _ptr(du), _ptr(x), _ptr(y), _ptr(z), _sz3(x.shape), _h3(h))
    return du


def gradfm_adj(x, y, z, G, h, init=None):
    x, y, z = _F(x), _F(y), _F(z)
    g = _F(np.asarray(G) != 0, np.uint8)
    du = _z(x) if init is None else _F(init).copy(order="F")
    lib().orc_gradfm_adj(_code(x.dtype), _ptr(du), _ptr(x), _ptr(y), _ptr(z), _ptr(g),
                         _sz3(x.shape), _h3(h))
    return du


def _vec3(name, u, h, G=None, init=None):
    u = _F(u)
    d = [_z(u) if init is None else _F(init[c]).copy(order="F") for c in range(3)]
    if G is None:
        getattr(lib(), name)(_code(u.dtype), *map(_ptr, d), _ptr(u), _sz3(u.shape), _h3(h))
    else:
        g = _F(np.asarray(G) != 0, np.uint8)
        getattr(lib(), name)(_code(u.dtype), *map(_ptr, d), _ptr(u), _ptr(g), _sz3(u.shape),
                             _h3(h))
    return tuple(d)


def gradb(u, h, init=None):
    return _vec3("orc_gradb", u, h, None, init)


def gradbm(u, G, h, init=None):
    return _vec3("orc_gradbm", u, h, G, init)


def gradc(u, h, init=None):
    return _vec3("orc_gradc", u, h, None, init)


def gradcm(u, G, h, init=None):
    return _vec3("orc_gradcm", u, h, G, init)


def gradb_adj(x, y, z, h, init=None):
    x, y, z = _F(x), _F(y), _F(z)
    du = _z(x) if init is None else _F(init).copy(order="F")
    lib().orc_gradb_adj(_code(x.dtype), _ptr(du), _ptr(x), _ptr(y), _ptr(z), _sz3(x.shape), _h3(h))
    return du


def gradbm_adj(x, y, z, G, h, init=None):
    x, y, z = _F(x), _F(y), _F(z)
    g = _F(np.asarray(G) != 0, np.uint8)
    du = _z(x) if init is None else _F(init).copy(order="F")
    lib().orc_gradbm_adj(_code(x.dtype), _ptr(du), _ptr(x), _ptr(y), _ptr(z), _ptr(g),
                         _sz3(x.shape), _h3(h))
    return du


def lap1(u, G, h, init=None):
    u = _F(u)
    g = _F(np.asarray(G) != 0, np.uint8)
    d = _z(u) if init is None else _F(init).copy(order="F")
    lib().orc_lap1(_code(u.dtype), _ptr(d), _ptr(u), _ptr(g), _sz3(u.shape), _h3(h))
    return d


def lapw(u, h, init=None):
    u = _F(u)
    d = _z(u) if init is None else _F(init).copy(order="F")
    lib().orc_lapw(_code(u.dtype), _ptr(d), _ptr(u), _sz3(u.shape), _h3(h))
    return d
