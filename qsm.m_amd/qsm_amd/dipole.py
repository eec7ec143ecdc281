"""FFT dipole convolution and 3-D FFTs on the GPU (the toolbox idioms
real(ifft3(D .* fft3(padfastfft(chi, pad)))) with D from dipoleKernel.m, and
fft3.m / ifft3.m).  Thin bindings of qsm_dipole_conv* / qsm_fft3_dev."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, dtype_code, fortran, h3, lib, ptr, sz3

_sz = C.c_size_t
_p = C.c_void_p
_dp = C.POINTER(C.c_double)
_lib.register("qsm_nextfastfft", _sz, [_sz])
_lib.register("qsm_fft3_dev", C.c_int, [C.c_int, _p, C.POINTER(_sz), C.c_int, _p])
_lib.register("qsm_dipole_pad_size", C.c_int, [C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz)])
_lib.register("qsm_dipole_work_bytes", _sz, [C.c_int, C.POINTER(_sz), C.POINTER(_sz)])
_lib.register("qsm_dipole_conv_dev", C.c_int, [C.c_int, _p, C.POINTER(_sz), _dp, _dp, C.POINTER(_sz), _p, _p, _p])
_lib.register("qsm_dipole_conv", C.c_int, [C.c_int, _p, C.POINTER(_sz), _dp, _dp, C.POINTER(_sz), _p])


def _d3(v):
    v = np.asarray(v, dtype=np.float64).reshape(-1)
    return (C.c_double * 3)(*v.tolist())


def nextfastfft(n: int) -> int:
    return int(lib().qsm_nextfastfft(int(n)))


def pad_size(shape, pad):
    M, off = (_sz * 3)(), (_sz * 3)()
    check(lib().qsm_dipole_pad_size(sz3(shape), sz3(pad), M, off))
    return tuple(M), tuple(off)


def work_bytes(dtype, shape, pad) -> int:
    """Bytes of device work space dipole_conv_dev needs (dtype: numpy or
    torch float32/float64): the half spectrum along x of the padded
    volume."""
    code = _lib.F32 if str(dtype).endswith("float32") else _lib.F64
    return int(lib().qsm_dipole_work_bytes(code, sz3(shape), sz3(pad)))


def dipole_conv(chi, vsz=(1.0, 1.0, 1.0), bdir=(0.0, 0.0, 1.0), pad=(0, 0, 0)):
    """field = real(ifft3(D .* fft3(padfastfft(chi, pad)))) cropped back."""
    chi = fortran(chi)
    if chi.dtype not in (np.float32, np.float64):
        chi = chi.astype(np.float64, order="F")
    out = np.zeros(chi.shape, dtype=chi.dtype, order="F")
    check(lib().qsm_dipole_conv(dtype_code(chi.dtype), ptr(chi), sz3(chi.shape), h3(vsz),
                                _d3(bdir), sz3(pad), ptr(out)))
    return out


def dipole_conv_dev(chi, vsz, bdir, pad, out, work=None, stream=None):
    """Device form on torch tensors (flat, Fortran order); allocates the
    complex work space with torch when none is given."""
    import torch

    if chi.dim() != 3:
        raise ValueError("chi must be a 3-D tensor (use a Fortran-order view)")
    shape = tuple(int(s) for s in chi.shape)
    code = _lib.F64 if chi.dtype == torch.float64 else _lib.F32
    nb = work_bytes(chi.dtype, shape, pad)
    own = work is None
    if own:
        work = torch.empty(nb, dtype=torch.uint8, device=chi.device)
    elif work.numel() * work.element_size() < nb:
        raise ValueError(f"work space too small: {nb} bytes needed")
    sp = stream.cuda_stream if stream is not None else None
    check(lib().qsm_dipole_conv_dev(code, chi.data_ptr(), sz3(shape), h3(vsz), _d3(bdir),
                                    sz3(pad), out.data_ptr(), work.data_ptr(), sp))
    if own and stream is not None:
        # the call is asynchronous: keep the block from being reused by
        # other streams before this one has run it
        work.record_stream(stream)
    return out


def fft3(x, inverse=False):
    """fft3.m / ifft3.m of a complex host array (column-major), on the GPU."""
    import torch

    x = np.asfortranarray(np.asarray(x))
    if x.dtype not in (np.complex64, np.complex128):
        x = x.astype(np.complex128, order="F")
    shape = x.shape
    flat = np.ascontiguousarray(x.reshape(-1, order="F"))
    t = torch.from_numpy(flat.view(np.float64 if x.dtype == np.complex128 else np.float32)).cuda()
    code = _lib.F64 if x.dtype == np.complex128 else _lib.F32
    check(lib().qsm_fft3_dev(code, t.data_ptr(), sz3(shape), 1 if inverse else 0, None))
    out = t.cpu().numpy().view(x.dtype)
    return out.reshape(shape, order="F")


def ifft3(X):
    return fft3(X, inverse=True)
