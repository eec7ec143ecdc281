"""The device utilities behind the bandwidth references: qsm_copy_dev (the
plain copy the HBM counters are calibrated against) and qsm_stream_dev (the
15-in, 15-out state traffic of a fused pass, its access pattern's ceiling)
compute what they claim."""
import ctypes as C

import numpy as np
import pytest

from qsm_amd import QsmError, _lib
from qsm_amd._lib import check, lib

pytestmark = pytest.mark.gpu

_lib.register("qsm_copy_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p])
_lib.register("qsm_stream_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p])


@pytest.mark.parametrize("nbytes", [16, 4096 + 7, (1 << 20) + 13])
def test_copy_dev(cuda, nbytes):
    import torch

    a = torch.randint(0, 255, (nbytes,), dtype=torch.uint8, device="cuda")
    b = torch.zeros(nbytes + 16, dtype=torch.uint8, device="cuda")
    check(lib().qsm_copy_dev(b.data_ptr(), a.data_ptr(), nbytes, None))
    torch.cuda.synchronize()
    assert torch.equal(b[:nbytes], a) and int(b[nbytes:].sum()) == 0


@pytest.mark.parametrize("narr,n", [(15, 100003), (1, 4097), (16, 256)])
def test_stream_dev(cuda, narr, n):
    import torch

    a = torch.randn(narr * n, dtype=torch.float32, device="cuda")
    b = torch.zeros_like(a)
    check(lib().qsm_stream_dev(b.data_ptr(), a.data_ptr(), n, narr, None))
    torch.cuda.synchronize()
    x = a.view(narr, n).cpu().numpy()
    want = x * np.float32(1.0001) + np.roll(x, -1, axis=0)
    np.testing.assert_array_equal(b.view(narr, n).cpu().numpy(), want)


def test_stream_dev_rejects_narr(cuda):
    with pytest.raises(QsmError, match="narr"):
        check(lib().qsm_stream_dev(1, 1, 4, 17, None))
