"""Exhaustive gfx950 checks of the fast correctly rounded square root and
reciprocal used by the shading kernel's exact quantities: every float of the
tested binades must equal the compiler's IEEE sqrtf / division bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "libhvx_testkernels.so")


def _run(first_bits, n, which):
    import torch
    lib = C.CDLL(LIB)
    lib.hvxt_fastmath_check.argtypes = [C.c_uint32, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p]
    mism = torch.zeros(1, dtype=torch.int64, device="cuda")
    first = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    assert lib.hvxt_fastmath_check(first_bits, n, which, mism.data_ptr(), first.data_ptr()) == 0
    torch.cuda.synchronize()
    return int(mism.item()), int(first.item()) & 0xFFFFFFFF


def _bits(x):
    return int(np.array([x], np.float32).view(np.uint32)[0])


@pytest.mark.parametrize("lo,hi", [(1.0, 4.0), (2.0 ** -90, 2.0 ** -88), (2.0 ** 100, 2.0 ** 102)])
def test_sqrt_rn_normal_exhaustive(gpu, lo, hi):
    a, b = _bits(lo), _bits(hi)
    mism, first = _run(a, b - a, 0)
    assert mism == 0, f"{mism} mismatches, first at bits {first:#x}"


@pytest.mark.parametrize("lo,hi", [(1.0, 2.0), (2.0 ** -120, 2.0 ** -119), (2.0 ** 120, 2.0 ** 121),
                                   (-2.0, -1.0)])
def test_rcp_rn_normal_exhaustive(gpu, lo, hi):
    a, b = sorted((_bits(lo), _bits(hi)))
    mism, first = _run(a, b - a, 1)
    assert mism == 0, f"{mism} mismatches, first at bits {first:#x}"
