"""Exhaustive gfx950 checks of the fast correctly rounded square root and
reciprocal used by the shading kernel's exact quantities and the specular
filter's normalisations: every float of their stated domains must equal the
compiler's IEEE sqrtf / division bit for bit."""
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


def _binades(lo_exp, hi_exp, sign=0):
    """Every float with exponent in [lo_exp, hi_exp), one launch per binade."""
    for e in range(lo_exp, hi_exp):
        yield (sign << 31) | ((e + 127) << 23), 1 << 23


def test_sqrt_rn_normal_exhaustive(gpu):
    """sqrt_rn_normal equals sqrtf on every float of [2^-102, 2^128)."""
    bad = []
    for first, n in _binades(-102, 128):
        mism, at = _run(first, n, 0)
        if mism:
            bad.append((hex(first), mism, hex(at)))
    assert not bad, f"binades with mismatches (first bits, count, first bad): {bad[:8]}"


def test_rcp_rn_normal_exhaustive(gpu):
    """rcp_rn_normal equals 1.0f / b on every float with |b| in [2^-125, 2^125)."""
    bad = []
    for sign in (0, 1):
        for first, n in _binades(-125, 125, sign):
            mism, at = _run(first, n, 1)
            if mism:
                bad.append((hex(first), mism, hex(at)))
    assert not bad, f"binades with mismatches (first bits, count, first bad): {bad[:8]}"


@pytest.mark.parametrize("bf16", [0, 1], ids=["f16", "bf16"])
def test_mfma_16x16x32_lane_layout(gpu, bf16):
    """The operand/result lane layout the IBL contraction kernels assume for
    v_mfma_f32_16x16x32_{f16,bf16}: the product of exactly representable
    operands must equal the float64 matrix product."""
    import torch
    lib = C.CDLL(LIB)
    lib.hvxt_mfma_layout.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(3 + bf16)
    # small integers times powers of two: exact in both f16 and bf16, products exact in f32
    a = (rng.integers(-8, 9, (16, 32)) * 2.0 ** rng.integers(-3, 3, (16, 32))).astype(np.float32)
    b = (rng.integers(-8, 9, (32, 16)) * 2.0 ** rng.integers(-3, 3, (32, 16))).astype(np.float32)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    dc = torch.full((16, 16), np.nan, dtype=torch.float32, device="cuda")
    assert lib.hvxt_mfma_layout(bf16, da.data_ptr(), db.data_ptr(), dc.data_ptr()) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dc.cpu().numpy(), (a.astype(np.float64) @ b.astype(np.float64)).astype(np.float32))


def test_mlp_encoding_matches_cpu_bits(gpu):
    """The radiance network's FP32 encoding is the same function on the GPU
    and in the oracle's BF16 emulation: bit-identical sin/cos for every
    frequency over a dense grid of [-1, 1] and random floats in it."""
    import torch
    from oracle import hvx_oracle as O
    lib = C.CDLL(LIB)
    lib.hvxt_mlp_encode.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
    rng = np.random.default_rng(17)
    x = np.concatenate([np.linspace(-1, 1, 100001), rng.uniform(-1, 1, 100000), [0.0, -0.0, 1e-30]]).astype(np.float32)
    dx = torch.from_numpy(x).cuda()
    dsc = torch.full((x.size * 32,), np.nan, dtype=torch.float32, device="cuda")
    assert lib.hvxt_mlp_encode(dx.data_ptr(), x.size, dsc.data_ptr()) == 0
    torch.cuda.synchronize()
    got = dsc.cpu().numpy().reshape(-1, 16, 2)
    ref = O.mlp_sincos(x)
    diff = np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32))
    assert diff == 0, f"{diff} of {ref.size} values differ"
