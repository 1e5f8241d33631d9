"""Tonemap post step (Tonemap.frag): pins of the oracle by closed forms, and
the GPU kernel against the FP32 oracle bit for bit (NaN positions included)."""
import numpy as np
import pytest

from hellovulkan_amd import abi
from oracle import hvx_oracle as O
from tests.helpers import to_dev


def test_reinhard_with_pure_white_one_is_identity():
    """L (1 + L / 1) / (1 + L) = L: the shader's Reinhard leaves the colour
    unchanged (to rounding); zero luminance gives 0/0 = NaN."""
    rng = np.random.default_rng(1)
    x = np.exp(rng.normal(0, 2, (3, 2000))).astype(np.float32)
    y = O.tonemap(x, 0, gamma=False, dtype=np.float64)
    assert np.abs(y / x - 1).max() < 1e-6
    z = O.tonemap(np.zeros((3, 4), np.float32), 0, gamma=False)
    assert np.all(np.isnan(z))


def test_curves_closed_forms():
    aces = O.tonemap(np.array([[0.0, 1e6], [0.0, 1e6], [0.0, 1e6]], np.float32), 1, gamma=False, dtype=np.float64)
    assert np.all(aces[:, 0] == 0.0) and np.all(aces[:, 1] == 1.0)   # clamp at 1 (limit a/c = 1.033)
    film = O.tonemap(np.full((3, 2), 0.004, np.float32), 2, gamma=False, dtype=np.float64)
    assert np.all(film == 0.0)
    unc = O.tonemap(np.full((3, 1), 5.6, np.float32), 3, gamma=False, dtype=np.float64)
    assert np.allclose(unc, 1.0, atol=1e-6)   # exposure 2 maps 5.6 to the white point 11.2
    g = O.tonemap(np.array([[1.0, 0.25], [1.0, 0.25], [1.0, 0.25]], np.float32), 0, dtype=np.float64)
    assert np.allclose(g[:, 0], 1.0) and np.allclose(g[:, 1], 0.25 ** (1 / 2.2), rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("gamma", [True, False])
def test_tonemap_gpu_matches_oracle(ctx, op, gamma):
    import torch
    W, H = 173, 61
    rng = np.random.default_rng(op * 2 + gamma)
    x = np.exp(rng.normal(-1, 3, (3, W * H))).astype(np.float32)
    x[:, :5] = 0.0                   # black pixels
    x[:, 5:9] = [[1e-30], [1e30], [0.5]]  # extreme luminance
    d = to_dev(x)
    out = torch.empty(3 * W * H, dtype=torch.float32, device="cuda")
    packed = torch.empty(W * H, dtype=torch.int32, device="cuda")
    ctx.tonemap(d, W, H, op, gamma, out, packed)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(3, -1)
    ref = O.tonemap(x, op, gamma)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.array_equal(got[ok].view(np.uint32), ref[ok].view(np.uint32))
    # packed B8G8R8A8_UNORM
    pk = packed.cpu().numpy().view(np.uint32)
    exp = np.where(np.isnan(ref), 0.0, np.clip(ref, 0.0, 1.0))
    q = np.rint(exp * np.float32(255.0)).astype(np.uint32)
    assert np.array_equal(pk & 0xFF, q[2]) and np.array_equal((pk >> 8) & 0xFF, q[1])
    assert np.array_equal((pk >> 16) & 0xFF, q[0]) and np.all(pk >> 24 == 255)
