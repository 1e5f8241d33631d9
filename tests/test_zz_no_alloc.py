"""hvx_cf_cull_lights_noalloc (the culling entry point that allocates
nothing, on the library's fixed per-GPU scratch) against hvx_cf_cull_lights,
through the C ABI: two calls in a row, the second on a smaller input, each
bit for bit equal to the allocating entry point's lists, counts, offsets,
total and overflow count; then a call past the fixed scratch is refused
before launch. Collected last, so a failure here hides no other test."""
import numpy as np
import pytest

from hellovulkan_amd import abi
from tests.helpers import SMALL_FRAME, UHD_FRAME, camera, to_dev

pytestmark = pytest.mark.gpu


def _cull(ctx, ubo, d_aabbs, d_lights, n, cap, noalloc):
    import torch
    Cn = ubo.slice_count_x * ubo.slice_count_y * ubo.slice_count_z
    cells = torch.full((Cn * 2,), -7, dtype=torch.int32, device="cuda")
    idx = torch.full((Cn * cap,), -7, dtype=torch.int32, device="cuda")
    tot = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    ov = torch.full((1,), -7, dtype=torch.int32, device="cuda")
    f = ctx.cull_lights_noalloc if noalloc else ctx.cull_lights
    f(ubo, d_aabbs, d_lights, n, cap, cells, idx, Cn * cap, tot, ov)
    torch.cuda.synchronize()
    t = int(tot.item())
    return cells.cpu().numpy(), idx.cpu().numpy()[:t], t, int(ov.item())


def _aabbs(ctx, ubo):
    import torch
    Cn = ubo.slice_count_x * ubo.slice_count_y * ubo.slice_count_z
    d = torch.empty(Cn * 8, dtype=torch.float32, device="cuda")
    ctx.build_aabbs(ubo, d)
    return d


def test_noalloc_cull_equals_allocating_cull_twice(gpu):
    """A fresh context (its own growable scratch still empty) calls the
    allocation-free entry point at 3840x2160 / 65,536 lights, then at
    2560x1600 / 1,024 lights; each output equals hvx_cf_cull_lights'. The
    second call reuses the same fixed array, which still holds the first
    call's lists past the smaller grid."""
    ctx = abi.Context(0)
    try:
        for cfg, seed in ((UHD_FRAME, 31), (SMALL_FRAME, 32)):
            _, _, ubo, _ = camera(cfg["width"], cfg["height"], cfg["grid"])
            n, cap = cfg["n_lights"], cfg["cap"]
            print(f"noalloc cull {cfg['width']}x{cfg['height']} lights {n} cap {cap}", flush=True)
            lights = abi.scene_lights(seed, n)
            d_aabbs = _aabbs(ctx, ubo)
            d_lights = to_dev(lights.view(np.uint8))
            gen0 = ctx.scratch_generation
            got = _cull(ctx, ubo, d_aabbs, d_lights, n, cap, noalloc=True)
            assert ctx.scratch_generation == gen0  # the context's scratch was not touched
            ref = _cull(ctx, ubo, d_aabbs, d_lights, n, cap, noalloc=False)
            assert got[2] == ref[2] and got[2] > 0
            assert got[3] == ref[3]
            assert np.array_equal(got[0], ref[0])
            assert np.array_equal(got[1], ref[1])
    finally:
        ctx.close()


def test_noalloc_cull_refuses_past_fixed_scratch(ctx):
    """A call whose scratch exceeds the fixed array is refused with
    HVX_ERR_CAPACITY and launches nothing (outputs untouched)."""
    import torch
    lib = abi.load_library()
    limit = lib.hvx_cf_cull_noalloc_scratch_limit()
    assert limit == 24 << 20
    _, _, ubo, _ = camera(3840, 2160, (32, 18, 48))
    Cn = 32 * 18 * 48
    assert lib.hvx_cf_cull_noalloc_scratch_bytes(Cn, 128, 65536) <= limit
    cap = 256  # 27,648 clusters x 256 x 4 B = 28.3 MB of lists alone
    assert lib.hvx_cf_cull_noalloc_scratch_bytes(Cn, cap, 1024) > limit
    lights = abi.scene_lights(33, 1024)
    d_aabbs = _aabbs(ctx, ubo)
    d_lights = to_dev(lights.view(np.uint8))
    cells = torch.full((Cn * 2,), -7, dtype=torch.int32, device="cuda")
    idx = torch.full((Cn * cap,), -7, dtype=torch.int32, device="cuda")
    with pytest.raises(abi.HvxError) as e:
        ctx.cull_lights_noalloc(ubo, d_aabbs, d_lights, 1024, cap, cells, idx, Cn * cap)
    assert e.value.status == abi.HVX_ERR_CAPACITY
    torch.cuda.synchronize()
    assert int(cells.min()) == -7 and int(cells.max()) == -7
