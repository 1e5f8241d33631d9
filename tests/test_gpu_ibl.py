"""GPU parity of the IBL stages (BRDF LUT, equirect->cube, mip chain, diffuse
irradiance, GGX specular prefilter) against the FP32 CPU oracle through the C
ABI: within 1e-5 relative (the kernels reproduce the oracle's arithmetic, so
they are expected to match it bit for bit), at reduced sizes in full and at
the benchmark sizes on texel subsets."""
import numpy as np
import pytest

from hellovulkan_amd import abi
from oracle import hvx_oracle as O
from tests.helpers import to_dev

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _report(name, got, ref):
    err = O.rel_err(got, ref)
    exact = float(np.mean(np.asarray(got) == np.asarray(ref)))
    print(f"{name}: max rel err {err:.3e}, bit-exact fraction {exact:.5f}")
    return err


@pytest.mark.parametrize("W,H,n,rows", [(64, 64, 1024, None), (512, 512, 1024, 24), (100, 37, 333, None)])
def test_brdf_lut(ctx, W, H, n, rows):
    import torch
    d = torch.empty(W * H * 2, dtype=torch.float32, device="cuda")
    ctx.brdf_lut(W, H, n, d)
    torch.cuda.synchronize()
    got = d.cpu().numpy().reshape(-1, 2)
    if rows:
        rng = np.random.default_rng(W + n)
        ys = np.unique(np.concatenate([[0, H - 1], rng.integers(0, H, rows)]))
        ids = (ys[:, None] * W + np.arange(W)[None, :]).ravel().astype(np.uint32)
    else:
        ids = np.arange(W * H, dtype=np.uint32)
    ref = O.brdf_lut(W, H, n, ids)
    assert _report(f"brdf {W}x{H}x{n}", got[ids], ref) <= TOL


def _equirect(seed, w, h):
    return abi.scene_equirect(seed, w, h)


@pytest.mark.parametrize("w,h,S,sub", [(256, 128, 64, None), (2048, 1024, 1024, 60000)])
def test_equirect_to_cube(ctx, w, h, S, sub):
    import torch
    eq = _equirect(7, w, h)
    cube = torch.empty(6 * S * S * 4, dtype=torch.float32, device="cuda")
    ctx.equirect_to_cube(to_dev(eq), w, h, cube, S)
    torch.cuda.synchronize()
    got = cube.cpu().numpy().reshape(-1, 4)
    ids = np.arange(6 * S * S, dtype=np.uint32)
    if sub:
        ids = np.unique(np.random.default_rng(S).integers(0, 6 * S * S, sub)).astype(np.uint32)
    ref = O.equirect_to_cube(eq, S, ids)
    assert _report(f"equirect {S}", got[ids], ref) <= TOL


def test_cube_mips(ctx):
    import torch
    S, L = 64, 7
    rng = np.random.default_rng(3)
    chain = np.zeros(abi.cube_chain_floats(S, L), np.float32)
    chain[: 6 * S * S * 4] = rng.uniform(0, 100, 6 * S * S * 4)
    d = to_dev(chain)
    ctx.cube_mips(d, S, L)
    torch.cuda.synchronize()
    ref = O.cube_mips(chain, S, L)
    assert np.array_equal(d.cpu().numpy(), ref)


def _cube_from_equirect(ctx, seed, w, h, S):
    import torch
    eq = _equirect(seed, w, h)
    d = torch.empty(6 * S * S * 4, dtype=torch.float32, device="cuda")
    ctx.equirect_to_cube(to_dev(eq), w, h, d, S)
    torch.cuda.synchronize()
    return d


@pytest.mark.parametrize("S,Sd,sub", [(64, 8, None), (1024, 32, 96)])
def test_diffuse(ctx, S, Sd, sub):
    import torch
    cube = _cube_from_equirect(ctx, 9, 2 * S, S, S)
    out = torch.empty(6 * Sd * Sd * 4, dtype=torch.float32, device="cuda")
    ctx.filter(abi.FILTER_DIFFUSE, cube, S, out, Sd, 1, 0)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(-1, 4)
    ids = np.arange(6 * Sd * Sd, dtype=np.uint32)
    if sub:
        ids = np.unique(np.random.default_rng(Sd).integers(0, 6 * Sd * Sd, sub)).astype(np.uint32)
    ref = O.diffuse(cube.cpu().numpy(), S, Sd, ids)
    assert _report(f"diffuse {S}->{Sd}", got[ids], ref) <= TOL


def test_diffuse_staged_and_gathered_paths_equal(ctx):
    """The diffuse filter stages every bilinear footprint of level 0 in
    context scratch (one 64-byte entry each, edge-folded and corner taps
    included); captured into a graph on a context that holds no staging
    buffer yet, it gathers the taps from the cube instead, allocating
    nothing. Both give the same bits: graph replays, with eager work between
    them, equal the staged eager call; a second eager call (staging buffer
    reused) equals the first; sampled texels equal the FP32 oracle."""
    import torch
    S, Sd = 128, 16
    cube = _cube_from_equirect(ctx, 21, 2 * S, S, S)
    ref = torch.empty(6 * Sd * Sd * 4, dtype=torch.float32, device="cuda")
    ctx.filter(abi.FILTER_DIFFUSE, cube, S, ref, Sd, 1, 0)
    torch.cuda.synchronize()
    again = torch.full_like(ref, np.nan)
    ctx.filter(abi.FILTER_DIFFUSE, cube, S, again, Sd, 1, 0)
    torch.cuda.synchronize()
    assert torch.equal(again.view(torch.int32), ref.view(torch.int32))
    fresh = abi.Context(0)
    try:
        out = torch.full_like(ref, np.nan)
        s = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            fresh.filter(abi.FILTER_DIFFUSE, cube, S, out, Sd, 1, 0, stream=s)
        for _ in range(3):
            out.fill_(np.nan)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(out.view(torch.int32), ref.view(torch.int32))
            torch.zeros(1 << 20, device="cuda").add_(1.0)
        del g
    finally:
        fresh.close()
    ids = np.arange(0, 6 * Sd * Sd, 5, dtype=np.uint32)
    got = ref.cpu().numpy().reshape(-1, 4)[ids]
    assert np.array_equal(got.view(np.uint32), O.diffuse(cube.cpu().numpy(), S, Sd, ids).view(np.uint32))


@pytest.mark.parametrize("S,So,mips,n,emulate,sub", [
    (64, 16, 5, 256, False, None),
    (64, 16, 5, 64, True, None),
    (1024, 512, 10, 1024, False, 400),
])
def test_specular(ctx, S, So, mips, n, emulate, sub):
    import torch
    cube = _cube_from_equirect(ctx, 13, 2 * S, S, S)
    out = torch.empty(abi.cube_chain_floats(So, mips), dtype=torch.float32, device="cuda")
    ctx.filter(abi.FILTER_SPECULAR, cube, S, out, So, mips, n, emulate_static_viewport=emulate)
    torch.cuda.synchronize()
    host = out.cpu().numpy()
    c0 = cube.cpu().numpy()
    rng = np.random.default_rng(So + mips)
    worst = 0.0
    for m in range(mips):
        Sm = max(So >> m, 1)
        off = abi.cube_level_offset_floats(So, m)
        got = host[off: off + 6 * Sm * Sm * 4].reshape(-1, 4)
        ids = np.arange(6 * Sm * Sm, dtype=np.uint32)
        if sub and ids.size > sub:
            ids = np.unique(rng.integers(0, ids.size, sub)).astype(np.uint32)
        ref = O.specular(c0, S, So, mips, m, n, ids, emulate=emulate)
        worst = max(worst, _report(f"specular mip {m}", got[ids], ref))
    assert worst <= TOL


def test_filter_rejects_bad_arguments(ctx):
    import torch
    cube = torch.zeros(6 * 16 * 16 * 4, dtype=torch.float32, device="cuda")
    out = torch.zeros(abi.cube_chain_floats(8, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(abi.HvxError) as e:
        ctx.filter(abi.FILTER_DIFFUSE, cube, 16, out, 8, 2, 0)
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    with pytest.raises(abi.HvxError) as e:
        ctx.filter(abi.FILTER_SPECULAR, cube, 16, out, 8, 5, 64)  # 8 has 4 mips
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    with pytest.raises(abi.HvxError) as e:
        ctx.filter(abi.FILTER_SPECULAR, cube, 16, out, 8, 4, 0)
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    with pytest.raises(abi.HvxError) as e:
        ctx.filter(7, cube, 16, out, 8, 1, 1)
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    with pytest.raises(abi.HvxError) as e:
        ctx.cube_mips(cube, 12, 2)  # not a power of two
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
