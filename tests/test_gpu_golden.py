"""The GPU kernels against the committed golden fixtures (tests/golden/,
written by the CPU oracle from seeded inputs): light lists bit for bit, the
plain IBL stages and shading within 1e-5 relative, the contraction filter
within 2e-3 of its FP64 twin, the network's inference within the BF16
tolerance of the FP64 forward, tone mapping bit for bit."""
import hashlib
import os

import numpy as np
import pytest

from hellovulkan_amd import abi
from oracle import hvx_oracle as O
from tests.helpers import SMALL_FRAME, camera, to_dev

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")
FW, FH, FGRID, FNL, FCAP = 64, 40, (4, 4, 8), 256, 32
ENV_SEED, S, SD, SO, MIPS, NSAMP, LUT = 11, 16, 4, 8, 4, 64, 32
CS, CSO, CMIPS = 32, 16, 5


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN)


def test_golden_culling_2560x1600(ctx, gold):
    import torch
    _, _, ubo, _ = camera(SMALL_FRAME["width"], SMALL_FRAME["height"], SMALL_FRAME["grid"])
    Cn = 16 * 9 * 24
    lights = abi.scene_lights(7, SMALL_FRAME["n_lights"])
    aabb = torch.empty(Cn * 8, dtype=torch.float32, device="cuda")
    ctx.build_aabbs(ubo, aabb)
    cells = torch.zeros(Cn * 2, dtype=torch.int32, device="cuda")
    idx = torch.zeros(Cn * SMALL_FRAME["cap"], dtype=torch.int32, device="cuda")
    tot = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.cull_lights(ubo, aabb, to_dev(lights.view(np.uint8)), SMALL_FRAME["n_lights"], SMALL_FRAME["cap"], cells, idx,
                    Cn * SMALL_FRAME["cap"], tot)
    torch.cuda.synchronize()
    t = int(tot.item())
    assert t == int(gold["cull_total"][0])
    assert np.array_equal(cells.cpu().numpy().view(np.uint32), gold["cull_cells"])
    got = idx.cpu().numpy().view(np.uint32)[:t]
    assert np.array_equal(got[:4096], gold["cull_head"])
    assert hashlib.sha256(np.ascontiguousarray(got).tobytes()).hexdigest() == str(gold["cull_digest"])


def _ibl(ctx):
    import torch
    eq = abi.scene_equirect(ENV_SEED, 2 * S, S, threads=4)
    cube = torch.empty(abi.cube_chain_floats(S, abi.mip_count(S)), dtype=torch.float32, device="cuda")
    ctx.equirect_to_cube(to_dev(eq), 2 * S, S, cube, S)
    diff = torch.empty(6 * SD * SD * 4, dtype=torch.float32, device="cuda")
    spec = torch.empty(abi.cube_chain_floats(SO, MIPS), dtype=torch.float32, device="cuda")
    lut = torch.empty(LUT * LUT * 2, dtype=torch.float32, device="cuda")
    ctx.filter(abi.FILTER_DIFFUSE, cube, S, diff, SD, 1, 0)
    ctx.filter(abi.FILTER_SPECULAR, cube, S, spec, SO, MIPS, NSAMP)
    ctx.brdf_lut(LUT, LUT, NSAMP, lut)
    torch.cuda.synchronize()
    return cube, diff, spec, lut


def test_golden_ibl_stages(ctx, gold):
    cube, diff, spec, lut = _ibl(ctx)
    for name, got in (("eq2cube", cube[:6 * S * S * 4]), ("diffuse", diff), ("specular", spec), ("brdf_lut", lut)):
        g = got.cpu().numpy()
        e = O.rel_err(g, gold[name])
        print(f"{name}: max rel err {e:.3e}, bit-identical {np.array_equal(g.view(np.uint32), gold[name].view(np.uint32))}")
        assert e <= 1e-5, name


def test_golden_shading(ctx, gold):
    import torch
    _, diff, spec, lut = _ibl(ctx)
    desc, cam, ubo, ivp = camera(FW, FH, FGRID)
    fl = abi.scene_lights(13, FNL)
    fr = abi.scene_fragments(14, desc, ivp, threads=4)
    Cn = FGRID[0] * FGRID[1] * FGRID[2]
    aabb = torch.empty(Cn * 8, dtype=torch.float32, device="cuda")
    ctx.build_aabbs(ubo, aabb)
    d_l = to_dev(fl.view(np.uint8))
    cells = torch.zeros(Cn * 2, dtype=torch.int32, device="cuda")
    idx = torch.zeros(Cn * FCAP, dtype=torch.int32, device="cuda")
    ctx.cull_lights(ubo, aabb, d_l, FNL, FCAP, cells, idx, Cn * FCAP)
    planes = {k: to_dev(v) for k, v in fr.items()}
    m = abi.ibl_maps(spec, SO, MIPS, diff, SD, lut, LUT, LUT)
    rgb = torch.empty(3 * FW * FH, dtype=torch.float32, device="cuda")
    ctx.shade(abi.frag_soa(planes, FW, FH, ivp), cam, ubo, abi.pbr_defaults(), d_l, FNL, cells, idx, m, rgb)
    torch.cuda.synchronize()
    got = rgb.cpu().numpy().reshape(3, -1).T
    e = O.rel_err(got, gold["shade_rgb"])
    print(f"shading: max rel err {e:.3e}")
    assert e <= 1e-5


def test_golden_contraction(ctx, gold):
    import torch
    eq = abi.scene_equirect(ENV_SEED, 2 * CS, CS, threads=4)
    levels = abi.mip_count(CS)
    cube = torch.empty(abi.cube_chain_floats(CS, levels), dtype=torch.float32, device="cuda")
    ctx.equirect_to_cube(to_dev(eq), 2 * CS, CS, cube, CS)
    ctx.cube_mips(cube, CS, levels)
    plan = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, CS, CSO, CMIPS)
    out = torch.empty(abi.cube_chain_floats(CSO, CMIPS), dtype=torch.float32, device="cuda")
    plan.filter(cube, levels, out)
    torch.cuda.synchronize()
    plan.close()
    o = out.cpu().numpy()
    got = np.concatenate([o[abi.cube_chain_floats(CSO, m):abi.cube_chain_floats(CSO, m + 1)].reshape(-1, 4)[:, :3]
                          for m in range(1, CMIPS)])
    e = O.rel_err(got, gold["contract"])
    print(f"contraction vs FP64 twin: max rel err {e:.3e}")
    assert e <= 2e-3


def test_golden_network_and_tonemap(ctx, gold):
    import torch
    mlp = abi.Mlp(ctx, 1.0, 1)
    mlp.set_params(gold["mlp_params"])
    x = gold["mlp_inputs"]
    rgb = torch.empty(x.shape[0] * 3, dtype=torch.float32, device="cuda")
    mlp.eval(to_dev(x), x.shape[0], rgb)
    torch.cuda.synchronize()
    y = np.log(rgb.cpu().numpy().reshape(-1, 3).astype(np.float64) + 1e-3)
    e = np.abs(y - gold["mlp_logrgb"]).max()
    print(f"network log-radiance vs FP64: max abs err {e:.3e}")
    assert e <= 5e-2
    mlp.close()
    src = to_dev(np.ascontiguousarray(gold["tonemap_in"]).reshape(-1))
    for op in range(4):
        out = torch.empty(3 * 64, dtype=torch.float32, device="cuda")
        ctx.tonemap(src, 64, 1, op, True, out, None)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().reshape(3, 64).view(np.uint32), gold["tonemap_out"][op].view(np.uint32)), op
