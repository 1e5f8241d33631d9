"""GPU tests at the benchmark workloads' full sizes, through the C ABI:

* the batched form: 256 procedural probes with faces drawn from
  {64, 128, 256, 512} (GGX prefilter to S/2, full chain, 1024 samples) and
  16 views of the 3840x2160 / 32x18x48 frame with 1k-64k lights (cap 128),
  batched eagerly and as replays of one captured graph, equal to the
  per-item entry points bit for bit; one 512^2 probe checked against the
  FP32 oracle; the contraction mode equal to single-probe contraction plans;
* the learned radiance network of the 512^2 x 10 GGX chain inferred inside
  3840x2160 / 65,536-light shading: fragments it does not serve are shaded
  bit-identically to hvx_cf_shade;
* the 3840x2160 frame shaded with the IBL maps the library builds from the
  2048x1024 procedural equirect (not small random maps) against the oracle.

Sizes follow the benchmark; the oracle checks sample texels and fragments.
"""
import numpy as np
import pytest

from hellovulkan_amd import abi
from oracle import hvx_oracle as O
from tests.helpers import TARGET, UHD_FRAME, camera

pytestmark = pytest.mark.gpu

N_SAMPLES, CAP, N_VIEWS, N_PROBES = 1024, 128, 16, 256


def _batch_workload(ctx, dev):
    """The 256-probe / 16-view workload (the benchmark's generator, seeded)."""
    import torch
    rng = np.random.default_rng(2024)
    faces = rng.choice([64, 128, 256, 512], N_PROBES).tolist()
    probes = []
    for k, S in enumerate(faces):
        eq = torch.from_numpy(abi.scene_equirect(5000 + k, 2 * S, S, threads=16)).to(dev)
        cube = torch.empty(6 * S * S * 4, dtype=torch.float32, device=dev)
        ctx.equirect_to_cube(eq, 2 * S, S, cube, S)
        So = S // 2
        probes.append(dict(S=S, So=So, mips=abi.mip_count(So), cube=cube))
    W, H, grid = UHD_FRAME["width"], UHD_FRAME["height"], UHD_FRAME["grid"]
    Cn = grid[0] * grid[1] * grid[2]
    counts = np.exp(rng.uniform(np.log(1024), np.log(65536), N_VIEWS)).astype(np.int64).tolist()
    views = []
    for k, n in enumerate(counts):
        ang = 2 * np.pi * k / N_VIEWS
        eye = (TARGET[0] + 6.0 * np.sin(ang), 1.0, TARGET[2] + 6.0 * np.cos(ang))
        _, _, ubo, _ = abi.camera_build(eye, TARGET, 0.785398, 0.1, 100.0, W, H, grid)
        aabb = torch.empty(Cn * 8, dtype=torch.float32, device=dev)
        ctx.build_aabbs(ubo, aabb)
        lights = torch.from_numpy(abi.scene_lights(3000 + k, n).view(np.uint8)).to(dev)
        views.append(dict(ubo=ubo, aabb=aabb, lights=lights, n=n, Cn=Cn))
    torch.cuda.synchronize()
    return probes, views


def _view_bufs(dev, Cn):
    import torch
    return [torch.full((Cn * 2,), -1, dtype=torch.int32, device=dev),
            torch.full((Cn * CAP,), -1, dtype=torch.int32, device=dev),
            torch.full((1,), -1, dtype=torch.int32, device=dev)]


def _same_view(got, ref):
    import torch
    t = int(ref[2].item())
    return int(got[2].item()) == t and torch.equal(got[0], ref[0]) and torch.equal(got[1][:t], ref[1][:t])


def test_batch_256_probes_16_views_equal_per_item(ctx):
    import torch
    dev = torch.device("cuda", 0)
    probes, views = _batch_workload(ctx, dev)
    print(f"batch: {N_PROBES} probes (faces {sorted(set(p['S'] for p in probes))}), "
          f"{N_VIEWS} views ({min(v['n'] for v in views)}-{max(v['n'] for v in views)} lights)", flush=True)
    # per-item entry points
    for p in probes:
        p["ref"] = torch.empty(abi.cube_chain_floats(p["So"], p["mips"]), dtype=torch.float32, device=dev)
        ctx.filter(abi.FILTER_SPECULAR, p["cube"], p["S"], p["ref"], p["So"], p["mips"], N_SAMPLES)
        p["out"] = torch.full_like(p["ref"], -1.0)
    for v in views:
        v["ref"] = _view_bufs(dev, v["Cn"])
        ctx.cull_lights(v["ubo"], v["aabb"], v["lights"], v["n"], CAP, v["ref"][0], v["ref"][1], v["Cn"] * CAP,
                        v["ref"][2])
        v["out"] = _view_bufs(dev, v["Cn"])
    b = abi.Batch(ctx, [abi.probe(p["cube"], p["S"], p["out"], p["So"], p["mips"]) for p in probes], N_SAMPLES,
                  [abi.view(v["ubo"], v["aabb"], v["lights"], v["n"], CAP, v["out"][0], v["out"][1], v["Cn"] * CAP,
                            v["out"][2]) for v in views])
    try:
        def check(what):
            for k, p in enumerate(probes):
                assert torch.equal(p["out"].view(torch.int32), p["ref"].view(torch.int32)), f"{what}: probe {k}"
            for k, v in enumerate(views):
                assert _same_view(v["out"], v["ref"]), f"{what}: view {k}"

        def poison():
            for p in probes:
                p["out"].fill_(-1.0)
            for v in views:
                for t in v["out"]:
                    t.fill_(-1)
            torch.cuda.synchronize()

        b.run()
        torch.cuda.synchronize()
        check("eager run")
        s = torch.cuda.Stream(dev)
        b.capture(s)
        for r in range(3):
            poison()
            b.launch(s)
            s.synchronize()
            check(f"graph replay {r}")
            # eager work between replays
            v = views[r]
            ctx.cull_lights(v["ubo"], v["aabb"], v["lights"], v["n"], CAP, v["ref"][0], v["ref"][1],
                            v["Cn"] * CAP, v["ref"][2])
            torch.cuda.synchronize()
        # one 512^2 probe against the oracle, texels sampled in every mip
        p = next(p for p in probes if p["S"] == 512)
        cube = p["cube"].cpu().numpy()
        got = p["out"].cpu().numpy()
        rng = np.random.default_rng(8)
        for m in range(p["mips"]):
            Sm = max(p["So"] >> m, 1)
            ids = np.unique(rng.integers(0, 6 * Sm * Sm, 64)).astype(np.uint32)
            ref = O.specular(cube, 512, p["So"], p["mips"], m, N_SAMPLES, ids)
            off = abi.cube_level_offset_floats(p["So"], m)
            g = got[off: off + 6 * Sm * Sm * 4].reshape(-1, 4)[ids]
            assert np.array_equal(g.view(np.uint32), ref.view(np.uint32)), f"512^2 probe, mip {m}"
        # contraction mode equals single-probe contraction plans
        b.set_contract(True, abi.PRECISION_FP16_SCALED)
        poison()
        b.run()
        torch.cuda.synchronize()
        plans, chains = {}, {}
        for k, p in enumerate(probes):
            key = (p["S"], p["So"], p["mips"])
            if key not in plans:
                plans[key] = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, *key, abi.PRECISION_FP16_SCALED)
                chains[key] = torch.empty(abi.cube_chain_floats(p["S"], abi.mip_count(p["S"])), dtype=torch.float32,
                                          device=dev)
            chain = chains[key]
            chain[: p["cube"].numel()].copy_(p["cube"])
            ctx.cube_mips(chain, p["S"], abi.mip_count(p["S"]))
            single = torch.full_like(p["ref"], -3.0)
            plans[key].filter(chain, abi.mip_count(p["S"]), single)
            torch.cuda.synchronize()
            assert torch.equal(p["out"].view(torch.int32), single.view(torch.int32)), f"contraction mode: probe {k}"
        for v in views:
            assert _same_view(v["out"], v["ref"])
        for pl in plans.values():
            pl.close()
    finally:
        b.close()


def _cfg2_maps(ctx, dev):
    """The IBL maps the library builds from the 2048x1024 procedural
    equirect: 6x1024^2 cube, 6x32^2 irradiance, 6x512^2 x 10 GGX chain,
    512^2 BRDF LUT (1024 samples each)."""
    import torch
    eq = torch.from_numpy(abi.scene_equirect(77, 2048, 1024, threads=16)).to(dev)
    cube = torch.empty(abi.cube_chain_floats(1024, abi.mip_count(1024)), dtype=torch.float32, device=dev)
    ctx.equirect_to_cube(eq, 2048, 1024, cube, 1024)
    spec = torch.empty(abi.cube_chain_floats(512, 10), dtype=torch.float32, device=dev)
    diff = torch.empty(6 * 32 * 32 * 4, dtype=torch.float32, device=dev)
    lut = torch.empty(512 * 512 * 2, dtype=torch.float32, device=dev)
    ctx.filter(abi.FILTER_DIFFUSE, cube, 1024, diff, 32, 1, 0)
    ctx.filter(abi.FILTER_SPECULAR, cube, 1024, spec, 512, 10, N_SAMPLES)
    ctx.brdf_lut(512, 512, N_SAMPLES, lut)
    torch.cuda.synchronize()
    return spec, diff, lut


def _frame(ctx, dev):
    import torch
    W, H, grid, NL = UHD_FRAME["width"], UHD_FRAME["height"], UHD_FRAME["grid"], UHD_FRAME["n_lights"]
    desc, cam, ubo, ivp = camera(W, H, grid)
    Cn = grid[0] * grid[1] * grid[2]
    lights = abi.scene_lights(1001, NL)
    fr = abi.scene_fragments(1002, desc, ivp, threads=16)
    d_lights = torch.from_numpy(lights.view(np.uint8)).to(dev)
    d_fr = {k: torch.from_numpy(v).to(dev) for k, v in fr.items()}
    aabb = torch.empty(Cn * 8, dtype=torch.float32, device=dev)
    ctx.build_aabbs(ubo, aabb)
    cells = torch.empty(Cn * 2, dtype=torch.int32, device=dev)
    idx = torch.empty(Cn * CAP, dtype=torch.int32, device=dev)
    tot = torch.zeros(1, dtype=torch.int32, device=dev)
    ctx.cull_lights(ubo, aabb, d_lights, NL, CAP, cells, idx, Cn * CAP, tot)
    torch.cuda.synchronize()
    return dict(W=W, H=H, cam=cam, ubo=ubo, ivp=ivp, lights=lights, fr=fr, d_lights=d_lights, NL=NL,
                frs=abi.frag_soa(d_fr, W, H, ivp), d_fr=d_fr, cells=cells, idx=idx, tot=tot)


def test_shade_3840x2160_with_library_ibl_maps_vs_oracle(ctx):
    """The frame shaded with the 512^2 x 10 GGX chain, 32^2 irradiance and
    512^2 LUT the library built: 30,000 sampled fragments within 1e-5 of the
    FP32 oracle shading them with the same maps."""
    import torch
    dev = torch.device("cuda", 0)
    spec, diff, lut = _cfg2_maps(ctx, dev)
    f = _frame(ctx, dev)
    maps = abi.ibl_maps(spec, 512, 10, diff, 32, lut, 512, 512)
    rgb = torch.empty(3 * f["W"] * f["H"], dtype=torch.float32, device=dev)
    pc = abi.pbr_defaults()
    ctx.shade(f["frs"], f["cam"], f["ubo"], pc, f["d_lights"], f["NL"], f["cells"], f["idx"], maps, rgb)
    torch.cuda.synchronize()
    tot = int(f["tot"].item())
    host_maps = dict(spec=spec.cpu().numpy(), spec_size=512, spec_mips=10, diff=diff.cpu().numpy(), diff_size=32,
                     lut=lut.cpu().numpy(), lut_w=512, lut_h=512)
    ids = np.random.default_rng(12).choice(f["W"] * f["H"], 30000, replace=False).astype(np.uint32)
    cells = f["cells"].cpu().numpy().view(abi.CELL_DTYPE)
    idx = f["idx"].cpu().numpy().view(np.uint32)[:tot]
    ref = O.shade(f["fr"], f["W"], f["H"], f["ivp"], f["cam"], f["ubo"], pc, f["lights"], cells, idx, host_maps, ids,
                  threads=16)
    got = rgb.cpu().numpy().reshape(3, -1).T[ids]
    err = O.rel_err(got, ref)
    print(f"shading with the library's IBL maps: max rel err {err:.2e} over {ids.size} fragments")
    assert err <= 1e-5


def test_learned_shading_3840x2160_unserved_fragments_unchanged(ctx):
    """The network trained on the 512^2 x 10 chain (4M samples, 15,000
    steps) inferred inside 3840x2160 / 65,536-light shading with the 2e-3
    fallback. In the exact arithmetic mode (whose results do not depend on
    how fragments are grouped into waves) every fragment the network does
    not serve equals hvx_cf_shade's bit for bit, and only served fragments
    change. Prints the served fraction and the fraction of fragments whose
    radiance moved by more than 2e-3."""
    import torch
    dev = torch.device("cuda", 0)
    spec, diff, lut = _cfg2_maps(ctx, dev)
    f = _frame(ctx, dev)
    maps = abi.ibl_maps(spec, 512, 10, diff, 32, lut, 512, 512)
    pc = abi.pbr_defaults()
    mlp = abi.Mlp(ctx, float(pc.max_reflection_lod), 2024)
    try:
        n, batch, steps = 1 << 22, 32768, 15000
        x = torch.empty(n * 4, dtype=torch.float32, device=dev)
        t = torch.empty(n * 4, dtype=torch.float32, device=dev)
        mlp.sample(spec, 512, 10, 99, n, x, t)
        mlp.train(x, t, n, batch, steps, 4e-3)
        cf, bins, tol = 128, 4, 2e-3
        cell_err = torch.empty(bins * 6 * cf * cf, dtype=torch.float32, device=dev)
        mlp.cell_errors(spec, 512, 10, cf, bins, cell_err)
        F = f["W"] * f["H"]
        ctx.set_option(abi.OPT_SHADE_EXACT, 1)
        try:
            learned = torch.empty(3 * F, dtype=torch.float32, device=dev)
            plain = torch.empty(3 * F, dtype=torch.float32, device=dev)
            served = torch.zeros(1, dtype=torch.int32, device=dev)
            ctx.shade_learned(f["frs"], f["cam"], f["ubo"], pc, f["d_lights"], f["NL"], f["cells"], f["idx"], maps,
                              mlp, cell_err, cf, bins, tol, learned, served)
            ctx.shade(f["frs"], f["cam"], f["ubo"], pc, f["d_lights"], f["NL"], f["cells"], f["idx"], maps, plain)
            torch.cuda.synchronize()
        finally:
            ctx.set_option(abi.OPT_SHADE_EXACT, 0)
        n_served = int(served.item())
        changed = (learned.view(3, -1).view(torch.int32) != plain.view(3, -1).view(torch.int32)).any(dim=0)
        n_changed = int(changed.sum().item())
        rel = ((learned - plain).abs() / plain.abs().clamp_min(1e-30)).view(3, -1).amax(dim=0)
        beyond = float((rel > tol).float().mean().item())
        print(f"learned shading: served {n_served / F:.4f} of fragments, changed {n_changed / F:.4f}, "
              f"max rel diff {float(rel.max().item()):.2e}, fraction beyond 2e-3 {beyond:.2e}")
        assert 0 < n_served < F
        assert n_changed <= n_served  # nothing the network did not serve changed
        assert n_changed >= 0.9 * n_served
        # the margin on the measured cell errors keeps every served fragment
        # within the bound on this frame
        assert int((rel > tol).sum().item()) == 0
    finally:
        mlp.close()
