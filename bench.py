#!/usr/bin/env python3
"""Benchmark of the clustered-forward + IBL compute path on one MI355X.

One step = one frame of the 3840x2160 workload: light culling of 65,536
lights into the 32x18x48 cluster grid (cap 128) followed by clustered
Cook-Torrance + split-sum IBL shading of 8,294,400 struct-of-arrays fragments
(the per-frame stages; cluster AABBs are rebuilt only when the camera's
projection changes, as in the reference, and are timed separately). Inputs
are synthetic and seeded, resident in HBM before timing starts. The IBL maps
the shading samples are produced by this library's own stages from the
2048x1024 procedural equirect (cube 6x1024^2, irradiance 6x32^2, 10-mip GGX
prefilter from 512^2, BRDF LUT 512^2 x 1024 samples); those stages are timed
once each and reported beside the frame metric.

Prints one JSON line. `value` is fragments shaded per second (Mfrag/s) over
whole frames (cull + shade).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cluster-light tests/s + Mfrag/s shaded; IBL filter ms vs plain FP32 @<=2e-3 err"
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32 vector peak (datasheet)
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E peak (datasheet)
# algorithmic work per fragment of the shading kernel: 160 FP32 ops of per-
# fragment set-up + ambient, 140 per listed light; 45 bytes of fragment data
# in and out (depth 4, normal 12, albedo 12, roughness 4, metallic 1, RGB 12)
OPS_FRAG, OPS_PAIR, BYTES_FRAG = 160, 140, 45
# texels the contraction filters' bright-texel split may sum exactly
HOT_TEXELS = 2048


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-ibl", action="store_true", help="skip the one-off IBL stage timings")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample duration")
    p.add_argument("--no-batch", action="store_true", help="skip the batched probes + views workload")
    p.add_argument("--no-contract", action="store_true", help="skip the contraction (MFMA) filter timings")
    p.add_argument("--no-learned", action="store_true", help="skip the learned-radiance (MLP) workload")
    p.add_argument("--mlp-steps", type=int, default=15000)
    p.add_argument("--batch-probes", type=int, default=256)
    p.add_argument("--calibrate", action="store_true",
                   help="run one 256 MiB elementwise copy first (known HBM traffic, for counter scaling)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this path runs on a single MI355X. "
              "Re-run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import torch
    from hellovulkan_amd import abi
    from tests.helpers import UHD_FRAME, camera

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    ctx = abi.Context(0)
    stream = torch.cuda.current_stream(dev)

    W, H, grid, NL, CAP = UHD_FRAME["width"], UHD_FRAME["height"], UHD_FRAME["grid"], UHD_FRAME["n_lights"], UHD_FRAME["cap"]
    F = W * H
    Cn = grid[0] * grid[1] * grid[2]
    desc, cam, ubo, ivp = camera(W, H, grid)

    # ---------------- IBL maps (2048x1024 equirect, 1024^2 cube, 512^2 x 10 chain) built by the library's own stages
    t0 = time.time()
    EQW, EQH, CUBE, DIFF, SPEC, SPEC_MIPS, LUT, LUT_N, SPEC_N = 2048, 1024, 1024, 32, 512, 10, 512, 1024, 1024
    eq = abi.scene_equirect(77, EQW, EQH, threads=16)
    d_eq = torch.from_numpy(eq).to(dev)
    d_cube = torch.empty(abi.cube_chain_floats(CUBE, abi.mip_count(CUBE)), dtype=torch.float32, device=dev)
    d_diff = torch.empty(6 * DIFF * DIFF * 4, dtype=torch.float32, device=dev)
    d_spec = torch.empty(abi.cube_chain_floats(SPEC, SPEC_MIPS), dtype=torch.float32, device=dev)
    d_lut = torch.empty(LUT * LUT * 2, dtype=torch.float32, device=dev)

    def timed(fn, reps=1):
        fn()  # warm (also allocates scratch)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    if args.calibrate:  # y = x + 0: reads and writes 256 MiB each
        cx = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        cy = torch.empty_like(cx)
        torch.add(cx, 0.0, out=cy)
        torch.cuda.synchronize()
        del cx, cy

    ibl = {}
    ibl["equirect_to_cube_ms"] = timed(lambda: ctx.equirect_to_cube(d_eq, EQW, EQH, d_cube, CUBE))
    ibl["cube_mips_ms"] = timed(lambda: ctx.cube_mips(d_cube, CUBE, abi.mip_count(CUBE)))
    if args.no_ibl:
        ctx.filter(abi.FILTER_DIFFUSE, d_cube, CUBE, d_diff, DIFF, 1, 0)
        ctx.filter(abi.FILTER_SPECULAR, d_cube, CUBE, d_spec, SPEC, SPEC_MIPS, SPEC_N)
        ctx.brdf_lut(LUT, LUT, LUT_N, d_lut)
        torch.cuda.synchronize()
    else:
        ibl["diffuse_plain_fp32_ms"] = timed(lambda: ctx.filter(abi.FILTER_DIFFUSE, d_cube, CUBE, d_diff, DIFF, 1, 0))
        ibl["specular_plain_fp32_ms"] = timed(
            lambda: ctx.filter(abi.FILTER_SPECULAR, d_cube, CUBE, d_spec, SPEC, SPEC_MIPS, SPEC_N))
        ibl["brdf_lut_ms"] = timed(lambda: ctx.brdf_lut(LUT, LUT, LUT_N, d_lut))
    log(f"IBL maps ready in {time.time() - t0:.1f}s: {ibl}")

    # ---------------- frame inputs
    lights = abi.scene_lights(1001, NL)
    fr = abi.scene_fragments(1002, desc, ivp, threads=16)
    d_lights = torch.from_numpy(lights.view(np.uint8)).to(dev)
    d_fr = {k: torch.from_numpy(v).to(dev) for k, v in fr.items()}
    frs = abi.frag_soa(d_fr, W, H, ivp)
    maps = abi.ibl_maps(d_spec, SPEC, SPEC_MIPS, d_diff, DIFF, d_lut, LUT, LUT)
    pc = abi.pbr_defaults()
    d_aabb = torch.empty(Cn * 8, dtype=torch.float32, device=dev)
    d_cells = torch.empty(Cn * 2, dtype=torch.int32, device=dev)
    d_idx = torch.empty(Cn * CAP, dtype=torch.int32, device=dev)
    d_tot = torch.zeros(1, dtype=torch.int32, device=dev)
    d_rgb = torch.empty(3 * F, dtype=torch.float32, device=dev)
    aabb_ms = timed(lambda: ctx.build_aabbs(ubo, d_aabb), reps=10)

    def cull():
        ctx.cull_lights(ubo, d_aabb, d_lights, NL, CAP, d_cells, d_idx, Cn * CAP, d_tot)

    def shade():
        ctx.shade(frs, cam, ubo, pc, d_lights, NL, d_cells, d_idx, maps, d_rgb)

    # ---------------- warmup + timed frames, per-stage HIP events on the launch stream
    for _ in range(args.warmup):
        cull()
        shade()
    torch.cuda.synchronize()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    # HIP events around the shading kernel itself (inside the shading stage's
    # five launches, on the same stream), for the roofline of that kernel
    ctx.set_option(abi.OPT_SHADE_KERNEL_EVENTS, min(args.steps, 256))
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        cull()
        ev[k][1].record(stream)
        shade()
        ev[k][2].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    cull_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in ev]))
    shade_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in ev]))
    shade_kernel_ms = float(np.mean(ctx.kernel_times()))
    ctx.set_option(abi.OPT_SHADE_KERNEL_EVENTS, 0)
    ms_per_step = wall * 1e3 / args.steps

    # ---------------- algorithmic work of the dominant (shading) kernel
    cells = d_cells.cpu().numpy().view(np.uint32).reshape(-1, 2)
    counts = cells[:, 1].astype(np.int64)
    z = fr["depth"].astype(np.float64)
    lin = (0.1 * 100.0) / (100.0 + z * (0.1 - 100.0))
    zi = np.clip(np.floor(np.maximum(np.log2(lin) * ubo.slice_scaling + ubo.slice_bias, 0)), 0, grid[2] - 1)
    px = np.arange(F) % W
    py = np.arange(F) // W
    cx = ((px + 0.5) / (W / grid[0])).astype(np.int64)
    cy = ((py + 0.5) / (H / grid[1])).astype(np.int64)
    lf = counts[cx + cy * grid[0] + zi.astype(np.int64) * grid[0] * grid[1]]
    pairs = int(lf.sum())
    ops = OPS_FRAG * F + OPS_PAIR * pairs
    achieved_tflops = ops / (shade_kernel_ms * 1e-3) / 1e12
    bytes_alg = BYTES_FRAG * F
    roofline = {
        "kernel": "shade_kernel",
        "bound": "valu",
        "achieved": round(achieved_tflops, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 4),
        "traffic": None,
        "ops_per_launch": ops,
        "fragment_light_pairs": pairs,
        "kernel_ms": round(shade_kernel_ms, 4),
        "hbm_term_GBs": round(bytes_alg / (shade_kernel_ms * 1e-3) / 1e9, 1),
        "hbm_term_frac": round(bytes_alg / (shade_kernel_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
    }
    # HBM traffic per launch of the shading kernel, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload (calibrated)
    tpath = os.path.join(ROOT, "profiles", "shade_kernel_3840x2160_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            roofline["traffic"] = int(json.load(fh)["traffic_bytes"])
    tests = Cn * NL
    result = {
        "metric": METRIC,
        "value": round(F / (ms_per_step * 1e-3) / 1e6, 2),
        "unit": "Mfrag/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "clustered-forward frame 3840x2160, 32x18x48 clusters, 65536 lights, cap 128, "
                               "8294400 SoA fragments; IBL maps from a 2048x1024 procedural equirect",
                   "width": W, "height": H, "clusters": list(grid), "lights": NL, "cap": CAP, "fragments": F},
        "cull_ms": round(cull_ms, 4),
        "shade_ms": round(shade_ms, 4),
        "aabb_ms": round(aabb_ms, 4),
        "cluster_light_tests_per_s": round(tests / (cull_ms * 1e-3), 1),
        "list_entries": int(d_tot.item()),
        "ibl_stage_ms": {k: round(v, 4) for k, v in ibl.items()},
        "roofline": roofline,
    }

    contract_out = {}
    if not args.no_contract and not args.no_ibl:
        result["ibl_contract"] = contract_bench(ctx, stream, d_cube, CUBE, SPEC, SPEC_MIPS, DIFF,
                                                ibl["specular_plain_fp32_ms"], ibl["diffuse_plain_fp32_ms"],
                                                contract_out)
    if not args.no_learned:
        result["learned"] = learned_bench(ctx, stream, args.mlp_steps, d_spec, SPEC, SPEC_MIPS, frs, cam, ubo, pc,
                                          d_lights, NL, d_cells, d_idx, maps, F, shade_ms)
    if not args.no_batch:
        result["batch"] = batch_bench(ctx, dev, stream, args.batch_probes)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args, ubo, cam, desc, ivp, lights, fr, d_cube, CUBE, d_spec, d_diff,
                                              d_lut, F, pc, contract_out)
    print(json.dumps(result), flush=True)
    ctx.close()


def _ev_ms(fn, stream, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def contract_bench(ctx, stream, d_cube, S, So, mips, Sd, plain_spec_ms, plain_diff_ms, outputs):
    """The dense-contraction (FP16 MFMA) filters on the same 6x1024^2 cube as
    the plain filters: plan (weight) construction once, then per-call time.
    Their outputs are kept in `outputs` for the error report of the CPU leg."""
    import torch
    from hellovulkan_amd import abi
    levels = abi.mip_count(S)
    out = {}
    for name, ftype, so, m, plain, prec, hot in [
            ("specular", abi.FILTER_SPECULAR, So, mips, plain_spec_ms, abi.PRECISION_FP16_SCALED, 0),
            ("specular_split_bf16", abi.FILTER_SPECULAR, So, mips, plain_spec_ms, abi.PRECISION_BF16_SPLIT, 0),
            ("specular_hot", abi.FILTER_SPECULAR, So, mips, plain_spec_ms, abi.PRECISION_FP16_SCALED, HOT_TEXELS),
            ("diffuse", abi.FILTER_DIFFUSE, Sd, 1, plain_diff_ms, abi.PRECISION_FP16_SCALED, 0),
            ("diffuse_hot", abi.FILTER_DIFFUSE, Sd, 1, plain_diff_ms, abi.PRECISION_FP16_SCALED, HOT_TEXELS)]:
        t0 = time.time()
        plan = abi.ContractPlan(ctx, ftype, S, so, m, prec)
        if hot:
            plan.set_hot(hot)
        torch.cuda.synchronize()
        build_s = time.time() - t0
        dst = torch.empty(abi.cube_chain_floats(so, m), dtype=torch.float32, device=d_cube.device)
        ms = _ev_ms(lambda: plan.filter(d_cube, levels, dst), stream, 10)
        outputs[name] = dst
        info = [plan.info(k) for k in range(m)]
        wbytes = sum(i.weight_bytes for i in info)
        out[name] = {"ms": round(ms, 4), "plain_fp32_ms": round(plain, 4), "speedup": round(plain / ms, 2),
                     "plan_build_s": round(build_s, 3),
                     "weight_MB": round(wbytes / 1e6, 1),
                     "weight_stream_TBs": round(wbytes / (ms * 1e-3) / 1e12, 3),
                     "mips": [{"kind": ["delta", "full", "near+far"][i.kind], "in": i.in_face, "far": i.far_face,
                               "weight_MB": round(i.weight_bytes / 1e6, 1)} for i in info]}
        if hot:
            T, n_hot = plan.hot_info()
            out[name].update(bright_texels_max=hot, threshold=T, bright_texels=n_hot)
        plan.close()
    log(f"contraction filters: {out['specular']['ms']} ms specular ({out['specular']['speedup']}x), "
        f"{out['specular_hot']['ms']} ms with the bright-texel split ({out['specular_hot']['speedup']}x), "
        f"{out['diffuse']['ms']} ms diffuse ({out['diffuse']['speedup']}x)")
    return out


def learned_bench(ctx, stream, steps, d_spec, SPEC, SPEC_MIPS, frs, cam, ubo, pc, d_lights, NL, d_cells, d_idx, maps,
                  F, plain_shade_ms):
    """The per-environment network: 4M samples of the prefiltered chain over
    the lod range the shading reads (roughness * maxReflectionLod), BF16 MFMA
    training, the 2e-3 fallback mask, then the frame shaded with the network
    inferred inside the shading kernel."""
    import torch
    from hellovulkan_amd import abi
    lod_max = float(pc.max_reflection_lod)
    n, batch = 1 << 22, 32768
    mlp = abi.Mlp(ctx, lod_max, 2024)
    x = torch.empty(n * 4, dtype=torch.float32, device=d_spec.device)
    t = torch.empty(n * 4, dtype=torch.float32, device=d_spec.device)
    mlp.sample(d_spec, SPEC, SPEC_MIPS, 99, n, x, t)
    curve = torch.zeros(steps, dtype=torch.float32, device=d_spec.device)
    torch.cuda.synchronize()
    t0 = time.time()
    mlp.train(x, t, n, batch, steps, 4e-3, curve)
    torch.cuda.synchronize()
    train_s = time.time() - t0
    c = curve.cpu().numpy()
    # held-out samples
    nh = 1 << 18
    xh = torch.empty(nh * 4, dtype=torch.float32, device=d_spec.device)
    th = torch.empty(nh * 4, dtype=torch.float32, device=d_spec.device)
    mlp.sample(d_spec, SPEC, SPEC_MIPS, 100, nh, xh, th)
    rgb = torch.empty(nh * 3, dtype=torch.float32, device=d_spec.device)
    mlp.eval(xh, nh, rgb)
    rgb_n = torch.empty(n * 3, dtype=torch.float32, device=d_spec.device)
    infer_ms = _ev_ms(lambda: mlp.eval(x, n, rgb_n), stream, 5)
    del rgb_n
    ref = torch.exp(th.view(-1, 4)[:, :3]) - 1e-3

    def rel_q(pred):
        rel = ((pred - ref).abs() / ref.abs().clamp_min(1e-6)).amax(dim=1)
        return rel, torch.quantile(rel, torch.tensor([0.5, 0.9, 0.99], device=rel.device)).cpu().numpy()

    rel, relq = rel_q(rgb.view(-1, 3))
    # the classical representation at comparable storage: mips 0-1 of the
    # chain box-downsampled, sampled at the same held-out points
    S0 = 6 * SPEC * SPEC * 4
    L0 = d_spec[:S0].view(6, SPEC, SPEC, 4)
    L1 = d_spec[S0:S0 + 6 * (SPEC // 2) ** 2 * 4].view(6, SPEC // 2, SPEC // 2, 4)

    def down(L, s):
        f = L.shape[1] // s
        return L.reshape(6, s, f, s, f, 4).mean(dim=(2, 4)).reshape(-1)

    eq_maps = {}
    for Se in (16, 32, 64):
        small = torch.cat([down(L0, Se), down(L1, Se // 2)]).contiguous()
        xs, ts = torch.empty_like(xh), torch.empty_like(th)
        mlp.sample(small, Se, 2, 100, nh, xs, ts)
        q = rel_q(torch.exp(ts.view(-1, 4)[:, :3]) - 1e-3)[1]
        eq_maps[f"{Se}^2"] = {"bytes": small.numel() * 4, "rel_err_median": float(q[0]), "rel_err_p90": float(q[1]),
                              "rel_err_p99": float(q[2])}
    # measured cell errors, the learned frame and its difference from plain shading
    cf, bins, tol = 128, 4, 2e-3
    cell_err = torch.empty(bins * 6 * cf * cf, dtype=torch.float32, device=d_spec.device)
    mlp.cell_errors(d_spec, SPEC, SPEC_MIPS, cf, bins, cell_err)
    out = torch.empty(3 * F, dtype=torch.float32, device=d_spec.device)
    plain = torch.empty(3 * F, dtype=torch.float32, device=d_spec.device)
    served = torch.zeros(1, dtype=torch.int32, device=d_spec.device)
    ctx.shade_learned(frs, cam, ubo, pc, d_lights, NL, d_cells, d_idx, maps, mlp, cell_err, cf, bins, tol, out, served)
    ctx.shade(frs, cam, ubo, pc, d_lights, NL, d_cells, d_idx, maps, plain)
    torch.cuda.synchronize()
    served_n = int(served.item())
    diff = ((out - plain).abs() / plain.abs().clamp_min(1e-30)).view(3, -1).amax(dim=0)
    learned_ms = _ev_ms(lambda: ctx.shade_learned(frs, cam, ubo, pc, d_lights, NL, d_cells, d_idx, maps, mlp,
                                                  cell_err, cf, bins, tol, out), stream, 5)
    res = {
        "workload": f"network per environment over lod [0, {lod_max}] of the 6x512^2 x10 GGX chain; "
                    f"{n} samples, batch {batch}, {steps} Adam steps; fallback where the network could move the "
                    f"shaded radiance by more than 2e-3 relative",
        "train_s": round(train_s, 3),
        "train_samples_per_s": round(steps * batch / train_s, 1),
        "loss_first": float(c[0]), "loss_last": float(c[-50:].mean()),
        "heldout_rel_err_median": float(relq[0]), "heldout_rel_err_p90": float(relq[1]),
        "heldout_rel_err_p99": float(relq[2]),
        "downsampled_map_rel_err": eq_maps,
        "heldout_within_2e-3": round(float((rel <= 2e-3).float().mean().item()), 4),
        "cell_err_median": float(cell_err.median().item()),
        "served_fraction": round(served_n / F, 4),
        "frame_max_rel_diff_vs_plain": float(diff.max().item()),
        "frame_fraction_beyond_2e-3": float((diff > tol).float().mean().item()),
        "infer_Mquery_per_s": round(n / (infer_ms * 1e-3) / 1e6, 1),
        "shade_learned_ms": round(learned_ms, 4), "shade_plain_ms": round(plain_shade_ms, 4),
        "network_bytes": abi.MLP_PARAMS * 4, "map_bytes": abi.cube_chain_floats(SPEC, SPEC_MIPS) * 4,
    }
    log(f"learned: loss {res['loss_first']:.3f} -> {res['loss_last']:.4f} in {train_s:.1f}s; served "
        f"{res['served_fraction']:.3f}; learned shading {learned_ms:.3f} ms vs plain {plain_shade_ms:.3f} ms")
    mlp.close()
    return res


def batch_bench(ctx, dev, stream, n_probes):
    """The batched workload: n_probes procedural environment probes with cube
    faces drawn uniformly from {64,128,256,512}, each GGX-prefiltered
    (output face S/2, full mip chain, 1024 samples), and 16 views of the
    3840x2160 / 32x18x48 frame around the scene with light counts
    log-uniform in [1024, 65536], culled with cap 128. Timed three ways on
    the same outputs: per-item entry-point launches, one eager batched run,
    and replays of the captured batch graph (both branches in one graph)."""
    import torch
    from hellovulkan_amd import abi
    from tests.helpers import TARGET
    rng = np.random.default_rng(2024)
    N_SAMPLES, CAP, NV = 1024, 128, 16
    faces = rng.choice([64, 128, 256, 512], n_probes)
    t0 = time.time()
    probes, batch_out, per_item_out, cubes, specs = [], [], [], [], []
    for k, S in enumerate(faces.tolist()):
        eq = torch.from_numpy(abi.scene_equirect(5000 + k, 2 * S, S, threads=16)).to(dev)
        cube = torch.empty(6 * S * S * 4, dtype=torch.float32, device=dev)
        ctx.equirect_to_cube(eq, 2 * S, S, cube, S)
        So = S // 2
        mips = abi.mip_count(So)
        out = torch.empty(abi.cube_chain_floats(So, mips), dtype=torch.float32, device=dev)
        ref = torch.empty_like(out)
        cubes.append(cube)
        per_item_out.append(ref)
        specs.append((S, So, mips))
        batch_out.append(out)
        probes.append(abi.probe(cube, S, out, So, mips))
    W, H, grid = 3840, 2160, (32, 18, 48)
    Cn = grid[0] * grid[1] * grid[2]
    counts = np.exp(rng.uniform(np.log(1024), np.log(65536), NV)).astype(np.int64).tolist()
    views, view_items, keep = [], [], []
    for k, n in enumerate(counts):
        ang = 2 * np.pi * k / NV
        eye = (TARGET[0] + 6.0 * np.sin(ang), 1.0, TARGET[2] + 6.0 * np.cos(ang))
        desc, cam, ubo, ivp = abi.camera_build(eye, TARGET, 0.785398, 0.1, 100.0, W, H, grid)
        aabb = torch.empty(Cn * 8, dtype=torch.float32, device=dev)
        ctx.build_aabbs(ubo, aabb)
        lights = torch.from_numpy(abi.scene_lights(3000 + k, n).view(np.uint8)).to(dev)
        bufs = [torch.empty(Cn * 2, dtype=torch.int32, device=dev), torch.empty(Cn * CAP, dtype=torch.int32, device=dev),
                torch.zeros(1, dtype=torch.int32, device=dev)]
        ref = [torch.empty_like(b) for b in bufs]
        keep += [aabb, lights, *bufs, *ref]
        views.append(abi.view(ubo, aabb, lights, n, CAP, bufs[0], bufs[1], Cn * CAP, bufs[2]))
        view_items.append((ubo, aabb, lights, n, bufs, ref))
    torch.cuda.synchronize()
    log(f"batch inputs ready in {time.time() - t0:.1f}s ({n_probes} probes, {NV} views)")

    def per_item():
        for cube, ref, (S, So, mips) in zip(cubes, per_item_out, specs):
            ctx.filter(abi.FILTER_SPECULAR, cube, S, ref, So, mips, N_SAMPLES)
        for ubo, aabb, lights, n, bufs, ref in view_items:
            ctx.cull_lights(ubo, aabb, lights, n, CAP, ref[0], ref[1], Cn * CAP, ref[2])

    def ev_time(fn, s, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        torch.cuda.synchronize()
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    per_item_ms = ev_time(per_item, stream, 2)
    b = abi.Batch(ctx, probes, N_SAMPLES, views)
    t0 = time.time()
    tun = b.autotune(stream)
    tune_s = time.time() - t0
    log(f"autotuned in {tune_s:.1f}s: tile {tun.probe_tile} x {tun.probe_wgs_per_cu}/CU, "
        f"cull shape {tun.cull_shape} x {tun.cull_wgs_per_cu}/CU")
    eager_ms = ev_time(lambda: b.run(stream), stream, 2)
    gs = torch.cuda.Stream(dev)
    b.capture(gs)
    graph_ms = ev_time(lambda: b.launch(gs), gs, 3)
    same = all(torch.equal(o.view(torch.int32), r.view(torch.int32)) for o, r in zip(batch_out, per_item_out))
    for *_, bufs, ref in view_items:
        t = int(ref[2].item())
        same = same and int(bufs[2].item()) == t and torch.equal(bufs[0], ref[0]) and torch.equal(bufs[1][:t], ref[1][:t])
    texels = sum(6 * sum(max(So >> m, 1) ** 2 for m in range(mips)) for _, So, mips in specs)
    # the same batch with the probes through the FP16 MFMA contraction filter
    # (one plan per probe size; its error against FP64 is the contraction tests')
    t0 = time.time()
    b.set_contract(True)
    plans_s = time.time() - t0
    c_eager_ms = ev_time(lambda: b.run(stream), stream, 3)
    b.close()
    log(f"contraction-mode batch: {c_eager_ms:.2f} ms per eager run (plans built in {plans_s:.1f}s)")
    return {
        "workload": f"{n_probes} probes (faces {{64,128,256,512}} -> GGX prefilter S/2, full chain, 1024 samples) "
                    f"+ {NV} views 3840x2160 32x18x48 (lights log-uniform [1024,65536], cap 128)",
        "probe_faces": {str(f): int((faces == f).sum()) for f in (64, 128, 256, 512)},
        "view_lights": counts,
        "output_texels": texels,
        "per_item_ms": round(per_item_ms, 3),
        "batched_eager_ms": round(eager_ms, 3),
        "batched_graph_ms": round(graph_ms, 3),
        "speedup_graph_vs_per_item": round(per_item_ms / graph_ms, 3),
        "tuning": {"probe_tile": tun.probe_tile, "probe_wgs_per_cu": tun.probe_wgs_per_cu,
                   "cull_shape": ["4x4x1", "8x4x1", "4x4x2", "2x2x4"][tun.cull_shape],
                   "cull_wgs_per_cu": tun.cull_wgs_per_cu, "probe_ms": round(tun.probe_ms, 3),
                   "cull_ms": round(tun.cull_ms, 3), "autotune_s": round(tune_s, 1)},
        "outputs_equal_per_item": bool(same),
        "contraction_mode": {"eager_ms": round(c_eager_ms, 3),
                             "speedup_vs_sampled_graph": round(graph_ms / c_eager_ms, 2), "plans_s": round(plans_s, 2)},
    }


def host_info():
    """The host the CPU leg runs on: logical CPUs, those this process may use, model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable_cpus": usable, "model": model}


def _median_runs(fn, runs):
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), ts


def cpu_baseline(args, ubo, cam, desc, ivp, lights, fr, d_cube, CUBE, d_spec, d_diff, d_lut, F, pc, contract_out):
    """The FP32 oracle (-O2, -ffp-contract=off) on the host cores, and the
    per-mip errors of the GPU filters against the FP64 oracle.

    value: the 3840x2160 frame (full light culling + shading of a fixed
    seeded fragment sample, scaled to every fragment), median of 5 runs on
    all usable cores (at most 16). Also single-thread, the 2560x1600 /
    1,024-light culling pass and the 512^2 x 1024 BRDF LUT (a row sample,
    scaled), and per-filter times of the 1024^2 -> 512^2 x 10 / 32^2 filters
    (texel samples, scaled). Every scaled figure says what it was scaled from."""
    from oracle import hvx_oracle as O
    from tests.helpers import SMALL_FRAME, camera
    host = host_info()
    threads = max(1, min(16, host["usable_cpus"]))
    rng = np.random.default_rng(5)
    aabbs = O.build_aabbs(ubo)
    maps = dict(spec=d_spec.cpu().numpy(), spec_size=512, spec_mips=10, diff=d_diff.cpu().numpy(), diff_size=32,
                lut=d_lut.cpu().numpy(), lut_w=512, lut_h=512)
    out = {"host": host}

    # ---- the 3840x2160 frame: culling + shading sample, all cores, median of 5
    culled = {}

    def cull_all():
        culled["r"] = O.cull_lights(ubo, aabbs, lights, 128, threads=threads)

    t_cull, cull_runs = _median_runs(cull_all, 5)
    cells, idx, _, _ = culled["r"]
    probe_ids = np.sort(rng.choice(F, 1000, replace=False)).astype(np.uint32)
    t = time.perf_counter()
    O.shade(fr, desc.width, desc.height, ivp, cam, ubo, pc, lights, cells, idx, maps, probe_ids, threads=threads)
    per = (time.perf_counter() - t) / probe_ids.size
    n = int(min(F, max(2000, (args.cpu_seconds / 5.0) / max(per, 1e-9))))
    ids = np.sort(rng.choice(F, n, replace=False)).astype(np.uint32)
    t_shade, _ = _median_runs(lambda: O.shade(fr, desc.width, desc.height, ivp, cam, ubo, pc, lights, cells, idx,
                                              maps, ids, threads=threads), 5)
    frame_s = t_cull + t_shade / n * F
    # single thread: one culling pass and a 300-fragment shading sample
    t = time.perf_counter()
    O.cull_lights(ubo, aabbs, lights, 128, threads=1)
    t_cull1 = time.perf_counter() - t
    ids1 = ids[:300]
    t = time.perf_counter()
    O.shade(fr, desc.width, desc.height, ivp, cam, ubo, pc, lights, cells, idx, maps, ids1, threads=1)
    t_shade1 = time.perf_counter() - t
    frame1_s = t_cull1 + t_shade1 / ids1.size * F
    out.update({
        "value": round(F / frame_s / 1e6, 4),
        "unit": "Mfrag/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP32 oracle (-O2 -ffp-contract=off) on {threads} threads of a {host['usable_cpus']}-CPU share "
                  f"({host['model']}): median of 5 runs of the full culling of the 3840x2160 frame "
                  f"({t_cull:.2f} s) and of shading {n} seeded random fragments ({t_shade:.2f} s); frame time = "
                  f"cull + shading time scaled to {F} fragments",
        "cull_s": round(t_cull, 3),
        "cull_runs_s": [round(x, 3) for x in cull_runs],
        "cluster_light_tests_per_s": round(27648 * 65536 / t_cull, 1),
        "single_thread": {"value": round(F / frame1_s / 1e6, 4), "unit": "Mfrag/s", "cull_s": round(t_cull1, 3),
                          "sample": f"one run: full culling + shading of {ids1.size} fragments, scaled"},
    })

    # ---- the 2560x1600 / 1,024-light culling pass and the 512^2 BRDF LUT
    _, _, ubo0, _ = camera(SMALL_FRAME["width"], SMALL_FRAME["height"], SMALL_FRAME["grid"])
    lights0 = abi_scene_lights(11, SMALL_FRAME["n_lights"])
    aabbs0 = O.build_aabbs(ubo0)
    C0 = SMALL_FRAME["grid"][0] * SMALL_FRAME["grid"][1] * SMALL_FRAME["grid"][2]
    c0_all, _ = _median_runs(lambda: O.cull_lights(ubo0, aabbs0, lights0, 128, threads=threads), 5)
    c0_one, _ = _median_runs(lambda: O.cull_lights(ubo0, aabbs0, lights0, 128, threads=1), 5)
    rows = 8
    lut_ids = np.arange(rows * 512, dtype=np.uint32) + 252 * 512  # 8 middle rows
    lut_all, _ = _median_runs(lambda: O.brdf_lut(512, 512, 1024, lut_ids, threads=threads), 3)
    lut_one, _ = _median_runs(lambda: O.brdf_lut(512, 512, 1024, lut_ids, threads=1), 3)
    out["small_frame"] = {
        "workload": "2560x1600, 16x9x24 clusters, 1024 lights, cap 128; BRDF LUT 512^2 x 1024 samples",
        "cull_ms": round(c0_all * 1e3, 3), "cull_ms_1thread": round(c0_one * 1e3, 3),
        "cluster_light_tests_per_s": round(C0 * SMALL_FRAME["n_lights"] / c0_all, 1),
        "cluster_light_tests_per_s_1thread": round(C0 * SMALL_FRAME["n_lights"] / c0_one, 1),
        "brdf_lut_ms": round(lut_all * 512 / rows * 1e3, 1), "brdf_lut_ms_1thread": round(lut_one * 512 / rows * 1e3, 1),
        "sample": f"culling: median of 5 full passes; LUT: median of 3 runs of {rows} of 512 rows, scaled",
    }

    # ---- the 1024^2 -> 512^2 x 10 / 32^2 filters: CPU time and per-mip error of the GPU outputs
    cube0 = d_cube[:6 * CUBE * CUBE * 4].cpu().numpy()
    spec = maps["spec"]
    SPEC, MIPS, NS = 512, 10, 1024
    per_mip, spec_ms_all, spec_ms_one = [], 0.0, 0.0
    outs = {k: v.cpu().numpy() for k, v in contract_out.items()}
    lv = O.contract_levels(1, CUBE, SPEC, MIPS)
    for m in range(MIPS):
        Sm = max(SPEC >> m, 1)
        T = 6 * Sm * Sm
        ids = np.unique(rng.integers(0, T, min(T, 64))).astype(np.uint32)
        t = time.perf_counter()
        ref32 = O.specular(cube0, CUBE, SPEC, MIPS, m, NS, ids, threads=threads)
        spec_ms_all += (time.perf_counter() - t) / ids.size * T * 1e3
        t = time.perf_counter()
        O.specular(cube0, CUBE, SPEC, MIPS, m, NS, ids[:16], threads=1)
        spec_ms_one += (time.perf_counter() - t) / min(16, ids.size) * T * 1e3
        ref64 = O.specular(cube0, CUBE, SPEC, MIPS, m, NS, ids, dtype=np.float64, threads=threads)[:, :3]
        off = abi_cube_level_offset(SPEC, m)
        plain = spec[off: off + 4 * T].reshape(-1, 4)[ids, :3].astype(np.float64)
        row = {"mip": m, "texels_checked": int(ids.size),
               "plain_vs_fp32_oracle_bitexact": bool(np.array_equal(
                   spec[off: off + 4 * T].reshape(-1, 4)[ids].view(np.uint32), ref32.view(np.uint32))),
               "plain_fp32_vs_fp64_oracle": _max_rel(plain, ref64)}
        if "specular" in outs:
            twin = O.contract(1, cube0, CUBE, SPEC, MIPS, m, ids, threads=threads)
            twin_hot = O.contract(1, cube0, CUBE, SPEC, MIPS, m, ids, threads=threads, max_hot=HOT_TEXELS)
            exact_ids = ids[:8]
            ideal = O.ideal(1, cube0, CUBE, lv[m]["alpha2"], Sm, exact_ids, threads=threads) if m > 0 else None
            for key, name in (("specular", "fp16"), ("specular_split_bf16", "split_bf16"),
                              ("specular_hot", "fp16_bright_split")):
                if key not in outs:
                    continue
                got = outs[key][off: off + 4 * T].reshape(-1, 4)[ids, :3].astype(np.float64)
                row[f"contract_{name}_vs_fp64_twin"] = _max_rel(got, twin_hot if key == "specular_hot" else twin)
                if ideal is not None:
                    row[f"contract_{name}_vs_exact_integral"] = _max_rel(got[:exact_ids.size], ideal)
            if ideal is not None:
                row["plain_fp32_vs_exact_integral"] = _max_rel(plain[:exact_ids.size], ideal)
        per_mip.append(row)
    dids = np.unique(rng.integers(0, 6 * 32 * 32, 48)).astype(np.uint32)
    t = time.perf_counter()
    dref32 = O.diffuse(cube0, CUBE, 32, dids, threads=threads)
    diff_ms_all = (time.perf_counter() - t) / dids.size * 6 * 32 * 32 * 1e3
    t = time.perf_counter()
    O.diffuse(cube0, CUBE, 32, dids[:4], threads=1)
    diff_ms_one = (time.perf_counter() - t) / 4 * 6 * 32 * 32 * 1e3
    dref64 = O.diffuse(cube0, CUBE, 32, dids, dtype=np.float64, threads=threads)[:, :3]
    dplain = maps["diff"].reshape(-1, 4)[dids]
    diffuse_err = {"texels_checked": int(dids.size),
                   "plain_vs_fp32_oracle_bitexact": bool(np.array_equal(dplain.view(np.uint32),
                                                                        dref32.view(np.uint32))),
                   "plain_fp32_vs_fp64_oracle": _max_rel(dplain[:, :3].astype(np.float64), dref64)}
    if "diffuse" in outs:
        dideal = O.ideal(0, cube0, CUBE, 0.0, 32, dids[:8], threads=threads)
        for key, name, hot in (("diffuse", "fp16", 0), ("diffuse_hot", "fp16_bright_split", HOT_TEXELS)):
            if key not in outs:
                continue
            got = outs[key].reshape(-1, 4)[dids, :3].astype(np.float64)
            diffuse_err[f"contract_{name}_vs_fp64_twin"] = _max_rel(got, O.contract(0, cube0, CUBE, 32, 1, 0, dids,
                                                                                    threads=threads, max_hot=hot))
            diffuse_err[f"contract_{name}_vs_exact_integral"] = _max_rel(got[:8], dideal)
    out["ibl_filters"] = {
        "workload": "6x1024^2 cube -> 10-mip GGX chain from 512^2 (1024 samples) and 6x32^2 irradiance",
        "specular_ms": round(spec_ms_all, 1), "specular_ms_1thread": round(spec_ms_one, 1),
        "diffuse_ms": round(diff_ms_all, 1), "diffuse_ms_1thread": round(diff_ms_one, 1),
        "sample": "up to 64 texels per specular mip and 48 irradiance texels (16 and 4 on one thread), scaled to "
                  "every texel; errors: max over the checked texels and channels of |x - ref| / |ref|, exact "
                  "integral (sum over every level-0 texel, FP64) on the first 8",
        "specular_per_mip_err": per_mip,
        "diffuse_err": diffuse_err,
    }
    return out


def _max_rel(got, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref) / np.maximum(np.abs(ref), 1e-30)))


def abi_scene_lights(seed, n):
    from hellovulkan_amd import abi
    return abi.scene_lights(seed, n)


def abi_cube_level_offset(S, m):
    from hellovulkan_amd import abi
    return abi.cube_level_offset_floats(S, m)


if __name__ == "__main__":
    main()
