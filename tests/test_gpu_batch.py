"""GPU parity of the batched path (hvx_batch_*): probes of uneven size
prefiltered by the persistent probe kernel and views of uneven light counts
culled by the persistent culling kernel must equal the per-item entry points
(hvx_ibl_filter type 1, hvx_cf_cull_lights) bit for bit — eagerly, under
every tile shape, after autotuning and on repeated graph replays. The
per-item entry points are themselves pinned to the oracle (test_gpu_cf,
test_gpu_ibl); one small probe and one view are also checked against the
oracle directly here."""
import numpy as np
import pytest

from hellovulkan_amd import abi
from oracle import hvx_oracle as O
from tests.helpers import camera, to_dev

pytestmark = pytest.mark.gpu

# (input face, output face, output mips)
PROBES = [(16, 8, 4), (32, 16, 5), (64, 32, 6), (24, 12, 4), (48, 48, 6), (8, 4, 3), (32, 1, 1), (40, 20, 5)]
# (width, height, grid, n_lights, cap, light seed)
VIEWS = [
    (2560, 1600, (16, 9, 24), 1024, 128, 21),
    (640, 480, (8, 4, 6), 0, 16, 22),
    (700, 500, (7, 5, 3), 300, 7, 23),
    (3840, 2160, (32, 18, 48), 8192, 128, 24),
    (1920, 1080, (16, 9, 24), 65536, 150, 25),
    (1280, 720, (4, 4, 4), 2000, 1024, 26),
]
SAMPLES = 128


class Scene:
    def __init__(self, ctx, probes=PROBES, views=VIEWS, samples=SAMPLES):
        import torch
        rng = np.random.default_rng(5)
        self.samples = samples
        self.probe_in, self.probe_out, self.probe_ref, self.probe_specs = [], [], [], list(probes)
        for S, So, mips in probes:
            c = rng.uniform(0.0, 4.0, 6 * S * S * 4).astype(np.float32)
            c[3::4] = 1.0
            cube = to_dev(c)
            ref = torch.empty(abi.cube_chain_floats(So, mips), dtype=torch.float32, device="cuda")
            ctx.filter(abi.FILTER_SPECULAR, cube, S, ref, So, mips, samples)
            self.probe_in.append(cube)
            self.probe_ref.append(ref)
            self.probe_out.append(torch.full_like(ref, -1.0))
        self.view_specs, self.view_args, self.view_ref, self.view_out = list(views), [], [], []
        for W, H, grid, n, cap, seed in views:
            _, _, ubo, _ = camera(W, H, grid)
            Cn = grid[0] * grid[1] * grid[2]
            aabbs = torch.empty(Cn * 8, dtype=torch.float32, device="cuda")
            ctx.build_aabbs(ubo, aabbs)
            lights_np = abi.scene_lights(seed, n)
            lights = to_dev(lights_np.view(np.uint8)) if n else torch.empty(48, dtype=torch.uint8, device="cuda")
            ref = self._view_buffers(Cn, cap)
            ctx.cull_lights(ubo, aabbs, lights, n, cap, ref[0], ref[1], Cn * cap, ref[2])
            self.view_args.append((ubo, aabbs, lights, n, cap, Cn, lights_np))
            self.view_ref.append(ref)
            self.view_out.append(self._view_buffers(Cn, cap))
        torch.cuda.synchronize()

    @staticmethod
    def _view_buffers(Cn, cap):
        import torch
        return (torch.full((Cn * 2,), -1, dtype=torch.int32, device="cuda"),
                torch.full((Cn * cap,), -1, dtype=torch.int32, device="cuda"),
                torch.full((1,), -1, dtype=torch.int32, device="cuda"))

    def batch(self, ctx):
        probes = [abi.probe(i, s[0], o, s[1], s[2]) for i, o, s in zip(self.probe_in, self.probe_out, self.probe_specs)]
        views = [abi.view(a[0], a[1], a[2], a[3], a[4], o[0], o[1], a[5] * a[4], o[2])
                 for a, o in zip(self.view_args, self.view_out)]
        return abi.Batch(ctx, probes, self.samples, views)

    def poison(self, stream=None):
        import torch
        with torch.cuda.stream(stream) if stream is not None else _nullctx():
            for o in self.probe_out:
                o.fill_(-1.0)
            for o in self.view_out:
                for t in o:
                    t.fill_(-1)

    def check(self, what):
        import torch
        torch.cuda.synchronize()
        for k, (got, ref) in enumerate(zip(self.probe_out, self.probe_ref)):
            assert torch.equal(got.view(torch.int32), ref.view(torch.int32)), f"{what}: probe {k} {self.probe_specs[k]}"
        for k, (got, ref) in enumerate(zip(self.view_out, self.view_ref)):
            tot = int(ref[2].item())
            assert int(got[2].item()) == tot, f"{what}: view {k} total"
            assert torch.equal(got[0], ref[0]), f"{what}: view {k} cells"
            assert torch.equal(got[1][:tot], ref[1][:tot]), f"{what}: view {k} indices"


class _nullctx:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


@pytest.fixture(scope="module")
def scene(ctx):
    return Scene(ctx)


def test_batch_matches_per_item(ctx, scene):
    b = scene.batch(ctx)
    scene.poison()
    b.run()
    scene.check("eager default tuning")
    b.close()


@pytest.mark.parametrize("tile,pw,shape,cw", [(32, 1, 1, 1), (100, 16, 2, 3), (4096, 3, 3, 16), (256, 4, 0, 8)])
def test_batch_every_tiling(ctx, scene, tile, pw, shape, cw):
    b = scene.batch(ctx)
    b.set_tuning(tile, pw, shape, cw)
    t = b.tuning
    assert (t.probe_tile, t.probe_wgs_per_cu, t.cull_shape, t.cull_wgs_per_cu) == (tile, pw, shape, cw)
    scene.poison()
    b.run()
    scene.check(f"tiling {tile},{pw},{shape},{cw}")
    b.close()


def test_batch_autotune_then_graph_replays(ctx, scene):
    import torch
    b = scene.batch(ctx)
    t = b.autotune()
    print(f"autotuned: tile {t.probe_tile} x {t.probe_wgs_per_cu}/CU ({t.probe_ms:.3f} ms), "
          f"cull shape {t.cull_shape} x {t.cull_wgs_per_cu}/CU ({t.cull_ms:.3f} ms)")
    assert t.probe_ms > 0 and t.cull_ms > 0
    scene.poison()
    b.run()
    scene.check("after autotune")
    s = torch.cuda.Stream()
    with pytest.raises(abi.HvxError):
        b.launch(s)  # not captured yet
    b.capture(s)
    for r in range(3):
        scene.poison(s)
        b.launch(s)
        s.synchronize()
        scene.check(f"graph replay {r}")
        torch.zeros(1 << 20, device="cuda").add_(1.0)  # eager work between replays
        torch.cuda.synchronize()
    b.close()


def test_batch_only_probes_or_only_views(ctx):
    only_p = Scene(ctx, probes=PROBES[:3], views=[])
    b = only_p.batch(ctx)
    only_p.poison()
    b.run()
    only_p.check("probes only")
    b.close()
    only_v = Scene(ctx, probes=[], views=VIEWS[:3])
    b = only_v.batch(ctx)
    only_v.poison()
    b.run()
    only_v.check("views only")
    b.close()


def test_batch_against_oracle(ctx):
    """A small probe and view through the batch, against the oracle."""
    sc = Scene(ctx, probes=[(16, 8, 4)], views=[VIEWS[2]], samples=64)
    b = sc.batch(ctx)
    sc.poison()
    b.run()
    sc.check("small")
    out = sc.probe_out[0].cpu().numpy()
    cube = sc.probe_in[0].cpu().numpy()
    for m in range(4):
        Sm = max(8 >> m, 1)
        off = abi.cube_level_offset_floats(8, m)
        ids = np.arange(6 * Sm * Sm, dtype=np.uint32)
        ref = O.specular(cube, 16, 8, 4, m, 64, ids)
        assert O.rel_err(out[off: off + ids.size * 4].reshape(-1, 4), ref) <= 1e-5
    ubo, _, _, n, cap, Cn, lights_np = sc.view_args[0]
    ref_cells, ref_idx, ref_tot, _ = O.cull_lights(ubo, O.build_aabbs(ubo), lights_np, cap)
    cells = sc.view_out[0][0].cpu().numpy().view(abi.CELL_DTYPE)
    assert int(sc.view_out[0][2].item()) == ref_tot
    assert np.array_equal(cells["count"], ref_cells["count"]) and np.array_equal(cells["offset"], ref_cells["offset"])
    assert np.array_equal(sc.view_out[0][1].cpu().numpy().view(np.uint32)[:ref_tot], ref_idx)
    b.close()


def test_batch_rejects_bad_arguments(ctx, scene):
    import torch
    a = scene.view_args[1]
    o = scene.view_out[1]
    v = abi.view(a[0], a[1], a[2], a[3], a[4], o[0], o[1], a[5] * a[4], o[2])
    with pytest.raises(abi.HvxError) as e:
        abi.Batch(ctx, [], SAMPLES, [v] * 65)
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    small = abi.view(a[0], a[1], a[2], a[3], a[4], o[0], o[1], a[5] * a[4] - 1, o[2])
    with pytest.raises(abi.HvxError) as e:
        abi.Batch(ctx, [], SAMPLES, [small])
    assert e.value.status == abi.HVX_ERR_CAPACITY
    out = torch.empty(abi.cube_chain_floats(8, 4), dtype=torch.float32, device="cuda")
    bad = abi.probe(scene.probe_in[0], 16, out, 8, 5)  # 8 has 4 mips
    with pytest.raises(abi.HvxError) as e:
        abi.Batch(ctx, [bad], SAMPLES, [])
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    b = abi.Batch(ctx, [abi.probe(scene.probe_in[0], 16, out, 8, 4)], SAMPLES, [])
    for tun in [(16, 4, 0, 8), (256, 0, 0, 8), (256, 4, 4, 8), (256, 4, 0, 17)]:
        with pytest.raises(abi.HvxError) as e:
            b.set_tuning(*tun)
        assert e.value.status == abi.HVX_ERR_INVALID_ARG
    b.close()


CONTRACT_PROBES = [(16, 8, 4), (32, 16, 5), (64, 32, 6), (64, 32, 6), (128, 64, 7), (32, 16, 5)]
# classes of 6 and 9 probes: the multi-probe GEMMs' second and third probe groups
CONTRACT_LARGE_CLASSES = [(32, 16, 5)] * 6 + [(64, 32, 6)] * 9
# 256^2 -> 128^2 x 8: mip 1 is a near/far mip (five probes, two groups)
CONTRACT_SPLIT_CLASS = [(256, 128, 8)] * 5


@pytest.mark.parametrize("precision", [abi.PRECISION_FP16_SCALED, abi.PRECISION_BF16_SPLIT])
@pytest.mark.parametrize("probes", [CONTRACT_PROBES, CONTRACT_LARGE_CLASSES, CONTRACT_SPLIT_CLASS],
                         ids=["mixed", "large-classes", "near-far-class"])
def test_batch_contraction_mode_matches_per_probe(ctx, precision, probes):
    """Contraction mode: every probe equals the per-probe contraction filter
    (its plan over the probe's level 0 and box mips) bit for bit on repeated
    eager runs, with the views still culled as per-item; switching back
    restores the sampled prefilter's outputs."""
    import torch
    sc = Scene(ctx, probes=probes, views=VIEWS[:2])
    refs = []
    for cube, (S, So, mips) in zip(sc.probe_in, sc.probe_specs):
        levels = abi.mip_count(S)
        chain = torch.empty(abi.cube_chain_floats(S, levels), dtype=torch.float32, device="cuda")
        chain[:6 * S * S * 4].copy_(cube)
        ctx.cube_mips(chain, S, levels)
        plan = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, S, So, mips, precision)
        out = torch.empty(abi.cube_chain_floats(So, mips), dtype=torch.float32, device="cuda")
        plan.filter(chain, levels, out)
        torch.cuda.synchronize()
        plan.close()
        refs.append(out)
    sampled = sc.probe_ref
    sc.probe_ref = refs
    b = sc.batch(ctx)
    b.set_contract(True, precision)
    for r in range(2):
        sc.poison()
        b.run()
        sc.check(f"contraction mode, eager run {r}")
    b.set_contract(False)
    sc.probe_ref = sampled
    sc.poison()
    b.run()
    sc.check("sampled prefilter again")
    b.close()


def test_batch_contraction_mode_rejects_unsupported_sizes(ctx):
    """The contraction filter needs power-of-two faces: a 24^2 probe makes
    hvx_batch_set_contract fail cleanly and the batch stays usable."""
    sc = Scene(ctx, probes=[(24, 12, 4), (32, 16, 5)], views=[])
    b = sc.batch(ctx)
    with pytest.raises(abi.HvxError):
        b.set_contract(True)
    sc.poison()
    b.run()
    sc.check("sampled prefilter after a refused switch")
    b.close()


@pytest.mark.skip(reason="graph capture of the contraction mode ended in a host-side crash (segmentation fault "
                         "inside hvx_batch_capture) whose cause is not established; the library refuses that "
                         "capture (HVX_ERR_UNSUPPORTED) until it is")
def test_batch_contraction_mode_graph_replays(ctx):
    import torch
    sc = Scene(ctx, probes=CONTRACT_PROBES, views=VIEWS[:2])
    b = sc.batch(ctx)
    b.set_contract(True)
    s = torch.cuda.Stream()
    b.capture(s)
    b.launch(s)
    s.synchronize()
    b.close()
