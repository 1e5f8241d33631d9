"""GPU parity of the contraction (MFMA) cube filter against its FP64 twin in
the oracle: every output mip within 2e-3 relative error, on the procedural
HDR environment (sun of radiance 1e4), at a reduced size in full and at the
2048x1024 -> 1024^2 -> 512^2 x 10 workload on texel subsets."""
import numpy as np
import pytest

from hellovulkan_amd import abi
from oracle import hvx_oracle as O
from tests.helpers import to_dev

pytestmark = pytest.mark.gpu

TOL = 2e-3  # relative, FP16-MFMA contraction vs its FP64 twin


def _env_chain(ctx, seed, S):
    import torch
    eq = abi.scene_equirect(seed, 2 * S, S)
    levels = abi.mip_count(S)
    chain = torch.empty(abi.cube_chain_floats(S, levels), dtype=torch.float32, device="cuda")
    ctx.equirect_to_cube(to_dev(eq), 2 * S, S, chain, S)
    ctx.cube_mips(chain, S, levels)
    torch.cuda.synchronize()
    return chain, levels


def _run(ctx, ftype, S, So, mips, sub, seed=13, precision=abi.PRECISION_FP16_SCALED, per_mip=None):
    import torch
    chain, levels = _env_chain(ctx, seed, S)
    plan = abi.ContractPlan(ctx, ftype, S, So, mips, precision)
    out = torch.full((abi.cube_chain_floats(So, mips),), np.nan, dtype=torch.float32, device="cuda")
    plan.filter(chain, levels, out)
    torch.cuda.synchronize()
    host = out.cpu().numpy()
    lvl0 = chain[: 6 * S * S * 4].cpu().numpy()
    ref_levels = O.contract_levels(ftype, S, So, mips)
    rng = np.random.default_rng(S + So + mips)
    worst = 0.0
    for m in range(mips):
        info = plan.info(m)
        L = ref_levels[m]
        assert (info.kind, info.out_face, info.in_face, info.far_face) == (L["kind"], L["So"], L["Si"], L["Sf"])
        Sm = max(So >> m, 1)
        off = abi.cube_level_offset_floats(So, m)
        got = host[off: off + 6 * Sm * Sm * 4].reshape(-1, 4)
        assert np.all(got[:, 3] == 1.0)
        ids = np.arange(6 * Sm * Sm, dtype=np.uint32)
        if sub and ids.size > sub:
            ids = np.unique(np.concatenate([[0, ids.size - 1], rng.integers(0, ids.size, sub)])).astype(np.uint32)
        ref = O.contract(ftype, lvl0, S, So, mips, m, ids)
        err = O.rel_err(got[ids, :3], ref)
        print(f"type {ftype} {S}->{So} precision {precision} mip {m} kind {info.kind} Si {info.in_face} "
              f"Sf {info.far_face} weights {info.weight_bytes / 1e6:.1f} MB: max rel err vs FP64 {err:.3e} "
              f"({ids.size} texels)")
        if per_mip is not None:
            per_mip.append((info.kind, err))
        worst = max(worst, err)
    plan.close()
    return worst


@pytest.mark.parametrize("S,So,mips,sub", [(256, 128, 8, 3000), (128, 64, 7, 3000), (64, 32, 6, None)])
def test_contract_specular_small(ctx, S, So, mips, sub):
    assert _run(ctx, abi.FILTER_SPECULAR, S, So, mips, sub) <= TOL


@pytest.mark.parametrize("S,So,mips,sub", [(256, 128, 8, 3000), (1024, 512, 10, 40)])
def test_contract_split_bf16(ctx, S, So, mips, sub):
    """Split-BF16 near fields: within the bound everywhere, and the
    low-roughness (near/far) mips an order of magnitude closer to FP64 than
    with FP16 (their far fields stay FP16)."""
    per = []
    assert _run(ctx, abi.FILTER_SPECULAR, S, So, mips, sub, precision=abi.PRECISION_BF16_SPLIT, per_mip=per) <= TOL
    # near and far fields of the low-roughness mips carry split-BF16 products
    # (about 16 significant bits): measured 2.7e-6 .. 4.2e-6 against the
    # 1.5e-4 .. 3.9e-4 of the FP16 mode on the same mips
    split_errs = [e for k, e in per if k == abi.CONTRACT_SPLIT]
    assert split_errs and max(split_errs) <= 2e-5


def test_contract_diffuse_small(ctx):
    assert _run(ctx, abi.FILTER_DIFFUSE, 128, 16, 1, None) <= TOL


def test_contract_specular_full_size(ctx):
    assert _run(ctx, abi.FILTER_SPECULAR, 1024, 512, 10, 40) <= TOL


def test_contract_diffuse_full_size(ctx):
    assert _run(ctx, abi.FILTER_DIFFUSE, 1024, 32, 1, 300) <= TOL


def test_contract_constant_environment(ctx):
    import torch
    S, So, mips = 128, 64, 7
    levels = abi.mip_count(S)
    chain = torch.full((abi.cube_chain_floats(S, levels),), 3.0, dtype=torch.float32, device="cuda")
    plan = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, S, So, mips)
    out = torch.empty(abi.cube_chain_floats(So, mips), dtype=torch.float32, device="cuda")
    plan.filter(chain, levels, out)
    torch.cuda.synchronize()
    rgb = out.cpu().numpy().reshape(-1, 4)[:, :3]
    assert np.abs(rgb / 3.0 - 1.0).max() <= TOL


def test_contract_rejects_bad_arguments(ctx):
    import torch
    with pytest.raises(abi.HvxError) as e:
        abi.ContractPlan(ctx, abi.FILTER_SPECULAR, 96, 32, 4)
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    with pytest.raises(abi.HvxError) as e:
        abi.ContractPlan(ctx, abi.FILTER_DIFFUSE, 64, 16, 2)
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    with pytest.raises(abi.HvxError) as e:
        abi.ContractPlan(ctx, abi.FILTER_SPECULAR, 64, 32, 6, 7)
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    plan = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, 64, 32, 6)
    need = plan.info(0).min_input_level + 1
    chain = torch.zeros(abi.cube_chain_floats(64, 7), dtype=torch.float32, device="cuda")
    out = torch.zeros(abi.cube_chain_floats(32, 6), dtype=torch.float32, device="cuda")
    with pytest.raises(abi.HvxError) as e:
        plan.filter(chain, need - 1, out)
    assert e.value.status == abi.HVX_ERR_INVALID_ARG
    plan.close()


def test_contract_graph_replays_equal_eager(ctx):
    """The filter allocates nothing and resets nothing with memset, so it can
    be captured: graph replays (with eager work between them) equal the eager
    call bit for bit."""
    import torch
    S, So, mips = 128, 64, 7
    chain, levels = _env_chain(ctx, 5, S)
    plan = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, S, So, mips)
    ref = torch.empty(abi.cube_chain_floats(So, mips), dtype=torch.float32, device="cuda")
    plan.filter(chain, levels, ref)
    torch.cuda.synchronize()
    out = torch.full_like(ref, np.nan)
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.filter(chain, levels, out, stream=s)
    for _ in range(3):
        out.fill_(np.nan)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out.view(torch.int32), ref.view(torch.int32))
        torch.zeros(1 << 20, device="cuda").add_(1.0)
    plan.close()


def test_contract_and_plain_against_exact_integral_512_chain(ctx):
    """The benchmark environment (2048x1024 procedural equirect with a sun disk of
    radiance 1e4, 6x1024^2 cube, 512^2 x 10 GGX chain): the FP16 contraction
    filter and the plain 1024-sample filter against the filter integral summed
    over every level-0 texel in FP64, on 24 texels of each rough mip. The
    contraction's quadrature over box-filtered levels stays within 8e-3 of it
    (medians within 2.5e-3; measured max 4.9e-3 at mip 1, 3.9e-3 at mip 2,
    the sun disk's blur in the coarse levels among the worst); the plain
    filter's Monte-Carlo estimate misses it by far more on every mip
    (measured max 0.24 .. 3.6)."""
    import torch
    S, So, mips = 1024, 512, 10
    eq = abi.scene_equirect(77, 2048, 1024, threads=16)
    cube = torch.empty(abi.cube_chain_floats(S, abi.mip_count(S)), dtype=torch.float32, device="cuda")
    ctx.equirect_to_cube(to_dev(eq), 2048, 1024, cube, S)
    ctx.cube_mips(cube, S, abi.mip_count(S))
    plain = torch.empty(abi.cube_chain_floats(So, mips), dtype=torch.float32, device="cuda")
    ctx.filter(abi.FILTER_SPECULAR, cube, S, plain, So, mips, 1024)
    plan = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, S, So, mips)
    con = torch.empty_like(plain)
    plan.filter(cube, abi.mip_count(S), con)
    torch.cuda.synchronize()
    lvl0 = cube.cpu().numpy()[: 6 * S * S * 4]
    P, Cn = plain.cpu().numpy(), con.cpu().numpy()
    rng = np.random.default_rng(5)
    for m in range(1, mips):
        Sm, off = So >> m, abi.cube_chain_floats(So, m)
        n = 6 * Sm * Sm
        ids = np.unique(rng.integers(0, n, min(n, 24))).astype(np.uint32)
        r = m / (mips - 1)
        ref = O.ideal(abi.FILTER_SPECULAR, lvl0, S, (r * r) ** 2, Sm, ids)
        ep = np.abs(P[off:off + 4 * n].reshape(-1, 4)[ids, :3] - ref) / np.abs(ref)
        ec = np.abs(Cn[off:off + 4 * n].reshape(-1, 4)[ids, :3] - ref) / np.abs(ref)
        print(f"mip {m}: plain vs exact max {ep.max():.3e} median {np.median(ep):.3e}; "
              f"contraction vs exact max {ec.max():.3e} median {np.median(ec):.3e}")
        assert ec.max() <= 8e-3 and np.median(ec) <= 2.5e-3
        assert ec.max() < ep.max() and np.median(ec) < np.median(ep)
    plan.close()


@pytest.mark.parametrize("precision", [abi.PRECISION_FP16_SCALED, abi.PRECISION_BF16_SPLIT])
@pytest.mark.parametrize("nprobe", [3, 4, 9])
def test_contract_multi_equals_single(ctx, precision, nprobe):
    """One multi-probe call over n probes of the same sizes equals n
    single-probe calls bit for bit; more probes than reserved is refused.
    The grouped GEMMs take probes in groups of four: n = 3 is one partial
    group, 4 one full group, 9 two full groups and a tail (group indices past
    the first). At 256^2 -> 128^2 x 8, mip 1 is a near/far mip and the
    others full-sphere sums, so both GEMM kinds take the grouped path."""
    import torch
    S, So, mips = 256, 128, 8
    levels = abi.mip_count(S)
    rng = np.random.default_rng(41)
    plan = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, S, So, mips, precision)
    chains, singles, multis = [], [], []
    for k in range(nprobe):
        c = rng.uniform(0.0, 1.0 + 5 * k, 6 * S * S * 4).astype(np.float32)
        chain = torch.empty(abi.cube_chain_floats(S, levels), dtype=torch.float32, device="cuda")
        chain[:c.size].copy_(to_dev(c))
        ctx.cube_mips(chain, S, levels)
        out = torch.full((abi.cube_chain_floats(So, mips),), -1.0, dtype=torch.float32, device="cuda")
        plan.filter(chain, levels, out)
        chains.append(chain)
        singles.append(out)
        multis.append(torch.full_like(out, -2.0))
    d_in = torch.tensor([c.data_ptr() for c in chains], dtype=torch.int64, device="cuda")
    d_out = torch.tensor([o.data_ptr() for o in multis], dtype=torch.int64, device="cuda")
    with pytest.raises(abi.HvxError):
        plan.filter_multi(nprobe, d_in, levels, d_out)   # nothing reserved yet
    plan.reserve(nprobe)
    plan.filter_multi(nprobe, d_in, levels, d_out)
    torch.cuda.synchronize()
    for k in range(nprobe):
        assert torch.equal(multis[k].view(torch.int32), singles[k].view(torch.int32)), f"probe {k}"
    plan.close()


# ----------------------------------------------------------- bright-texel split
def _sun_ids(Sm, n_sun, k=6):
    """Output texels of an Sm face chain nearest the direction n_sun."""
    c = (2.0 * np.arange(Sm) + 1.0) / Sm - 1.0
    u, v = np.meshgrid(c, c)
    one = np.ones_like(u)
    faces = [(one, -v, -u), (-one, -v, u), (u, one, v), (u, -one, -v), (u, -v, one), (-u, -v, -one)]
    d = np.stack([np.stack(f, axis=-1) for f in faces]).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.argsort(-(d @ n_sun))[:k].astype(np.uint32)


SUN = np.array([np.cos(np.radians(35)) * np.cos(np.radians(40)), np.sin(np.radians(35)),
                np.cos(np.radians(35)) * np.sin(np.radians(40))])


def test_contract_hot_split_workload(ctx):
    """The 2048x1024 -> 1024^2 -> 512^2 x 10 workload with the bright-texel
    split at 2048 texels: the threshold and hot count equal the oracle's;
    every mip is within 2e-3 of the split FP64 twin; against the level-0
    sum (24 random texels and the 6 nearest the sun per rough mip) the split
    contraction's worst error is no larger than the unsplit one's, at most
    2e-3 on mips 3-9 and 6e-3 on mips 1-2 (measured 1.5e-3 and 4.9e-3), and
    at the texels nearest the sun it is lower than the unsplit one's on every
    rough mip; the delta mip is bit-identical to the unsplit call's; graph
    replays equal the eager call."""
    import torch
    S, So, mips, K = 1024, 512, 10, 2048
    eq = abi.scene_equirect(77, 2048, 1024, threads=16)
    levels = abi.mip_count(S)
    cube = torch.empty(abi.cube_chain_floats(S, levels), dtype=torch.float32, device="cuda")
    ctx.equirect_to_cube(to_dev(eq), 2048, 1024, cube, S)
    ctx.cube_mips(cube, S, levels)
    plan = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, S, So, mips)
    base = torch.empty(abi.cube_chain_floats(So, mips), dtype=torch.float32, device="cuda")
    plan.filter(cube, levels, base)
    plan.set_hot(K)
    hot = torch.full_like(base, np.nan)
    plan.filter(cube, levels, hot)
    torch.cuda.synchronize()
    T, n_hot = plan.hot_info()
    lvl0 = cube[: 6 * S * S * 4].cpu().numpy()
    T_ref = O.hot_threshold(lvl0, S, K)
    m = np.maximum(np.maximum(lvl0[0::4], lvl0[1::4]), lvl0[2::4])
    assert T == T_ref and n_hot == int(((m > 0) & (m >= T_ref)).sum()) and 0 < n_hot <= K
    print(f"threshold {T}, hot texels {n_hot}")
    Hb, Hh = base.cpu().numpy(), hot.cpu().numpy()
    lv = O.contract_levels(abi.FILTER_SPECULAR, S, So, mips)
    rng = np.random.default_rng(5)
    for mi in range(mips):
        Sm, off = So >> mi, abi.cube_chain_floats(So, mi)
        n = 6 * Sm * Sm
        b = Hb[off:off + 4 * n].reshape(-1, 4)[:, :3]
        h = Hh[off:off + 4 * n].reshape(-1, 4)[:, :3]
        if lv[mi]["kind"] == 0:
            assert np.array_equal(b, h)
            continue
        sun = _sun_ids(Sm, SUN)
        ids = np.concatenate([sun, np.setdiff1d(rng.integers(0, n, 24), sun)]).astype(np.uint32)
        twin = O.contract(abi.FILTER_SPECULAR, lvl0, S, So, mips, mi, ids, threads=16, max_hot=K)
        e_twin = np.abs(h[ids] - twin) / np.abs(twin)
        ref = O.ideal(abi.FILTER_SPECULAR, lvl0, S, lv[mi]["alpha2"], Sm, ids, threads=16)
        eh = np.abs(h[ids] - ref) / np.abs(ref)
        eb = np.abs(b[ids] - ref) / np.abs(ref)
        ns = sun.size
        print(f"mip {mi}: split vs twin {e_twin.max():.2e}; vs level-0 sum: split {eh.max():.2e} "
              f"(median {np.median(eh):.1e}, at the sun {eh[:ns].max():.2e}), unsplit {eb.max():.2e} "
              f"(at the sun {eb[:ns].max():.2e})")
        assert e_twin.max() <= TOL
        assert eh.max() <= eb.max() * (1 + 1e-3)
        assert eh.max() <= (2e-3 if mi >= 3 else 6e-3)
        assert eh[:ns].max() < eb[:ns].max()
    # graph replays of the split call equal the eager result
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    out_g = torch.full_like(base, np.nan)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        plan.filter(cube, levels, out_g, stream=s)
    for _ in range(3):
        out_g.fill_(np.nan)
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out_g, hot)
        plan.filter(cube, levels, base)  # eager work between replays
        torch.cuda.synchronize()
    plan.close()


def test_contract_hot_split_arguments(ctx):
    """set_hot accepts 0..4096; multi-probe calls are refused while it is on
    and accepted again after it is turned off; a small diffuse plan with the
    split stays within 2e-3 of its split twin."""
    import torch
    S, So, mips = 128, 64, 7
    plan = abi.ContractPlan(ctx, abi.FILTER_SPECULAR, S, So, mips)
    with pytest.raises(abi.HvxError):
        plan.set_hot(4097)
    with pytest.raises(abi.HvxError):
        plan.hot_info()
    plan.reserve(2)
    chain, levels = _env_chain(ctx, 21, S)
    outs = [torch.empty(abi.cube_chain_floats(So, mips), dtype=torch.float32, device="cuda") for _ in range(2)]
    cp = torch.tensor([chain.data_ptr()] * 2, dtype=torch.int64, device="cuda")
    op = torch.tensor([o.data_ptr() for o in outs], dtype=torch.int64, device="cuda")
    plan.set_hot(64)
    with pytest.raises(abi.HvxError):
        plan.filter_multi(2, cp, levels, op)
    plan.set_hot(0)
    plan.filter_multi(2, cp, levels, op)
    torch.cuda.synchronize()
    plan.close()
    dplan = abi.ContractPlan(ctx, abi.FILTER_DIFFUSE, S, 32, 1)
    dplan.set_hot(256)
    out = torch.full((abi.cube_chain_floats(32, 1),), np.nan, dtype=torch.float32, device="cuda")
    dplan.filter(chain, levels, out)
    torch.cuda.synchronize()
    lvl0 = chain[: 6 * S * S * 4].cpu().numpy()
    ids = np.arange(6 * 32 * 32, dtype=np.uint32)
    twin = O.contract(abi.FILTER_DIFFUSE, lvl0, S, 32, 1, 0, ids, max_hot=256)
    got = out.cpu().numpy().reshape(-1, 4)[:, :3]
    assert (np.abs(got - twin) / np.abs(twin)).max() <= TOL
    dplan.close()
