"""Pins of the CPU oracle (CPU only). The reference ships no tests or golden
data for this path, so each oracle stage is checked against what the
mathematics fixes: closed forms, special cases, invariants and independent
float64 brute force, chosen so that a dropped term, a wrong sign or index, or
a transposed operand fails at least one of them."""
import math

import numpy as np
import pytest

from hellovulkan_amd import abi
from oracle import hvx_oracle as O
from tests.helpers import SMALL_FRAME, FAR, FOVY, NEAR, camera, cells_to_lists

F32 = np.float32


# ------------------------------------------------------------ a1 camera/UBO
def test_camera_ubo_closed_forms():
    desc, cam, ubo, ivp = camera(2560, 1600, (16, 9, 24))
    P = np.array(cam.projection, np.float64).reshape(4, 4).T   # row-major view of column-major
    th = math.tan(FOVY / 2)
    aspect = 2560 / 1600
    assert P[0, 0] == pytest.approx(1 / (aspect * th), rel=1e-6)
    assert P[1, 1] == pytest.approx(-1 / th, rel=1e-6)           # Vulkan Y flip
    assert P[2, 2] == pytest.approx(FAR / (NEAR - FAR), rel=1e-6)  # depth [0, 1]
    assert P[3, 2] == -1.0
    assert P[2, 3] == pytest.approx(-(FAR * NEAR) / (FAR - NEAR), rel=1e-6)
    # near plane -> depth 0, far plane -> depth 1
    for z, d in ((-NEAR, 0.0), (-FAR, 1.0)):
        clip = P @ np.array([0, 0, z, 1])
        assert clip[2] / clip[3] == pytest.approx(d, abs=1e-6)
    V = np.array(cam.view, np.float64).reshape(4, 4).T
    eye = np.array([0, 1, 6, 1.0])
    tgt = np.array([0, 2.5, 0, 1.0])
    assert np.allclose(V @ eye, [0, 0, 0, 1], atol=1e-6)
    vt = V @ tgt
    assert abs(vt[0]) < 1e-6 and abs(vt[1]) < 1e-6 and vt[2] < 0      # target straight ahead (-z)
    assert np.allclose(V[:3, :3] @ V[:3, :3].T, np.eye(3), atol=1e-6)
    # slice mapping: log2(near)*s + b = 0 and log2(far)*s + b = Z
    s, b = ubo.slice_scaling, ubo.slice_bias
    assert math.log2(NEAR) * s + b == pytest.approx(0.0, abs=1e-5)
    assert math.log2(FAR) * s + b == pytest.approx(24.0, abs=1e-5)
    invP = np.array(ubo.camera_inverse_projection, np.float64).reshape(4, 4).T
    assert np.allclose(invP @ P, np.eye(4), atol=1e-5)
    IVP = np.array(ivp, np.float64).reshape(4, 4).T
    assert np.allclose(IVP @ P @ V, np.eye(4), atol=1e-4)


# ----------------------------------------------------------------- a2 AABBs
def _frustum_aabbs_f64(ubo, width, height):
    """Independent float64 cluster AABBs from fovy/aspect (no inverse
    projection): tile corners on the near plane's NDC grid, rays through them
    cut by the exponential slice planes."""
    X, Y, Z = ubo.slice_count_x, ubo.slice_count_y, ubo.slice_count_z
    th = math.tan(FOVY / 2)
    asp = width / height
    out = np.zeros((X * Y * Z, 6))
    for z in range(Z):
        zn = -NEAR * (FAR / NEAR) ** (z / Z)
        zf = -NEAR * (FAR / NEAR) ** ((z + 1) / Z)
        for y in range(Y):
            for x in range(X):
                pts = []
                for sx, sy in ((x / X, y / Y), ((x + 1) / X, (y + 1) / Y)):
                    nx, ny = 2 * sx - 1, 2 * sy - 1
                    ray = np.array([nx * th * asp, -ny * th, -1.0])   # Vulkan: +y NDC is down
                    for d in (zn, zf):
                        pts.append(ray * (d / ray[2]))
                pts = np.array(pts)
                out[x + y * X + z * X * Y] = np.concatenate([pts.min(0), pts.max(0)])
    return out


def test_aabbs_against_independent_frustum():
    _, _, ubo, _ = camera(640, 400, (4, 3, 6))
    got = O.build_aabbs(ubo).astype(np.float64)
    ref = _frustum_aabbs_f64(ubo, 640, 400)
    g = np.concatenate([got[:, 0:3], got[:, 4:7]], axis=1)
    scale = np.maximum(np.abs(ref), 1e-3)
    assert np.max(np.abs(g - ref) / scale) < 2e-5
    assert np.all(got[:, 3] == 0) and np.all(got[:, 7] == 0)


def test_aabbs_invariants():
    _, _, ubo, _ = camera(SMALL_FRAME["width"], SMALL_FRAME["height"], SMALL_FRAME["grid"])
    a = O.build_aabbs(ubo)
    X, Y, Z = SMALL_FRAME["grid"]
    assert np.all(a[:, 0:3] <= a[:, 4:7])
    a = a.reshape(Z, Y, X, 8)
    # adjacent z-slices share their separating plane (same pow() input)
    assert np.allclose(a[1:, :, :, 6], a[:-1, :, :, 2], rtol=3e-7, atol=0)
    assert np.allclose(a[0, :, :, 6], -NEAR, rtol=1e-6)
    assert np.allclose(a[-1, :, :, 2], -FAR, rtol=1e-6)
    f64 = O.build_aabbs(ubo, dtype=np.float64).reshape(Z, Y, X, 8)
    assert np.max(np.abs(a - f64) / np.maximum(np.abs(f64), 1e-3)) < 1e-5


# --------------------------------------------------------------- a3 culling
def _sphere_aabb_f64(pv, r, box):
    lo, hi = box[:3], box[4:7]
    d = np.where(pv < lo, lo - pv, np.where(pv > hi, hi - pv, 0.0))
    return float(np.sum(d * d)) - float(r) * float(r)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cull_brute_force_f64(seed):
    _, _, ubo, _ = camera(320, 200, (4, 3, 4))
    lights = abi.scene_lights(seed, 64)
    aabbs = O.build_aabbs(ubo)
    cap = 16
    cells, idx, tot, ov = O.cull_lights(ubo, aabbs, lights, cap)
    lists = cells_to_lists(cells, idx)
    V = np.array(ubo.camera_view, np.float64).reshape(4, 4).T
    pvs = (V @ np.concatenate([lights["position"][:, :3].astype(np.float64), np.ones((64, 1))], 1).T).T[:, :3]
    n_over = 0
    for c, lst in enumerate(lists):
        box = aabbs[c].astype(np.float64)
        margins = np.array([_sphere_aabb_f64(pvs[i], lights["radius"][i], box) for i in range(64)])
        hits = [i for i in range(64) if margins[i] < 0]
        ambiguous = set(np.nonzero(np.abs(margins) < 1e-4)[0].tolist())
        clear = [i for i in hits if i not in ambiguous]
        expect = hits[:cap]
        if not ambiguous:
            assert list(lst) == expect
        else:  # every clear hit before the cap position must be listed
            for i in clear:
                if len([h for h in hits if h < i]) < cap - len(ambiguous):
                    assert i in lst
        assert list(lst) == sorted(lst) and len(lst) <= cap
        n_over += len(hits) > cap and not ambiguous
    assert tot == sum(len(l) for l in lists)
    assert ov >= n_over


def test_cull_truncation_and_completeness():
    _, _, ubo, _ = camera(320, 200, (4, 3, 6))
    lights = abi.scene_lights(9, 500)
    aabbs = O.build_aabbs(ubo)
    full_cells, full_idx, _, ov_full = O.cull_lights(ubo, aabbs, lights, 500)
    assert ov_full == 0
    for cap in (1, 5, 33):
        cells, idx, _, ov = O.cull_lights(ubo, aabbs, lights, cap)
        for lf, lc in zip(cells_to_lists(full_cells, full_idx), cells_to_lists(cells, idx)):
            assert lc == lf[:cap]  # the first `cap` of the complete list
        assert ov == int(np.sum(full_cells["count"] > cap))
    # offsets are the exclusive prefix sum of the counts
    assert np.array_equal(full_cells["offset"][1:], np.cumsum(full_cells["count"])[:-1])


def test_cull_no_lights():
    _, _, ubo, _ = camera(320, 200, (4, 3, 6))
    cells, idx, tot, ov = O.cull_lights(ubo, O.build_aabbs(ubo), abi.scene_lights(1, 0), 8)
    assert tot == 0 and ov == 0 and np.all(cells["count"] == 0) and np.all(cells["offset"] == 0)


# ----------------------------------------------------------- a9 Hammersley
def test_hammersley_kat():
    h = O.hammersley(16)
    vdc = [0, .5, .25, .75, .125, .625, .375, .875, .0625, .5625, .3125, .8125, .1875, .6875, .4375, .9375]
    assert np.array_equal(h[:, 1], np.array(vdc, np.float32))
    assert np.array_equal(h[:, 0], (np.arange(16) / 16).astype(np.float32))


# ----------------------------------------------------------------- a10 GGX
def test_ggx_sample_properties():
    rng = np.random.default_rng(4)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        xi = rng.uniform(0, 1, 2).astype(F32)
        for rough in (0.0, 0.3, 1.0):
            h = O.ggx_sample(xi, n.astype(F32), rough)
            assert abs(np.linalg.norm(h) - 1) < 1e-6
            assert float(np.dot(h, n)) >= -1e-6
            if rough == 0.0:
                assert np.allclose(h, n, atol=2e-7)
    # rough = 1: cos(theta) = sqrt(1 - xi.y) (a^2 - 1 = 0): the pdf is uniform in cos^2
    h = O.ggx_sample(np.array([0.25, 0.36], F32), np.array([0, 0, 1], F32), 1.0)
    assert h[2] == pytest.approx(math.sqrt(1 - 0.36), rel=1e-6)
    # for N = +Z the frame is T = (0,-1,0), B = (1,0,0): world = (hy, -hx, hz),
    # i.e. azimuth phi - pi/2; phi = 2 pi * 0.25 lands on +X
    assert h[0] == pytest.approx(math.sqrt(0.36), rel=1e-6) and abs(h[1]) < 1e-6


# ------------------------------------------------------------- a11 BRDF LUT
def test_brdf_lut_closed_forms():
    W, H, n = 16, 4096, 256
    # top row: roughness = 0.5/H ~ 1e-4 -> H = N, G = 1: (1-(1-NoV)^5, (1-NoV)^5)
    ids = (H - 1) * W + np.arange(W)
    lut = O.brdf_lut(W, H, n, ids.astype(np.uint32), dtype=np.float64)
    nov = (np.arange(W) + 0.5) / W
    fc = (1 - nov) ** 5
    assert np.allclose(lut[:, 0], 1 - fc, atol=2e-4)
    assert np.allclose(lut[:, 1], fc, atol=2e-4)
    full = O.brdf_lut(32, 32, 512, dtype=np.float64)
    assert np.all(full >= 0) and np.all(full.sum(1) <= 1 + 1e-9)   # x + y = mean(G_Vis) <= 1
    f32 = O.brdf_lut(32, 32, 512)
    # FP32 rounding only; G_Vis divides by NoV, so the grazing column (NoV = 1/64)
    # carries the largest FP32 error
    err = np.abs(f32 - full).reshape(32, 32, 2)
    assert err.max() < 5e-5 and err[:, 3:].max() < 8e-6


# ------------------------------------------------- a12 equirect -> cube, a16
def test_texel_dir_face_centres_and_layout():
    axes = {0: (1, 0, 0), 1: (-1, 0, 0), 2: (0, 1, 0), 3: (0, -1, 0), 4: (0, 0, 1), 5: (0, 0, -1)}
    for f, ax in axes.items():
        assert np.allclose(O.texel_dir(f, 0, 0, 1), ax, atol=1e-7)
    # the shaders' UVToXYZ + y flip lands output texel (x, y) on Vulkan face
    # f at (s, t) = ((x+.5)/S, (y+.5)/S): output and sampling layouts agree
    S = 8
    for f in range(6):
        for y in range(S):
            for x in range(S):
                face, st = O.cube_face_st(O.texel_dir(f, x, y, S))
                assert face == f
                assert np.allclose(st, [(x + .5) / S, (y + .5) / S], atol=1e-6)


def test_cube_edge_adjacency():
    S = 6
    def centre(f, i, j):
        d = O.texel_dir(f, i, j, S)
        return d / np.abs(d).max()   # point on the cube surface
    for f in range(6):
        for k in range(S):
            for (i, j), (bi, bj) in (((-1, k), (0, k)), ((S, k), (S - 1, k)), ((k, -1), (k, 0)), ((k, S), (k, S - 1))):
                nf, ni, nj = O.cube_edge_texel(f, i, j, S)
                assert nf != f and 0 <= ni < S and 0 <= nj < S
                # the texel across the edge touches the border texel: centres are
                # one texel apart along the surface (2/S on each face)
                a, b = centre(f, bi, bj), centre(nf, ni, nj)
                assert np.abs(a - b).sum() == pytest.approx(2.0 / S, rel=1e-5)
                # and the relation is symmetric
                back = [t for t in ((-1, None), (S, None), (None, -1), (None, S))]
                found = False
                for ii, jj in ((-1, nj), (S, nj), (ni, -1), (ni, S)):
                    if O.cube_edge_texel(nf, ii, jj, S) == (f, bi, bj):
                        found = True
                assert found


def _cube_from_fn(fn, S):
    cube = np.zeros((6, S, S, 4), np.float32)
    for f in range(6):
        for y in range(S):
            for x in range(S):
                cube[f, y, x, :3] = fn(O.texel_dir(f, x, y, S).astype(np.float64))
                cube[f, y, x, 3] = 1
    return cube.reshape(-1)


def test_cube_sampling_texel_centres_and_continuity():
    S = 8
    rng = np.random.default_rng(2)
    cube = rng.uniform(0, 1, 6 * S * S * 4).astype(np.float32)
    for f in range(6):
        for (x, y) in ((0, 0), (3, 5), (7, 7), (2, 6)):
            d = O.texel_dir(f, x, y, S)
            got = O.sample_cube(cube, S, d)[0]
            ref = cube.reshape(6, S, S, 4)[f, y, x, :3]
            assert np.allclose(got, ref, atol=2e-6)
    # a smooth environment sampled along a great circle crossing face edges
    smooth = _cube_from_fn(lambda d: np.array([1 + d[0], 1 + d[1], 1 + d[2]]), 32)
    t = np.linspace(0, 2 * np.pi, 4001)
    dirs = np.stack([np.cos(t), 0.3 * np.ones_like(t), np.sin(t)], 1)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    vals = O.sample_cube(smooth, 32, dirs.astype(np.float32))
    assert np.max(np.abs(np.diff(vals, axis=0))) < 0.01
    assert np.allclose(vals, 1 + dirs, atol=0.02)


def test_equirect_constant_and_linear_maps():
    w, h, S = 64, 32, 16
    const = np.ones((h, w, 4), np.float32) * np.array([2.0, 3.0, 4.0, 1.0], np.float32)
    out = O.equirect_to_cube(const, S)
    assert np.allclose(out[:, :3], [2, 3, 4], rtol=1e-6) and np.all(out[:, 3] == 1)
    # a map linear in the texture's t coordinate: bilinear reproduces it, so the
    # cube holds t(direction) = 1 - acos(y)/pi away from the REPEAT wrap rows
    tt = ((np.arange(h) + 0.5) / h).astype(np.float32)
    lin = np.zeros((h, w, 4), np.float32)
    lin[:, :, 0] = tt[:, None]
    out = O.equirect_to_cube(lin, S, dtype=np.float64)
    dirs = np.array([O.texel_dir(f, x, y, S) for f in range(6) for y in range(S) for x in range(S)], np.float64)
    t_exp = 1 - np.arccos(np.clip(dirs[:, 1], -1, 1)) / np.pi
    keep = (t_exp > 1.0 / h) & (t_exp < 1 - 1.0 / h)
    assert np.allclose(out[keep, 0], t_exp[keep], atol=1e-6)


# --------------------------------------------------------------- a13 mips
def test_mips_are_box_means():
    S, L = 16, 5
    rng = np.random.default_rng(3)
    chain = np.zeros(abi.cube_chain_floats(S, L), np.float32)
    chain[: 6 * S * S * 4] = rng.uniform(0, 10, 6 * S * S * 4)
    out = O.cube_mips(chain, S, L)
    for m in range(1, L):
        sp, sn = S >> (m - 1), S >> m
        src = out[abi.cube_level_offset_floats(S, m - 1):][: 6 * sp * sp * 4].reshape(6, sp, sp, 4)
        dst = out[abi.cube_level_offset_floats(S, m):][: 6 * sn * sn * 4].reshape(6, sn, sn, 4)
        a, b = src[:, 0::2, 0::2], src[:, 0::2, 1::2]
        c, d = src[:, 1::2, 0::2], src[:, 1::2, 1::2]
        assert np.array_equal(dst, ((a + b) + (c + d)) * F32(0.25))


# ------------------------------------------------------------ a14 diffuse
def _theta_grid():
    th, t = [], F32(0)
    while t < F32(np.pi) * F32(0.5):
        th.append(t)
        t = F32(t + F32(0.025))
    return np.array(th, np.float64)


def test_diffuse_constant_and_linear_environment():
    S, Sd = 32, 4
    c = np.array([1.5, 0.5, 2.0])
    cube = _cube_from_fn(lambda d: c, S)
    out = O.diffuse(cube, S, Sd, dtype=np.float64)
    th = _theta_grid()
    assert len(th) == 63
    expect = np.pi * c * np.mean(np.cos(th) * np.sin(th))      # the quadrature's closed form
    assert np.allclose(out[:, :3], expect, rtol=1e-6)
    # linear environment L(d) = b + a.d: irradiance/pi-normalised = b + (2/3) a.N
    a, b = np.array([0.7, -0.4, 0.25]), 2.0
    cube = _cube_from_fn(lambda d: np.full(3, b + a @ d), 64)
    ids = np.arange(6 * Sd * Sd, dtype=np.uint32)
    out = O.diffuse(cube, 64, Sd, ids, dtype=np.float64)
    N = np.array([O.texel_dir(f, x, y, Sd) for f in range(6) for y in range(Sd) for x in range(Sd)], np.float64)
    ref = b + (2.0 / 3.0) * (N @ a)
    assert np.max(np.abs(out[:, 0] - ref) / ref) < 5e-3


# ----------------------------------------------------------- a15 specular
def test_specular_special_cases():
    S, So, mips, n = 16, 8, 4, 64
    rng = np.random.default_rng(6)
    cube = rng.uniform(0.5, 2, 6 * S * S * 4).astype(np.float32)
    # mip 0 has roughness 0: every sample is L = N, the output is the bilinear sample at N
    out = O.specular(cube, S, So, mips, 0, n)
    dirs = np.array([O.texel_dir(f, x, y, So) for f in range(6) for y in range(So) for x in range(So)], np.float32)
    ref = O.sample_cube(cube, S, dirs)
    assert np.allclose(out[:, :3], ref, rtol=1e-5)
    # constant environment -> the constant, at every roughness
    const = np.tile(np.array([3.0, 1.0, 0.25, 1.0], np.float32), 6 * S * S)
    for m in range(mips):
        o = O.specular(const, S, So, mips, m, n)
        assert np.allclose(o[:, :3], [3.0, 1.0, 0.25], rtol=1e-5)
    # static-viewport emulation: mip 1 of 4 covers only uv in [-1, 0]
    e = O.specular(cube, S, So, mips, 1, n, emulate=True)
    ne = O.specular(cube, S, So, mips, 1, n, emulate=False)
    assert not np.allclose(e, ne)


def test_specular_f32_vs_f64():
    S, So, mips, n = 32, 8, 4, 128
    rng = np.random.default_rng(8)
    cube = rng.uniform(0.1, 3, 6 * S * S * 4).astype(np.float32)
    for m in range(mips):
        a = O.specular(cube, S, So, mips, m, n)
        b = O.specular(cube, S, So, mips, m, n, dtype=np.float64)
        assert O.rel_err(a, b) < 2e-5


# ----------------------------------------------------- a5-a8 shading forms
def _one_fragment_scene(rough, metal, albedo, light_d, radius, color):
    """1x1 view looking down -z from the origin; the fragment sits on the
    axis, its normal faces the camera and one light sits between them, so
    N = V = L = H and every dot product is 1."""
    desc, cam, ubo, ivp = abi.camera_build((0, 0, 0), (0, 0, -1), FOVY, NEAR, FAR, 1, 1, (1, 1, 1))
    lin = 5.0
    zndc = np.float32(FAR * (lin - NEAR) / (lin * (FAR - NEAR)))
    clip = np.array([0, 0, zndc, 1.0])
    IVP = np.array(ivp, np.float64).reshape(4, 4).T
    wh = IVP @ clip
    wp = wh[:3] / wh[3]
    fr = dict(depth=np.array([zndc], F32), nx=np.array([0], F32), ny=np.array([0], F32),
              nz=np.array([1], F32), ar=np.array([albedo[0]], F32), ag=np.array([albedo[1]], F32),
              ab=np.array([albedo[2]], F32), rough=np.array([rough], F32),
              metal=np.array([255 if metal else 0], np.uint8))
    lights = np.zeros(1, abi.LIGHT_DTYPE)
    lights["position"][0] = [wp[0], wp[1], wp[2] + light_d, 1]
    lights["color"][0] = [*color, 1]
    lights["radius"][0] = radius
    cells = np.zeros(1, abi.CELL_DTYPE)
    cells["count"] = 1
    idx = np.zeros(1, np.uint32)
    return fr, ivp, cam, ubo, lights, cells, idx, wp


def _const_maps(cd, cs, bx, by):
    spec = np.tile(np.array([*cs, 1], F32), abi.cube_chain_floats(4, 3) // 4)
    diff = np.tile(np.array([*cd, 1], F32), 6 * 2 * 2)
    lut = np.tile(np.array([bx, by], F32), 8 * 8)
    return dict(spec=spec, spec_size=4, spec_mips=3, diff=diff, diff_size=2, lut=lut, lut_w=8, lut_h=8)


@pytest.mark.parametrize("rough,metal", [(1.0, False), (1.0, True), (0.5, False)])
def test_shading_closed_form(rough, metal):
    albedo, color, d, r = (0.8, 0.4, 0.2), (1.0, 0.5, 0.25), 2.0, 4.0
    fr, ivp, cam, ubo, lights, cells, idx, wp = _one_fragment_scene(rough, metal, albedo, d, r, color)
    cd, cs, bx, by = (0.3, 0.2, 0.1), (0.6, 0.5, 0.4), 0.7, 0.05
    maps = _const_maps(cd, cs, bx, by)
    pc = abi.pbr_defaults()
    got = O.shade(fr, 1, 1, ivp, cam, ubo, pc, lights, cells, idx, maps, [0], dtype=np.float64)[0]
    m = 1.0 if metal else 0.0
    alb = np.array(albedo)
    F0 = pc.base_reflectivity * (1 - m) + alb * m
    a2 = (rough * rough) ** 2
    D = a2 / (np.pi * (1 * (a2 - 1) + 1) ** 2)            # NoH = 1
    k = (rough + 1) ** 2 / 8
    G = (1 / (1 * (1 - k) + k)) ** 2                        # NoL = NoV = 1
    F = F0                                                  # HoV = 1
    spec = D * G * F / (4 + 1e-4)
    kD = (1 - F) * (1 - m)
    att = (1 - d / r) / d ** pc.light_falloff
    rad = np.array(color) * att * pc.light_intensity
    Lo = alb * pc.albedo_multiplier + (kD * alb / np.pi + spec) * rad
    Fa = F0                                                 # NoV = 1: (1 - NoV)^5 = 0
    amb = (1 - Fa) * (1 - m) * np.array(cd) * alb + np.array(cs) * (Fa * bx + by)
    assert np.allclose(got, amb + Lo, rtol=1e-6)
    got32 = O.shade(fr, 1, 1, ivp, cam, ubo, pc, lights, cells, idx, maps, [0])[0]
    assert np.allclose(got32, got, rtol=2e-6)


def test_shading_closed_form_generic_angles():
    """Radiance.glsl / PBRHeader.glsl at a generic (N, V, L): metal = 1 and
    albedo = 1 make F0 = 1, so F == 1 and kD == 0 in both the light term and
    the ambient term, and the light's term reduces to
    D(N.H) G_L(N.L) G_V(N.V) / (4 N.V N.L + 1e-4) * rad * N.L. The normal and
    light are placed so that N.H, H.V, N.L and N.V all differ (by more than
    10 %), so a swap of N.H with H.V in D or of N.L with N.V in G or in the
    denominator changes the value by far more than the tolerance."""
    rough = 0.45
    fr, ivp, cam, ubo, lights, cells, idx, wp = _one_fragment_scene(rough, True, (1.0, 1.0, 1.0), 2.0, 6.0,
                                                                     (1.0, 0.5, 0.25))
    n = np.array([-0.6, -0.6, 1.0]); n /= np.linalg.norm(n)
    fr["nx"][:], fr["ny"][:], fr["nz"][:] = n.astype(F32)
    n = np.array([fr["nx"][0], fr["ny"][0], fr["nz"][0]], np.float64)
    ldir = np.array([0.3, 0.4, 0.8]); ldir /= np.linalg.norm(ldir)
    lpos = (wp + 2.0 * ldir).astype(F32)
    lights["position"][0, :3] = lpos
    cd, cs, bx, by = (0.3, 0.2, 0.1), (0.6, 0.5, 0.4), 0.7, 0.05
    maps = _const_maps(cd, cs, bx, by)
    pc = abi.pbr_defaults()
    got = O.shade(fr, 1, 1, ivp, cam, ubo, pc, lights, cells, idx, maps, [0], dtype=np.float64)[0]
    cp = np.array(cam.position[:3], np.float64)
    V = (cp - wp) / np.linalg.norm(cp - wp)
    d = lpos.astype(np.float64) - wp
    dist = np.linalg.norm(d)
    L = d / dist
    Hh = (V + L) / np.linalg.norm(V + L)
    NoH, NoL, NoV, HoV = n @ Hh, n @ L, n @ V, Hh @ V
    vals = sorted([NoH, NoL, NoV, HoV])
    assert min(b - a for a, b in zip(vals, vals[1:])) > 0.1 and vals[0] > 0.1
    a2 = (rough * rough) ** 2
    D = a2 / (np.pi * (NoH * NoH * (a2 - 1) + 1) ** 2)
    k = (rough + 1) ** 2 / 8
    G = (NoL / (NoL * (1 - k) + k)) * (NoV / (NoV * (1 - k) + k))
    spec = D * G / (4 * NoV * NoL + 1e-4)
    att = max(1 - dist / lights["radius"][0], 0) / dist ** pc.light_falloff
    rad = np.array([1.0, 0.5, 0.25]) * att * pc.light_intensity
    Lo = 1.0 * pc.albedo_multiplier + spec * rad * NoL
    amb = np.array(cs) * (bx + by)  # F = 1, kD = 0: only the prefiltered term
    assert np.allclose(got, amb + Lo, rtol=1e-7), (got, amb + Lo)
    # the mistakes this pins: each moves the light term by more than 5 %
    D_swap = a2 / (np.pi * (HoV * HoV * (a2 - 1) + 1) ** 2)
    spec_swap = D * G / (4 * NoV * NoV + 1e-4)
    for wrong in (D_swap * G / (4 * NoV * NoL + 1e-4), spec_swap):
        assert abs(wrong - spec) > 0.05 * spec
    got32 = O.shade(fr, 1, 1, ivp, cam, ubo, pc, lights, cells, idx, maps, [0])[0]
    assert np.allclose(got32, got, rtol=2e-6)


def test_shading_light_beyond_radius_contributes_nothing():
    fr, ivp, cam, ubo, lights, cells, idx, _ = _one_fragment_scene(0.5, False, (0.5, 0.5, 0.5), 3.0, 2.5, (1, 1, 1))
    maps = _const_maps((0.1, 0.1, 0.1), (0.2, 0.2, 0.2), 0.5, 0.1)
    pc = abi.pbr_defaults()
    with_l = O.shade(fr, 1, 1, ivp, cam, ubo, pc, lights, cells, idx, maps, [0], dtype=np.float64)
    cells0 = cells.copy()
    cells0["count"] = 0
    without = O.shade(fr, 1, 1, ivp, cam, ubo, pc, lights, cells0, idx, maps, [0], dtype=np.float64)
    assert np.array_equal(with_l, without)


def test_lut_repeat_wrap_in_shading_sampler():
    # NoV = 1 samples the LUT at s = 1: REPEAT blends the last and first columns
    fr, ivp, cam, ubo, lights, cells, idx, _ = _one_fragment_scene(0.5, False, (0.5, 0.5, 0.5), 1.0, 0.5, (1, 1, 1))
    cells["count"] = 0
    pc = abi.pbr_defaults()
    maps = _const_maps((0, 0, 0), (1, 1, 1), 0.0, 0.0)
    lut = np.zeros((8, 8, 2), F32)
    lut[:, 0, 1] = 1.0    # first column: by = 1
    maps["lut"] = lut.reshape(-1)
    got = O.shade(fr, 1, 1, ivp, cam, ubo, pc, lights, cells, idx, maps, [0], dtype=np.float64)[0]
    lo = np.array(fr["ar"].tolist() * 3) * 0  # albedo term is zero-weighted below
    # ambient specular = pre * (F*bx + by) with by = 0.5 (half of column 0)
    base = fr["ar"][0] * pc.albedo_multiplier
    assert np.allclose(got - base, 0.5, atol=1e-5)


def test_cluster_lookup_slices():
    # LinearDepth: z = 0 -> near, z = 1 -> far; the slice of a depth in the
    # middle of slice k is k
    desc, cam, ubo, ivp = camera(64, 32, (2, 1, 24))
    Z = 24
    for k in range(Z):
        lin = NEAR * (FAR / NEAR) ** ((k + 0.5) / Z)
        z = FAR * (lin - NEAR) / (lin * (FAR - NEAR))
        l = (NEAR * FAR) / (FAR + z * (NEAR - FAR))
        assert l == pytest.approx(lin, rel=1e-9)
        assert int(max(math.log2(l) * ubo.slice_scaling + ubo.slice_bias, 0)) == k
