"""The C-ABI library on the CPU: it loads, exports every entry point the
headers declare, keeps the reference's struct layouts, and its host-side
logic (camera/UBO construction, seeded generators, size helpers, error
paths that need no GPU) behaves. No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from hellovulkan_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("include/hvx.h", "include/hvx_scene.h"):
        text = open(os.path.join(ROOT, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(hvx_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = declared_functions()
    assert len(names) >= 19
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in abi.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_struct_layouts_match_reference():
    # LightData std430 48 B, AABB 32 B, LightCell 8 B, ClusterForwardUBO 176 B,
    # CameraUBO 160 B, PushConstPBR 20 B
    assert C.sizeof(abi.Light) == 48 and C.sizeof(abi.AABB) == 32 and C.sizeof(abi.LightCell) == 8
    assert C.sizeof(abi.CFUBO) == 176 and C.sizeof(abi.CameraUBO) == 160 and C.sizeof(abi.PBRPC) == 20
    assert C.sizeof(abi.Probe) == 32 and C.sizeof(abi.View) == 232 and C.sizeof(abi.BatchTuning) == 24
    assert abi.View.n_lights.offset == 192 and abi.View.indices_capacity.offset == 216
    assert abi.CFUBO.screen_size.offset == 128 and abi.CFUBO.slice_count_x.offset == 152
    assert abi.Light.radius.offset == 32


def test_status_strings_and_version():
    lib = abi.load_library()
    assert lib.hvx_abi_version() == 1
    assert lib.hvx_status_str(abi.HVX_OK) == b"ok"
    assert lib.hvx_status_str(abi.HVX_ERR_CAPACITY) == b"output buffer too small"
    assert lib.hvx_status_str(12345) == b"unknown status"


def test_mip_count_and_chain_sizes():
    # Utility::MipMapCount
    assert [abi.mip_count(s) for s in (1, 2, 32, 128, 512, 1024, 1000)] == [1, 2, 6, 8, 10, 11, 10]
    assert abi.cube_chain_floats(4, 3) == 6 * 4 * (16 + 4 + 1)
    assert abi.cube_level_offset_floats(512, 1) == 6 * 512 * 512 * 4


def test_null_context_is_rejected():
    lib = abi.load_library()
    ubo = abi.CFUBO()
    assert lib.hvx_cf_build_aabbs(None, C.byref(ubo), None, None) == abi.HVX_ERR_NULL
    assert lib.hvx_ctx_create(0, None) == abi.HVX_ERR_NULL


def test_camera_build_validation():
    lib = abi.load_library()
    d = abi.CameraDesc()
    d.fovy, d.znear, d.zfar, d.width, d.height = 0.8, 0.1, 100.0, 16, 16
    d.target[2] = -1
    d.slice_count_x = d.slice_count_y = d.slice_count_z = 2
    assert lib.hvx_camera_build(C.byref(d), None, None, None) == abi.HVX_OK
    d.znear = 200.0
    assert lib.hvx_camera_build(C.byref(d), None, None, None) == abi.HVX_ERR_INVALID_ARG


def test_generators_deterministic_and_in_range():
    a = abi.scene_lights(5, 1000)
    b = abi.scene_lights(5, 1000)
    assert a.tobytes() == b.tobytes()
    assert abi.scene_lights(6, 1000).tobytes() != a.tobytes()
    p = a["position"]
    assert p[:, 0].min() >= -15 and p[:, 0].max() <= 15 and p[:, 1].min() >= 0 and p[:, 1].max() <= 12
    assert p[:, 2].min() >= -7 and p[:, 2].max() <= 7 and np.all(p[:, 3] == 1)
    assert a["radius"].min() >= 0.5 and a["radius"].max() <= 3.0
    assert a["color"][:, :3].min() >= 0 and a["color"][:, :3].max() <= 1
    # the first elements of a longer stream are the shorter stream (counter-based)
    assert abi.scene_lights(5, 10).tobytes() == a[:10].tobytes()


def test_fragments_distribution():
    desc, cam, ubo, ivp = abi.camera_build((0, 1, 6), (0, 2.5, 0), 0.785398, 0.1, 100, 160, 100, (4, 3, 6))
    f = abi.scene_fragments(3, desc, ivp, threads=2)
    g = abi.scene_fragments(3, desc, ivp, threads=1)
    for k in f:
        assert f[k].tobytes() == g[k].tobytes()
    z = f["depth"].astype(np.float64)
    assert z.min() >= 0 and z.max() < 1
    lin = (0.1 * 100) / (100 + z * (0.1 - 100))
    u = np.log(lin / 0.1) / np.log(1000)          # log-uniform depth -> u uniform
    assert abs(u.mean() - 0.5) < 0.02 and abs(u.std() - np.sqrt(1 / 12)) < 0.02
    n = np.stack([f["nx"], f["ny"], f["nz"]], 1).astype(np.float64)
    assert np.allclose(np.linalg.norm(n, axis=1), 1, atol=1e-6)
    # normals face the camera
    W, H = 160, 100
    px, py = np.arange(W * H) % W, np.arange(W * H) // W
    clip = np.stack([((px + .5) / W) * 2 - 1, ((py + .5) / H) * 2 - 1, z, np.ones_like(z)], 0)
    IVP = np.array(ivp, np.float64).reshape(4, 4).T
    wh = IVP @ clip
    wp = (wh[:3] / wh[3]).T
    v = np.array([0, 1, 6.0]) - wp
    assert np.all(np.sum(n * v, 1) >= -1e-6)
    assert 0.27 < f["metal"].mean() / 255 < 0.33 and set(np.unique(f["metal"])) <= {0, 255}
    assert f["rough"].min() >= 0.05 and f["rough"].max() <= 1


def test_equirect_generator():
    e = abi.scene_equirect(77, 256, 128, threads=4)
    assert e.shape == (128, 256, 4) and np.all(e[..., 3] == 1) and np.all(e[..., :3] >= 0)
    assert e[..., :3].max() >= 1e4          # the sun disk
    assert np.all(np.isfinite(e))
    assert abi.scene_equirect(77, 256, 128, threads=1).tobytes() == e.tobytes()
