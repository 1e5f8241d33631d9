"""Python front end of the CPU oracle (oracle/_build/libhvx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker. The product never imports it.
Host arrays only (numpy); each call mirrors one stage of the path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from hellovulkan_amd.abi import CameraUBO, CFUBO, FragSoA, IBLMaps, PBRPC, LIGHT_DTYPE, CELL_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libhvx_oracle.so")

P = C.POINTER
vp = C.c_void_p
u32 = C.c_uint32

_SIGS = [
    ("hvxo_math_f32", None, [C.c_int, vp, vp, vp, u32]),
    ("hvxo_math_f64", None, [C.c_int, vp, vp, vp, u32]),
    ("hvxo_hammersley_f32", None, [u32, vp]),
    ("hvxo_ggx_sample_f32", None, [C.c_float, C.c_float, vp, C.c_float, vp]),
    ("hvxo_cube_face_st_f32", C.c_int, [vp, vp]),
    ("hvxo_cube_edge_texel", None, [C.c_int, C.c_int, C.c_int, C.c_int, vp]),
    ("hvxo_sample_cube_f32", None, [vp, u32, vp, u32, vp]),
    ("hvxo_texel_dir_f32", None, [C.c_int, u32, u32, u32, vp]),
    ("hvxo_cf_build_aabbs_f32", None, [P(CFUBO), vp]),
    ("hvxo_cf_build_aabbs_f64", None, [P(CFUBO), vp]),
    ("hvxo_cf_cull_lights", None, [P(CFUBO), vp, vp, u32, u32, vp, vp, vp, vp, C.c_int]),
    ("hvxo_cf_cull_lights_batch", None, [P(CFUBO), vp, vp, u32, u32, u32, u32, vp, vp, C.c_uint64, vp, C.c_int]),
    ("hvxo_cf_shade_f32", None, [P(FragSoA), P(CameraUBO), P(CFUBO), P(PBRPC), vp, vp, vp, P(IBLMaps), vp, u32, vp,
                                 C.c_int]),
    ("hvxo_cf_shade_f64", None, [P(FragSoA), P(CameraUBO), P(CFUBO), P(PBRPC), vp, vp, vp, P(IBLMaps), vp, u32, vp,
                                 C.c_int]),
    ("hvxo_ibl_brdf_lut_f32", None, [u32, u32, u32, vp, u32, vp, C.c_int]),
    ("hvxo_ibl_brdf_lut_f64", None, [u32, u32, u32, vp, u32, vp, C.c_int]),
    ("hvxo_ibl_equirect_to_cube_f32", None, [vp, u32, u32, u32, vp, u32, vp, C.c_int]),
    ("hvxo_ibl_equirect_to_cube_f64", None, [vp, u32, u32, u32, vp, u32, vp, C.c_int]),
    ("hvxo_ibl_cube_mips_f32", None, [vp, u32, u32]),
    ("hvxo_ibl_diffuse_f32", None, [vp, u32, u32, vp, u32, vp, C.c_int]),
    ("hvxo_ibl_diffuse_f64", None, [vp, u32, u32, vp, u32, vp, C.c_int]),
    ("hvxo_ibl_specular_f32", None, [vp, u32, u32, u32, u32, u32, C.c_int, vp, u32, vp, C.c_int]),
    ("hvxo_ibl_specular_f64", None, [vp, u32, u32, u32, u32, u32, C.c_int, vp, u32, vp, C.c_int]),
    ("hvxo_contract_f64", None, [C.c_int, vp, u32, u32, u32, u32, vp, u32, vp, u32, C.c_int]),
    ("hvxo_hot_threshold", C.c_float, [vp, u32, u32]),
    ("hvxo_weight_integral", C.c_double, [C.c_int, C.c_double]),
    ("hvxo_ideal_f64", None, [C.c_int, vp, u32, C.c_double, u32, vp, u32, vp, C.c_int]),
    ("hvxo_contract_levels", C.c_int, [C.c_int, u32, u32, u32, vp, vp]),
    ("hvxo_mlp_forward_f64", None, [vp, vp, u32, C.c_double, C.c_int, vp]),
    ("hvxo_mlp_sincos", None, [vp, u32, vp]),
    ("hvxo_tonemap_f32", None, [vp, u32, C.c_int, C.c_int, vp]),
    ("hvxo_tonemap_f64", None, [vp, u32, C.c_int, C.c_int, vp]),
    ("hvxo_mlp_gradient_f64", None, [vp, vp, vp, u32, C.c_double, C.c_int, vp, vp]),
]

_lib = None

MATH_FN = {"log2": 0, "exp2": 1, "pow": 2, "sin": 3, "cos": 4, "atan2": 5, "acos": 6, "pow5": 7, "sqrt": 8,
           "atan": 9}


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run hellovulkan_amd/build.py")
        l = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype)


def _threads(threads):
    return threads if threads else (os.cpu_count() or 1)


def math_f32(fn: str, a, b=None) -> np.ndarray:
    a = _c(a, np.float32)
    b = a if b is None else _c(b, np.float32)
    out = np.empty_like(a)
    lib().hvxo_math_f32(MATH_FN[fn], a.ctypes.data, b.ctypes.data, out.ctypes.data, a.size)
    return out


def math_f64(fn: str, a, b=None) -> np.ndarray:
    a = _c(a, np.float64)
    b = a if b is None else _c(b, np.float64)
    out = np.empty_like(a)
    lib().hvxo_math_f64(MATH_FN[fn], a.ctypes.data, b.ctypes.data, out.ctypes.data, a.size)
    return out


def hammersley(n: int) -> np.ndarray:
    out = np.empty((n, 2), np.float32)
    lib().hvxo_hammersley_f32(n, out.ctypes.data)
    return out


def ggx_sample(xi, N, rough) -> np.ndarray:
    n = _c(N, np.float32)
    out = np.empty(3, np.float32)
    lib().hvxo_ggx_sample_f32(float(xi[0]), float(xi[1]), n.ctypes.data, float(rough), out.ctypes.data)
    return out


def cube_face_st(d):
    d = _c(d, np.float32)
    st = np.empty(2, np.float32)
    f = lib().hvxo_cube_face_st_f32(d.ctypes.data, st.ctypes.data)
    return f, st


def cube_edge_texel(face, i, j, S):
    out = np.empty(3, np.int32)
    lib().hvxo_cube_edge_texel(face, i, j, S, out.ctypes.data)
    return tuple(int(v) for v in out)


def sample_cube(level: np.ndarray, S: int, dirs) -> np.ndarray:
    level = _c(level, np.float32)
    dirs = _c(dirs, np.float32).reshape(-1, 3)
    out = np.empty((dirs.shape[0], 3), np.float32)
    lib().hvxo_sample_cube_f32(level.ctypes.data, S, dirs.ctypes.data, dirs.shape[0], out.ctypes.data)
    return out


def texel_dir(face, x, y, D) -> np.ndarray:
    out = np.empty(3, np.float32)
    lib().hvxo_texel_dir_f32(face, x, y, D, out.ctypes.data)
    return out


def build_aabbs(ubo: CFUBO, dtype=np.float32) -> np.ndarray:
    Cn = ubo.slice_count_x * ubo.slice_count_y * ubo.slice_count_z
    out = np.empty((Cn, 8), dtype)
    fn = lib().hvxo_cf_build_aabbs_f32 if dtype == np.float32 else lib().hvxo_cf_build_aabbs_f64
    fn(C.byref(ubo), out.ctypes.data)
    return out


def cull_lights(ubo: CFUBO, aabbs: np.ndarray, lights: np.ndarray, cap: int, threads=None):
    """Returns (cells[C] CELL_DTYPE, indices[total] u32, total, overflow_count)."""
    Cn = ubo.slice_count_x * ubo.slice_count_y * ubo.slice_count_z
    aabbs = _c(aabbs, np.float32).reshape(Cn, 8)
    lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
    cells = np.zeros(Cn, CELL_DTYPE)
    idx = np.zeros(Cn * cap, np.uint32)
    tot = np.zeros(1, np.uint32)
    ov = np.zeros(1, np.uint32)
    lib().hvxo_cf_cull_lights(C.byref(ubo), aabbs.ctypes.data, lights.ctypes.data, lights.shape[0], cap,
                              cells.ctypes.data, idx.ctypes.data, tot.ctypes.data, ov.ctypes.data,
                              _threads(threads))
    return cells, idx[: int(tot[0])].copy(), int(tot[0]), int(ov[0])


def cull_lights_batch(ubo: CFUBO, aabbs: np.ndarray, lights: np.ndarray, cap: int, flags: int = 0, batch: int = 576,
                      capacity=None, threads=None):
    """LightCullingBatch.comp restated (flags: hvx.h HVX_CULL_BATCH_REFERENCE_*).
    Returns (cells[C] CELL_DTYPE, indices[min(total, capacity)] u32, total)."""
    Cn = ubo.slice_count_x * ubo.slice_count_y * ubo.slice_count_z
    aabbs = _c(aabbs, np.float32).reshape(Cn, 8)
    lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
    if capacity is None:
        capacity = Cn * (max(cap, lights.shape[0] + batch) if flags & 1 else cap)
    cells = np.zeros(Cn, CELL_DTYPE)
    idx = np.zeros(max(capacity, 1), np.uint32)
    tot = np.zeros(1, np.uint32)
    lib().hvxo_cf_cull_lights_batch(C.byref(ubo), aabbs.ctypes.data, lights.ctypes.data, lights.shape[0], cap, flags,
                                    batch, cells.ctypes.data, idx.ctypes.data, capacity, tot.ctypes.data,
                                    _threads(threads))
    t = int(tot[0])
    return cells, idx[: min(t, capacity)].copy(), t


def shade(frags: dict, width, height, ivp, cam: CameraUBO, ubo: CFUBO, pc: PBRPC, lights, cells, indices,
          maps: dict, ids, dtype=np.float32, threads=None) -> np.ndarray:
    """Shades fragments `ids` on the host. frags: host planes as
    hellovulkan_amd.abi.scene_fragments returns; maps: host arrays with keys
    spec, spec_size, spec_mips, diff, diff_size, lut, lut_w, lut_h."""
    from hellovulkan_amd.abi import frag_soa, ibl_maps
    keep = {k: np.ascontiguousarray(v) for k, v in frags.items()}
    fr = frag_soa(keep, width, height, ivp)
    spec, diff, lut = (np.ascontiguousarray(maps[k], np.float32) for k in ("spec", "diff", "lut"))
    m = ibl_maps(spec, maps["spec_size"], maps["spec_mips"], diff, maps["diff_size"], lut, maps["lut_w"],
                 maps["lut_h"])
    lights = np.ascontiguousarray(lights, LIGHT_DTYPE)
    cells = np.ascontiguousarray(cells, CELL_DTYPE)
    indices = _c(indices, np.uint32)
    if indices.size == 0:
        indices = np.zeros(1, np.uint32)
    ids = _c(ids, np.uint32)
    out = np.empty((ids.size, 3), dtype)
    fn = lib().hvxo_cf_shade_f32 if dtype == np.float32 else lib().hvxo_cf_shade_f64
    fn(C.byref(fr), C.byref(cam), C.byref(ubo), C.byref(pc), lights.ctypes.data, cells.ctypes.data,
       indices.ctypes.data, C.byref(m), ids.ctypes.data, ids.size, out.ctypes.data, _threads(threads))
    return out


def brdf_lut(W, H, n, ids=None, dtype=np.float32, threads=None) -> np.ndarray:
    ids = np.arange(W * H, dtype=np.uint32) if ids is None else _c(ids, np.uint32)
    out = np.empty((ids.size, 2), dtype)
    fn = lib().hvxo_ibl_brdf_lut_f32 if dtype == np.float32 else lib().hvxo_ibl_brdf_lut_f64
    fn(W, H, n, ids.ctypes.data, ids.size, out.ctypes.data, _threads(threads))
    return out


def equirect_to_cube(eq: np.ndarray, S: int, ids=None, dtype=np.float32, threads=None) -> np.ndarray:
    eq = _c(eq, np.float32)
    h, w = eq.shape[0], eq.shape[1]
    ids = np.arange(6 * S * S, dtype=np.uint32) if ids is None else _c(ids, np.uint32)
    out = np.empty((ids.size, 4), dtype)
    fn = lib().hvxo_ibl_equirect_to_cube_f32 if dtype == np.float32 else lib().hvxo_ibl_equirect_to_cube_f64
    fn(eq.ctypes.data, w, h, S, ids.ctypes.data, ids.size, out.ctypes.data, _threads(threads))
    return out


def cube_mips(chain: np.ndarray, S: int, levels: int) -> np.ndarray:
    chain = np.array(chain, np.float32, copy=True)
    lib().hvxo_ibl_cube_mips_f32(chain.ctypes.data, S, levels)
    return chain


def diffuse(cube0: np.ndarray, S: int, Sd: int, ids=None, dtype=np.float32, threads=None) -> np.ndarray:
    cube0 = _c(cube0, np.float32)
    ids = np.arange(6 * Sd * Sd, dtype=np.uint32) if ids is None else _c(ids, np.uint32)
    out = np.empty((ids.size, 4), dtype)
    fn = lib().hvxo_ibl_diffuse_f32 if dtype == np.float32 else lib().hvxo_ibl_diffuse_f64
    fn(cube0.ctypes.data, S, Sd, ids.ctypes.data, ids.size, out.ctypes.data, _threads(threads))
    return out


def specular(cube0: np.ndarray, S: int, So: int, mips: int, mip: int, n: int, ids=None, emulate=False,
             dtype=np.float32, threads=None) -> np.ndarray:
    cube0 = _c(cube0, np.float32)
    Sm = max(So >> mip, 1)
    ids = np.arange(6 * Sm * Sm, dtype=np.uint32) if ids is None else _c(ids, np.uint32)
    out = np.empty((ids.size, 4), dtype)
    fn = lib().hvxo_ibl_specular_f32 if dtype == np.float32 else lib().hvxo_ibl_specular_f64
    fn(cube0.ctypes.data, S, So, mips, mip, n, 1 if emulate else 0, ids.ctypes.data, ids.size, out.ctypes.data,
       _threads(threads))
    return out


CONTRACT_DELTA, CONTRACT_FULL, CONTRACT_SPLIT = 0, 1, 2


def contract_levels(ftype: int, S: int, So: int, mips: int) -> list[dict]:
    """Per-mip levels of the contraction filter (kind, So, Si, Sf, alpha2, t0, t1)."""
    lv = np.zeros(4 * mips, np.uint32)
    win = np.zeros(3 * mips, np.float64)
    if lib().hvxo_contract_levels(ftype, S, So, mips, lv.ctypes.data, win.ctypes.data) != mips:
        raise ValueError("invalid contraction sizes")
    return [dict(kind=int(lv[4 * m]), So=int(lv[4 * m + 1]), Si=int(lv[4 * m + 2]), Sf=int(lv[4 * m + 3]),
                 alpha2=float(win[3 * m]), t0=float(win[3 * m + 1]), t1=float(win[3 * m + 2])) for m in range(mips)]


def contract(ftype: int, level0: np.ndarray, S: int, So: int, mips: int, mip: int, ids=None,
             threads=None, max_hot: int = 0) -> np.ndarray:
    """FP64 twin of the contraction filter (RGB) at output texels `ids` of mip `mip`;
    max_hot > 0: with the bright-texel split at that many texels."""
    level0 = _c(level0[: 6 * S * S * 4], np.float32)
    Sm = max(So >> mip, 1)
    ids = np.arange(6 * Sm * Sm, dtype=np.uint32) if ids is None else _c(ids, np.uint32)
    out = np.full((ids.size, 3), np.nan, np.float64)
    lib().hvxo_contract_f64(ftype, level0.ctypes.data, S, So, mips, mip, ids.ctypes.data, ids.size, out.ctypes.data,
                            max_hot, _threads(threads))
    return out


def hot_threshold(level0: np.ndarray, S: int, max_hot: int) -> float:
    """Threshold T of the bright-texel split (+inf: no texel split off)."""
    level0 = _c(level0[: 6 * S * S * 4], np.float32)
    return float(lib().hvxo_hot_threshold(level0.ctypes.data, S, max_hot))


def weight_integral(ftype: int, alpha2: float) -> float:
    """Integral of the filter weight over the sphere (closed form / series)."""
    return float(lib().hvxo_weight_integral(ftype, alpha2))


def ideal(ftype: int, level0: np.ndarray, S: int, alpha2: float, So_mip: int, ids, threads=None) -> np.ndarray:
    """The filter integral summed over every level-0 texel (RGB), FP64."""
    level0 = _c(level0[: 6 * S * S * 4], np.float32)
    ids = _c(ids, np.uint32)
    out = np.full((ids.size, 3), np.nan, np.float64)
    lib().hvxo_ideal_f64(ftype, level0.ctypes.data, S, alpha2, So_mip, ids.ctypes.data, ids.size, out.ctypes.data,
                         _threads(threads))
    return out


MLP_PARAMS = 16771


def tonemap(rgb_planar: np.ndarray, op: int, gamma: bool = True, dtype=np.float32) -> np.ndarray:
    """Tonemap.frag of planar RGB (3 x n), FP32 (parity) or FP64."""
    rgb = _c(rgb_planar, np.float32).reshape(3, -1)
    n = rgb.shape[1]
    out = np.empty((3, n), dtype)
    fn = lib().hvxo_tonemap_f32 if dtype == np.float32 else lib().hvxo_tonemap_f64
    fn(rgb.ctypes.data, n, op, 1 if gamma else 0, out.ctypes.data)
    return out


def mlp_sincos(x: np.ndarray) -> np.ndarray:
    """The kernels' FP32 encoding of hvx_mlp.h: (n, 16, 2) sin and cos of f_k x."""
    x = _c(x, np.float32).ravel()
    out = np.empty((x.size, 16, 2), np.float32)
    lib().hvxo_mlp_sincos(x.ctypes.data, x.size, out.ctypes.data)
    return out


def mlp_forward(params: np.ndarray, inputs: np.ndarray, lod_max: float, bf16: bool = False) -> np.ndarray:
    """FP64 log-radiance (n x 3) of the radiance network at inputs (n x 4: dir, lod);
    bf16=True rounds the kernels' BF16 operands and uses their FP32 encoding."""
    params = _c(params, np.float32)
    inputs = _c(inputs, np.float32).reshape(-1, 4)
    out = np.empty((inputs.shape[0], 3), np.float64)
    lib().hvxo_mlp_forward_f64(params.ctypes.data, inputs.ctypes.data, inputs.shape[0], lod_max, 1 if bf16 else 0,
                               out.ctypes.data)
    return out


def mlp_gradient(params, inputs, targets, lod_max: float, bf16: bool = False):
    """FP64 gradient (parameter layout) and loss of the mean squared log-radiance error;
    bf16=True rounds the BF16 GEMM operands of the GPU kernels at the same points."""
    params = _c(params, np.float32)
    inputs = _c(inputs, np.float32).reshape(-1, 4)
    targets = _c(targets, np.float32).reshape(-1, 4)
    g = np.empty(MLP_PARAMS, np.float64)
    loss = np.zeros(1, np.float64)
    lib().hvxo_mlp_gradient_f64(params.ctypes.data, inputs.ctypes.data, targets.ctypes.data, inputs.shape[0], lod_max,
                                1 if bf16 else 0, g.ctypes.data, loss.ctypes.data)
    return g, float(loss[0])


def rel_err(got, ref, floor_frac=1e-6) -> float:
    """max |got - ref| / max(|ref|, floor_frac * max|ref|)."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    if ref.size == 0:
        return 0.0
    floor = floor_frac * float(np.max(np.abs(ref))) if ref.size else 0.0
    den = np.maximum(np.abs(ref), max(floor, 1e-30))
    return float(np.max(np.abs(got - ref) / den))
