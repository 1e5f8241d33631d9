"""ctypes binding of the C ABI in include/hvx.h and include/hvx_scene.h.

The library is loaded from this package directory (built in-tree by
hellovulkan_amd/build.py). torch is imported first when available so that the
HIP runtime torch ships is the one libhvx.so binds to (both carry the soname
libamdhip64.so.7): device pointers and streams from torch are then valid in
libhvx calls. A missing library raises; there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

try:  # load torch's HIP runtime before libhvx.so resolves libamdhip64.so.7
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is present in this image
    torch = None

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
# HVX_LIBRARY selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("HVX_LIBRARY") or os.path.join(PKG_DIR, "libhvx.so")

HVX_OK = 0
HVX_ERR_NULL = 1
HVX_ERR_INVALID_ARG = 2
HVX_ERR_CAPACITY = 3
HVX_ERR_ALIGN = 4
HVX_ERR_HIP = 5
HVX_ERR_UNSUPPORTED = 6
HVX_ERR_NO_DEVICE = 7
HVX_ERR_STALE_GRAPH = 8

# hvx_cf_cull_lights_batch flags
CULL_BATCH_REFERENCE_CAP = 1
CULL_BATCH_REFERENCE_TAIL = 2

OPT_SHADE_EXACT = 1
OPT_SHADE_KERNEL_EVENTS = 2
OPT_MLP_SEPARATE_LAUNCHES = 3

FILTER_DIFFUSE = 0
FILTER_SPECULAR = 1
TONEMAP_REINHARD, TONEMAP_ACES, TONEMAP_FILMIC, TONEMAP_UNCHARTED2 = 0, 1, 2, 3

f4 = C.c_float * 4
f16 = C.c_float * 16


class Light(C.Structure):
    _fields_ = [("position", f4), ("color", f4), ("radius", C.c_float), ("_pad", C.c_float * 3)]


class AABB(C.Structure):
    _fields_ = [("min_point", f4), ("max_point", f4)]


class LightCell(C.Structure):
    _fields_ = [("offset", C.c_uint32), ("count", C.c_uint32)]


class CFUBO(C.Structure):
    _fields_ = [
        ("camera_inverse_projection", f16), ("camera_view", f16), ("screen_size", C.c_float * 2),
        ("slice_scaling", C.c_float), ("slice_bias", C.c_float), ("camera_near", C.c_float),
        ("camera_far", C.c_float), ("slice_count_x", C.c_uint32), ("slice_count_y", C.c_uint32),
        ("slice_count_z", C.c_uint32), ("_pad", C.c_uint32 * 3),
    ]


class CameraUBO(C.Structure):
    _fields_ = [("projection", f16), ("view", f16), ("position", f4), ("camera_near", C.c_float),
                ("camera_far", C.c_float), ("_pad", C.c_float * 2)]


class PBRPC(C.Structure):
    _fields_ = [("light_intensity", C.c_float), ("base_reflectivity", C.c_float),
                ("max_reflection_lod", C.c_float), ("light_falloff", C.c_float),
                ("albedo_multiplier", C.c_float)]


class FragSoA(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("normal", C.c_void_p * 3), ("albedo", C.c_void_p * 3),
                ("roughness", C.c_void_p), ("metallic", C.c_void_p), ("width", C.c_uint32),
                ("height", C.c_uint32), ("inv_view_proj", f16)]


class IBLMaps(C.Structure):
    _fields_ = [("specular", C.c_void_p), ("specular_size", C.c_uint32), ("specular_mips", C.c_uint32),
                ("diffuse", C.c_void_p), ("diffuse_size", C.c_uint32), ("_pad0", C.c_uint32),
                ("brdf_lut", C.c_void_p), ("lut_width", C.c_uint32), ("lut_height", C.c_uint32)]


class FilterOpts(C.Structure):
    _fields_ = [("emulate_static_viewport", C.c_uint32), ("_reserved", C.c_uint32 * 3)]


class CameraDesc(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("target", C.c_float * 3), ("fovy", C.c_float),
                ("znear", C.c_float), ("zfar", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32),
                ("slice_count_x", C.c_uint32), ("slice_count_y", C.c_uint32), ("slice_count_z", C.c_uint32)]


class ContractMipInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("out_face", C.c_uint32), ("in_face", C.c_uint32), ("far_face", C.c_uint32),
                ("window_t0", C.c_float), ("window_t1", C.c_float), ("weight_bytes", C.c_uint64),
                ("min_input_level", C.c_uint32), ("_pad", C.c_uint32)]


class Probe(C.Structure):
    _fields_ = [("cube", C.c_void_p), ("face_size", C.c_uint32), ("out_face_size", C.c_uint32),
                ("out_mips", C.c_uint32), ("_pad", C.c_uint32), ("out_chain", C.c_void_p)]


class View(C.Structure):
    _fields_ = [("ubo", CFUBO), ("clusters", C.c_void_p), ("lights", C.c_void_p), ("n_lights", C.c_uint32),
                ("cap", C.c_uint32), ("cells", C.c_void_p), ("indices", C.c_void_p),
                ("indices_capacity", C.c_size_t), ("total", C.c_void_p)]


class BatchTuning(C.Structure):
    _fields_ = [("probe_tile", C.c_uint32), ("probe_wgs_per_cu", C.c_uint32), ("cull_shape", C.c_uint32),
                ("cull_wgs_per_cu", C.c_uint32), ("probe_ms", C.c_float), ("cull_ms", C.c_float)]


# numpy views of the array-of-struct layouts
LIGHT_DTYPE = np.dtype([("position", "<f4", 4), ("color", "<f4", 4), ("radius", "<f4"), ("_pad", "<f4", 3)])
AABB_DTYPE = np.dtype([("min_point", "<f4", 4), ("max_point", "<f4", 4)])
CELL_DTYPE = np.dtype([("offset", "<u4"), ("count", "<u4")])
assert LIGHT_DTYPE.itemsize == C.sizeof(Light) == 48
assert C.sizeof(AABB) == 32 and C.sizeof(LightCell) == 8
assert C.sizeof(CFUBO) == 176 and C.sizeof(CameraUBO) == 160 and C.sizeof(PBRPC) == 20

vp = C.c_void_p
u32 = C.c_uint32
P = C.POINTER

# (name, restype, argtypes) of every exported entry point of include/hvx.h
# and include/hvx_scene.h
SIGNATURES = [
    ("hvx_abi_version", C.c_int, []),
    ("hvx_status_str", C.c_char_p, [C.c_int]),
    ("hvx_ctx_create", C.c_int, [C.c_int, P(vp)]),
    ("hvx_ctx_destroy", C.c_int, [vp]),
    ("hvx_ctx_last_error", C.c_char_p, [vp]),
    ("hvx_ctx_set_option", C.c_int, [vp, C.c_int, C.c_int64]),
    ("hvx_ctx_get_option", C.c_int, [vp, C.c_int, P(C.c_int64)]),
    ("hvx_ctx_kernel_times", C.c_int, [vp, P(C.c_float), u32, P(u32)]),
    ("hvx_ctx_scratch_generation", C.c_uint64, [vp]),
    ("hvx_cf_build_aabbs", C.c_int, [vp, P(CFUBO), vp, vp]),
    ("hvx_cf_cull_lights", C.c_int, [vp, P(CFUBO), vp, vp, u32, u32, vp, vp, C.c_size_t, vp, vp, vp]),
    ("hvx_cf_cull_lights_noalloc", C.c_int, [vp, P(CFUBO), vp, vp, u32, u32, vp, vp, C.c_size_t, vp, vp, vp]),
    ("hvx_cf_cull_noalloc_scratch_bytes", C.c_size_t, [u32, u32, u32]),
    ("hvx_cf_cull_noalloc_scratch_limit", C.c_size_t, []),
    ("hvx_cf_cull_lights_batch", C.c_int, [vp, P(CFUBO), vp, vp, u32, u32, u32, vp, vp, C.c_size_t, vp, vp]),
    ("hvx_cf_shade", C.c_int, [vp, P(FragSoA), P(CameraUBO), P(CFUBO), P(PBRPC), vp, u32, vp, vp,
                               P(IBLMaps), vp, vp]),
    ("hvx_ibl_brdf_lut", C.c_int, [vp, u32, u32, u32, vp, vp]),
    ("hvx_ibl_equirect_to_cube", C.c_int, [vp, vp, u32, u32, vp, u32, vp]),
    ("hvx_ibl_cube_mips", C.c_int, [vp, vp, u32, u32, vp]),
    ("hvx_ibl_filter", C.c_int, [vp, C.c_int, vp, u32, vp, u32, u32, u32, P(FilterOpts), vp]),
    ("hvx_cube_chain_floats", C.c_size_t, [u32, u32]),
    ("hvx_cube_level_offset_floats", C.c_size_t, [u32, u32]),
    ("hvx_mip_count", u32, [u32]),
    ("hvx_ibl_contract_plan_create", C.c_int, [vp, C.c_int, u32, u32, u32, C.c_int, P(vp)]),
    ("hvx_ibl_contract_plan_destroy", C.c_int, [vp]),
    ("hvx_ibl_contract_plan_query", C.c_int, [vp, u32, P(ContractMipInfo)]),
    ("hvx_ibl_filter_contract", C.c_int, [vp, vp, u32, vp, vp]),
    ("hvx_ibl_contract_plan_reserve", C.c_int, [vp, u32]),
    ("hvx_ibl_filter_contract_multi", C.c_int, [vp, u32, vp, u32, vp, vp]),
    ("hvx_ibl_contract_plan_set_hot", C.c_int, [vp, u32]),
    ("hvx_ibl_contract_hot_info", C.c_int, [vp, P(C.c_float), P(u32)]),
    ("hvx_mlp_create", C.c_int, [vp, C.c_float, C.c_uint64, P(vp)]),
    ("hvx_mlp_destroy", C.c_int, [vp]),
    ("hvx_mlp_get_params", C.c_int, [vp, vp]),
    ("hvx_mlp_set_params", C.c_int, [vp, vp]),
    ("hvx_mlp_sample", C.c_int, [vp, vp, u32, u32, C.c_uint64, u32, vp, vp, vp]),
    ("hvx_mlp_gradient", C.c_int, [vp, vp, vp, u32, vp, vp, vp]),
    ("hvx_mlp_train", C.c_int, [vp, vp, vp, u32, u32, u32, C.c_float, vp, vp]),
    ("hvx_mlp_eval", C.c_int, [vp, vp, u32, vp, vp]),
    ("hvx_mlp_cell_errors", C.c_int, [vp, vp, u32, u32, u32, u32, vp, vp]),
    ("hvx_cf_shade_learned", C.c_int, [vp, P(FragSoA), P(CameraUBO), P(CFUBO), P(PBRPC), vp, u32, vp, vp,
                                       P(IBLMaps), vp, vp, u32, u32, C.c_float, vp, vp, vp]),
    ("hvx_batch_create", C.c_int, [vp, P(Probe), u32, u32, P(View), u32, P(vp)]),
    ("hvx_batch_destroy", C.c_int, [vp]),
    ("hvx_batch_set_tuning", C.c_int, [vp, P(BatchTuning)]),
    ("hvx_batch_get_tuning", C.c_int, [vp, P(BatchTuning)]),
    ("hvx_batch_autotune", C.c_int, [vp, vp]),
    ("hvx_batch_set_contract", C.c_int, [vp, C.c_int, C.c_int]),
    ("hvx_batch_run", C.c_int, [vp, vp]),
    ("hvx_batch_capture", C.c_int, [vp, vp]),
    ("hvx_batch_launch", C.c_int, [vp, vp]),
    ("hvx_post_tonemap", C.c_int, [vp, vp, u32, u32, C.c_int, C.c_int, vp, vp, vp]),
    ("hvx_hdr_decode", C.c_int, [vp, C.c_size_t, C.c_int, P(u32), P(u32), vp]),
    ("hvx_camera_build", C.c_int, [P(CameraDesc), P(CameraUBO), P(CFUBO), P(C.c_float)]),
    ("hvx_scene_lights", C.c_int, [C.c_uint64, u32, vp]),
    ("hvx_scene_fragments", C.c_int, [C.c_uint64, P(CameraDesc), P(C.c_float), vp, vp, vp, vp, vp, vp, vp, vp,
                                      vp, C.c_int]),
    ("hvx_scene_equirect", C.c_int, [C.c_uint64, u32, u32, vp, C.c_int]),
]

_lib = None


def load_library() -> C.CDLL:
    """Loads libhvx.so from the package directory; raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run hellovulkan_amd/build.py "
                               "(or __graft_entry__.build()) first")
        lib = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class HvxError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        lib = load_library()
        msg = f"{where}: {lib.hvx_status_str(status).decode()} ({status})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


def hdr_decode(data: bytes, flip_vertically: bool = True) -> np.ndarray:
    """RGBA32F image (height x width x 4) of a Radiance .hdr file's bytes."""
    lib = load_library()
    buf = np.frombuffer(data, np.uint8)
    w, h = C.c_uint32(), C.c_uint32()
    st = lib.hvx_hdr_decode(buf.ctypes.data, buf.size, int(flip_vertically), C.byref(w), C.byref(h), None)
    if st:
        raise HvxError(st, "hvx_hdr_decode")
    out = np.empty((h.value, w.value, 4), np.float32)
    st = lib.hvx_hdr_decode(buf.ctypes.data, buf.size, int(flip_vertically), C.byref(w), C.byref(h),
                            out.ctypes.data)
    if st:
        raise HvxError(st, "hvx_hdr_decode")
    return out


def camera_build(eye, target, fovy, znear, zfar, width, height, grid):
    """Returns (CameraDesc, CameraUBO, CFUBO, inv_view_proj[16] as np.float32)."""
    lib = load_library()
    d = CameraDesc()
    d.eye[:] = [float(v) for v in eye]
    d.target[:] = [float(v) for v in target]
    d.fovy, d.znear, d.zfar = float(fovy), float(znear), float(zfar)
    d.width, d.height = int(width), int(height)
    d.slice_count_x, d.slice_count_y, d.slice_count_z = (int(g) for g in grid)
    cam, ubo = CameraUBO(), CFUBO()
    ivp = np.zeros(16, np.float32)
    st = lib.hvx_camera_build(C.byref(d), C.byref(cam), C.byref(ubo), ivp.ctypes.data_as(P(C.c_float)))
    if st:
        raise HvxError(st, "hvx_camera_build")
    return d, cam, ubo, ivp


def scene_lights(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, LIGHT_DTYPE)
    st = load_library().hvx_scene_lights(C.c_uint64(seed), n, out.ctypes.data)
    if st:
        raise HvxError(st, "hvx_scene_lights")
    return out


def scene_fragments(seed: int, desc: CameraDesc, ivp: np.ndarray, threads: int = 8) -> dict:
    W, H = desc.width, desc.height
    planes = {k: np.empty(W * H, np.float32) for k in ("depth", "nx", "ny", "nz", "ar", "ag", "ab", "rough")}
    metal = np.empty(W * H, np.uint8)
    st = load_library().hvx_scene_fragments(
        C.c_uint64(seed), C.byref(desc), np.ascontiguousarray(ivp, np.float32).ctypes.data_as(P(C.c_float)),
        *(planes[k].ctypes.data for k in ("depth", "nx", "ny", "nz", "ar", "ag", "ab", "rough")),
        metal.ctypes.data, threads)
    if st:
        raise HvxError(st, "hvx_scene_fragments")
    planes["metal"] = metal
    return planes


def scene_equirect(seed: int, w: int, h: int, threads: int = 8) -> np.ndarray:
    out = np.empty((h, w, 4), np.float32)
    st = load_library().hvx_scene_equirect(C.c_uint64(seed), w, h, out.ctypes.data, threads)
    if st:
        raise HvxError(st, "hvx_scene_equirect")
    return out


def cube_chain_floats(size: int, mips: int) -> int:
    return int(load_library().hvx_cube_chain_floats(size, mips))


def cube_level_offset_floats(size: int, level: int) -> int:
    return int(load_library().hvx_cube_level_offset_floats(size, level))


def mip_count(size: int) -> int:
    return int(load_library().hvx_mip_count(size))


def _ptr(t) -> int:
    """Device (or host) address of a torch tensor / numpy array; 0 for None."""
    if t is None:
        return 0
    if hasattr(t, "data_ptr"):
        return int(t.data_ptr())
    return int(t.ctypes.data)


class Context:
    """Owns an hvx_ctx; each method is one C-ABI call on the given HIP stream
    (a torch.cuda.Stream, a raw handle, or None for torch's current stream)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.hvx_ctx_create(device, C.byref(h))
        if st:
            raise HvxError(st, "hvx_ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.hvx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, stream):
        if stream is None and torch is not None:
            return torch.cuda.current_stream(self.device).cuda_stream
        if stream is not None and hasattr(stream, "cuda_stream"):
            return stream.cuda_stream
        return stream or 0

    def _check(self, st: int, where: str):
        if st:
            raise HvxError(st, where, self.lib.hvx_ctx_last_error(self.h).decode())

    def set_option(self, option: int, value: int):
        self._check(self.lib.hvx_ctx_set_option(self.h, option, value), "hvx_ctx_set_option")

    def kernel_times(self, max_n: int = 256) -> list:
        """Durations (ms) of the shading-kernel launches recorded since the
        last call (HVX_OPT_SHADE_KERNEL_EVENTS)."""
        buf = (C.c_float * max(1, max_n))()
        n = C.c_uint32()
        self._check(self.lib.hvx_ctx_kernel_times(self.h, buf, max_n, C.byref(n)), "hvx_ctx_kernel_times")
        return [float(buf[k]) for k in range(n.value)]

    def get_option(self, option: int) -> int:
        v = C.c_int64()
        self._check(self.lib.hvx_ctx_get_option(self.h, option, C.byref(v)), "hvx_ctx_get_option")
        return int(v.value)

    def build_aabbs(self, ubo: CFUBO, clusters, stream=None):
        self._check(self.lib.hvx_cf_build_aabbs(self.h, C.byref(ubo), _ptr(clusters), self._stream(stream)),
                    "hvx_cf_build_aabbs")

    def cull_lights(self, ubo: CFUBO, clusters, lights, n_lights: int, cap: int, cells, indices,
                    indices_capacity: int, total=None, overflow=None, stream=None):
        self._check(self.lib.hvx_cf_cull_lights(self.h, C.byref(ubo), _ptr(clusters), _ptr(lights), n_lights, cap,
                                                _ptr(cells), _ptr(indices), indices_capacity, _ptr(total),
                                                _ptr(overflow), self._stream(stream)), "hvx_cf_cull_lights")

    def cull_lights_noalloc(self, ubo: CFUBO, clusters, lights, n_lights: int, cap: int, cells, indices,
                            indices_capacity: int, total=None, overflow=None, stream=None):
        """hvx_cf_cull_lights on the library's fixed scratch (allocates nothing)."""
        self._check(self.lib.hvx_cf_cull_lights_noalloc(self.h, C.byref(ubo), _ptr(clusters), _ptr(lights), n_lights,
                                                        cap, _ptr(cells), _ptr(indices), indices_capacity,
                                                        _ptr(total), _ptr(overflow), self._stream(stream)),
                    "hvx_cf_cull_lights_noalloc")

    def cull_lights_batch(self, ubo: CFUBO, clusters, lights, n_lights: int, cap: int, flags: int, cells, indices,
                          indices_capacity: int, total=None, stream=None):
        """LightCullingBatch.comp semantics (flags: CULL_BATCH_REFERENCE_*)."""
        self._check(self.lib.hvx_cf_cull_lights_batch(self.h, C.byref(ubo), _ptr(clusters), _ptr(lights), n_lights,
                                                      cap, flags, _ptr(cells), _ptr(indices), indices_capacity,
                                                      _ptr(total), self._stream(stream)), "hvx_cf_cull_lights_batch")

    @property
    def scratch_generation(self) -> int:
        return int(self.lib.hvx_ctx_scratch_generation(self.h))

    def shade(self, frags: FragSoA, cam: CameraUBO, ubo: CFUBO, pc: PBRPC, lights, n_lights: int, cells, indices,
              maps: IBLMaps, rgb, stream=None):
        self._check(self.lib.hvx_cf_shade(self.h, C.byref(frags), C.byref(cam), C.byref(ubo), C.byref(pc),
                                          _ptr(lights), n_lights, _ptr(cells), _ptr(indices), C.byref(maps),
                                          _ptr(rgb), self._stream(stream)), "hvx_cf_shade")

    def shade_learned(self, frags: FragSoA, cam: CameraUBO, ubo: CFUBO, pc: PBRPC, lights, n_lights: int, cells,
                      indices, maps: IBLMaps, mlp, cell_err, cell_face: int, lod_bins: int, tol: float, rgb,
                      served=None, stream=None):
        self._check(self.lib.hvx_cf_shade_learned(self.h, C.byref(frags), C.byref(cam), C.byref(ubo), C.byref(pc),
                                                  _ptr(lights), n_lights, _ptr(cells), _ptr(indices), C.byref(maps),
                                                  mlp.h, _ptr(cell_err), cell_face, lod_bins, tol, _ptr(rgb),
                                                  _ptr(served), self._stream(stream)), "hvx_cf_shade_learned")

    def tonemap(self, rgb, width: int, height: int, op: int = 0, gamma: bool = True, out_rgb=None, out_bgra8=None,
                stream=None):
        self._check(self.lib.hvx_post_tonemap(self.h, _ptr(rgb), width, height, op, 1 if gamma else 0, _ptr(out_rgb),
                                              _ptr(out_bgra8), self._stream(stream)), "hvx_post_tonemap")

    def brdf_lut(self, w: int, h: int, samples: int, out, stream=None):
        self._check(self.lib.hvx_ibl_brdf_lut(self.h, w, h, samples, _ptr(out), self._stream(stream)),
                    "hvx_ibl_brdf_lut")

    def equirect_to_cube(self, eq, w: int, h: int, cube, face: int, stream=None):
        self._check(self.lib.hvx_ibl_equirect_to_cube(self.h, _ptr(eq), w, h, _ptr(cube), face,
                                                      self._stream(stream)), "hvx_ibl_equirect_to_cube")

    def cube_mips(self, chain, face: int, levels: int, stream=None):
        self._check(self.lib.hvx_ibl_cube_mips(self.h, _ptr(chain), face, levels, self._stream(stream)),
                    "hvx_ibl_cube_mips")

    def filter(self, ftype: int, cube, face: int, out, out_face: int, out_mips: int, samples: int,
               emulate_static_viewport: bool = False, stream=None):
        opts = FilterOpts()
        opts.emulate_static_viewport = 1 if emulate_static_viewport else 0
        self._check(self.lib.hvx_ibl_filter(self.h, ftype, _ptr(cube), face, _ptr(out), out_face, out_mips, samples,
                                            C.byref(opts), self._stream(stream)), "hvx_ibl_filter")


PRECISION_FP16_SCALED, PRECISION_BF16_SPLIT = 0, 1
CONTRACT_DELTA, CONTRACT_FULL, CONTRACT_SPLIT = 0, 1, 2


class ContractPlan:
    """Owns an hvx_contract_plan: the precomputed weights of the contraction
    (MFMA) cube filter for one (type, face, out_face, out_mips)."""

    def __init__(self, ctx: Context, ftype: int, face: int, out_face: int, out_mips: int,
                 precision: int = PRECISION_FP16_SCALED):
        self.ctx = ctx
        self.lib = ctx.lib
        self.out_face, self.out_mips = out_face, out_mips
        h = C.c_void_p()
        ctx._check(self.lib.hvx_ibl_contract_plan_create(ctx.h, ftype, face, out_face, out_mips, precision,
                                                         C.byref(h)), "hvx_ibl_contract_plan_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.hvx_ibl_contract_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self, mip: int) -> ContractMipInfo:
        i = ContractMipInfo()
        self.ctx._check(self.lib.hvx_ibl_contract_plan_query(self.h, mip, C.byref(i)), "hvx_ibl_contract_plan_query")
        return i

    def reserve(self, max_probes: int):
        self.ctx._check(self.lib.hvx_ibl_contract_plan_reserve(self.h, max_probes), "hvx_ibl_contract_plan_reserve")

    def filter_multi(self, n: int, d_chain_ptrs, chain_levels: int, d_out_ptrs, stream=None):
        """d_chain_ptrs / d_out_ptrs: device int64 tensors of n data pointers."""
        self.ctx._check(self.lib.hvx_ibl_filter_contract_multi(self.h, n, _ptr(d_chain_ptrs), chain_levels,
                                                                _ptr(d_out_ptrs), self.ctx._stream(stream)),
                        "hvx_ibl_filter_contract_multi")

    def set_hot(self, max_hot: int):
        """Bright-texel split for single-probe calls (0: off)."""
        self.ctx._check(self.lib.hvx_ibl_contract_plan_set_hot(self.h, max_hot), "hvx_ibl_contract_plan_set_hot")

    def hot_info(self):
        """(threshold, hot texels) of the last split call."""
        t, n = C.c_float(), C.c_uint32()
        self.ctx._check(self.lib.hvx_ibl_contract_hot_info(self.h, C.byref(t), C.byref(n)), "hvx_ibl_contract_hot_info")
        return t.value, n.value

    def filter(self, chain, chain_levels: int, out, stream=None):
        self.ctx._check(self.lib.hvx_ibl_filter_contract(self.h, _ptr(chain), chain_levels, _ptr(out),
                                                         self.ctx._stream(stream)), "hvx_ibl_filter_contract")


MLP_PARAMS = 16771


class Mlp:
    """Owns an hvx_mlp: the per-environment radiance network."""

    def __init__(self, ctx: Context, lod_max: float, seed: int = 1):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        ctx._check(self.lib.hvx_mlp_create(ctx.h, lod_max, seed, C.byref(h)), "hvx_mlp_create")
        self.h = h
        self.lod_max = lod_max

    def close(self):
        if self.h:
            self.lib.hvx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.empty(MLP_PARAMS, np.float32)
        self.ctx._check(self.lib.hvx_mlp_get_params(self.h, p.ctypes.data), "hvx_mlp_get_params")
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_PARAMS
        self.ctx._check(self.lib.hvx_mlp_set_params(self.h, p.ctypes.data), "hvx_mlp_set_params")

    def sample(self, chain, face: int, mips: int, seed: int, n: int, inputs, targets, stream=None):
        self.ctx._check(self.lib.hvx_mlp_sample(self.h, _ptr(chain), face, mips, seed, n, _ptr(inputs), _ptr(targets),
                                                self.ctx._stream(stream)), "hvx_mlp_sample")

    def gradient(self, inputs, targets, batch: int, grads, loss=None, stream=None):
        self.ctx._check(self.lib.hvx_mlp_gradient(self.h, _ptr(inputs), _ptr(targets), batch, _ptr(grads), _ptr(loss),
                                                  self.ctx._stream(stream)), "hvx_mlp_gradient")

    def train(self, inputs, targets, n: int, batch: int, steps: int, lr: float, loss_curve=None, stream=None):
        self.ctx._check(self.lib.hvx_mlp_train(self.h, _ptr(inputs), _ptr(targets), n, batch, steps, lr,
                                               _ptr(loss_curve), self.ctx._stream(stream)), "hvx_mlp_train")

    def cell_errors(self, chain, face: int, mips: int, cell_face: int, lod_bins: int, err, stream=None):
        self.ctx._check(self.lib.hvx_mlp_cell_errors(self.h, _ptr(chain), face, mips, cell_face, lod_bins, _ptr(err),
                                                     self.ctx._stream(stream)), "hvx_mlp_cell_errors")

    def eval(self, inputs, n: int, rgb, stream=None):
        self.ctx._check(self.lib.hvx_mlp_eval(self.h, _ptr(inputs), n, _ptr(rgb), self.ctx._stream(stream)),
                        "hvx_mlp_eval")


class Batch:
    """Owns an hvx_batch: probes (GGX prefilter) and views (light culling) run
    by the batched persistent kernels. `probes` is a list of Probe, `views` a
    list of View; the referenced device buffers must outlive the batch."""

    def __init__(self, ctx: Context, probes, sample_count: int, views):
        self.ctx = ctx
        self.lib = ctx.lib
        pa = (Probe * max(len(probes), 1))(*probes)
        va = (View * max(len(views), 1))(*views)
        h = C.c_void_p()
        ctx._check(self.lib.hvx_batch_create(ctx.h, pa, len(probes), sample_count, va, len(views), C.byref(h)),
                   "hvx_batch_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.hvx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def tuning(self) -> BatchTuning:
        t = BatchTuning()
        self.ctx._check(self.lib.hvx_batch_get_tuning(self.h, C.byref(t)), "hvx_batch_get_tuning")
        return t

    def set_tuning(self, probe_tile: int, probe_wgs_per_cu: int, cull_shape: int, cull_wgs_per_cu: int):
        t = BatchTuning(probe_tile, probe_wgs_per_cu, cull_shape, cull_wgs_per_cu, 0.0, 0.0)
        self.ctx._check(self.lib.hvx_batch_set_tuning(self.h, C.byref(t)), "hvx_batch_set_tuning")

    def autotune(self, stream=None) -> BatchTuning:
        self.ctx._check(self.lib.hvx_batch_autotune(self.h, self.ctx._stream(stream)), "hvx_batch_autotune")
        return self.tuning

    def set_contract(self, enable: bool, precision: int = 0):
        """Probes through the MFMA contraction filter (plans per size class) or back."""
        self.ctx._check(self.lib.hvx_batch_set_contract(self.h, 1 if enable else 0, precision),
                        "hvx_batch_set_contract")

    def run(self, stream=None):
        self.ctx._check(self.lib.hvx_batch_run(self.h, self.ctx._stream(stream)), "hvx_batch_run")

    def capture(self, stream):
        self.ctx._check(self.lib.hvx_batch_capture(self.h, self.ctx._stream(stream)), "hvx_batch_capture")

    def launch(self, stream):
        self.ctx._check(self.lib.hvx_batch_launch(self.h, self.ctx._stream(stream)), "hvx_batch_launch")


def probe(cube, face_size: int, out_chain, out_face_size: int, out_mips: int) -> Probe:
    return Probe(_ptr(cube), face_size, out_face_size, out_mips, 0, _ptr(out_chain))


def view(ubo: CFUBO, clusters, lights, n_lights: int, cap: int, cells, indices, indices_capacity: int,
         total=None) -> View:
    v = View()
    v.ubo = ubo
    v.clusters, v.lights, v.n_lights, v.cap = _ptr(clusters), _ptr(lights), n_lights, cap
    v.cells, v.indices, v.indices_capacity, v.total = _ptr(cells), _ptr(indices), indices_capacity, _ptr(total)
    return v


def frag_soa(planes: dict, width: int, height: int, ivp) -> FragSoA:
    """FragSoA over a dict of per-plane tensors/arrays (keys as scene_fragments)."""
    fr = FragSoA()
    fr.depth = _ptr(planes["depth"])
    fr.normal[0], fr.normal[1], fr.normal[2] = (_ptr(planes[k]) for k in ("nx", "ny", "nz"))
    fr.albedo[0], fr.albedo[1], fr.albedo[2] = (_ptr(planes[k]) for k in ("ar", "ag", "ab"))
    fr.roughness = _ptr(planes["rough"])
    fr.metallic = _ptr(planes["metal"])
    fr.width, fr.height = width, height
    fr.inv_view_proj[:] = [float(v) for v in np.asarray(ivp, np.float32)]
    return fr


def ibl_maps(spec, spec_size, spec_mips, diff, diff_size, lut, lut_w, lut_h) -> IBLMaps:
    m = IBLMaps()
    m.specular, m.specular_size, m.specular_mips = _ptr(spec), spec_size, spec_mips
    m.diffuse, m.diffuse_size = _ptr(diff), diff_size
    m.brdf_lut, m.lut_width, m.lut_height = _ptr(lut), lut_w, lut_h
    return m


def pbr_defaults() -> PBRPC:
    """Effective PushConstPBR of the clustered-forward app: UIData.h:22-29
    defaults with albedoMultipler = 0.01 (AppPBRClusterForward.cpp:21)."""
    pc = PBRPC()
    pc.light_intensity = 1.75
    pc.base_reflectivity = 0.01
    pc.max_reflection_lod = 1.0
    pc.light_falloff = 0.1
    pc.albedo_multiplier = 0.01
    return pc
