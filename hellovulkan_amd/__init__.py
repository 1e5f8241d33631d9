"""hellovulkan_amd — MI355X-native clustered-forward + image-based-lighting
compute path (HelloVulkan's AABB generation, light culling, clustered
Cook-Torrance shading, BRDF LUT, equirect-to-cube and cube filtering stages)
as hand-written HIP kernels for gfx950 behind the C ABI of include/hvx.h.

`hellovulkan_amd.abi` binds that ABI with ctypes; `hellovulkan_amd.build`
compiles it in-tree.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
