"""Writes tests/golden/golden.npz: outputs of the CPU oracle on small seeded
inputs (the library's host-side scene generators), kept as regression
fixtures for the oracle (tests/test_golden.py) and for the GPU kernels
(tests/test_gpu_golden.py). The reference ships no golden data for this
path; these fixtures record this build's oracle, which the other oracle
tests pin to closed forms, invariants and brute force.

Run from the repository root after building:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from hellovulkan_amd import abi  # noqa: E402
from oracle import hvx_oracle as O  # noqa: E402
from tests.helpers import SMALL_FRAME, camera  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")

# small frame for the shading fixture
FW, FH, FGRID, FNL, FCAP = 64, 40, (4, 4, 8), 256, 32
# IBL inputs: equirect 2S x S -> cube S
ENV_SEED, S, SD, SO, MIPS, NSAMP, LUT = 11, 16, 4, 8, 4, 64, 32
CS, CSO, CMIPS = 32, 16, 5   # contraction twin sizes


def index_digest(idx: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(idx, np.uint32).tobytes()).hexdigest()


def ibl_inputs():
    """Equirect (S x 2S), its cube (level 0), the irradiance map, the
    prefiltered chain and the BRDF LUT, all from the FP32 oracle."""
    eq = abi.scene_equirect(ENV_SEED, 2 * S, S, threads=4).reshape(S, 2 * S, 4)
    cube0 = O.equirect_to_cube(eq, S).reshape(-1)
    diff = O.diffuse(cube0, S, SD).reshape(-1)
    spec = np.concatenate([O.specular(cube0, S, SO, MIPS, m, NSAMP).reshape(-1) for m in range(MIPS)])
    lut = O.brdf_lut(LUT, LUT, NSAMP).reshape(-1)
    return eq, cube0, diff, spec, lut


def compute() -> dict:
    g = {}
    # light culling of the 2560x1600 frame (16x9x24 clusters, 2560x1600, 1024 lights, cap 128)
    _, _, ubo, _ = camera(SMALL_FRAME["width"], SMALL_FRAME["height"], SMALL_FRAME["grid"])
    lights = abi.scene_lights(7, SMALL_FRAME["n_lights"])
    cells, idx, tot, ovf = O.cull_lights(ubo, O.build_aabbs(ubo), lights, SMALL_FRAME["cap"])
    g["cull_cells"] = cells.view(np.uint32).reshape(-1)
    g["cull_total"] = np.array([tot, ovf], np.int64)
    g["cull_head"] = idx[:4096].astype(np.uint32)
    g["cull_digest"] = np.array(index_digest(idx))

    # IBL stages
    eq, cube0, diff, spec, lut = ibl_inputs()
    g["eq2cube"] = cube0.astype(np.float32)
    g["diffuse"] = diff.astype(np.float32)
    g["specular"] = spec.astype(np.float32)
    g["brdf_lut"] = lut.astype(np.float32)

    # shading of a small frame with those maps
    desc, cam, fubo, ivp = camera(FW, FH, FGRID)
    fl = abi.scene_lights(13, FNL)
    fr = abi.scene_fragments(14, desc, ivp, threads=4)
    fcells, fidx, _, _ = O.cull_lights(fubo, O.build_aabbs(fubo), fl, FCAP)
    maps = dict(spec=spec, spec_size=SO, spec_mips=MIPS, diff=diff, diff_size=SD, lut=lut, lut_w=LUT, lut_h=LUT)
    ids = np.arange(FW * FH, dtype=np.uint32)
    g["shade_rgb"] = O.shade(fr, FW, FH, ivp, cam, fubo, abi.pbr_defaults(), fl, fcells, fidx, maps, ids,
                             threads=4).astype(np.float32)

    # contraction filter FP64 twin (specular, every texel of mips 1..4)
    ceq = abi.scene_equirect(ENV_SEED, 2 * CS, CS, threads=4).reshape(CS, 2 * CS, 4)
    ccube = O.equirect_to_cube(ceq, CS).reshape(-1)
    g["contract"] = np.concatenate([O.contract(1, ccube, CS, CSO, CMIPS, m) for m in range(1, CMIPS)]).astype(np.float64)

    # radiance network FP64 forward at seeded parameters
    rng = np.random.default_rng(21)
    P = (rng.normal(size=O.MLP_PARAMS) * 0.2).astype(np.float32)
    d = rng.normal(size=(32, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x = np.concatenate([d, rng.uniform(0, 1, (32, 1))], axis=1).astype(np.float32)
    g["mlp_params"] = P
    g["mlp_inputs"] = x
    g["mlp_logrgb"] = O.mlp_forward(P, x, 1.0)

    # tone mapping (every operator, gamma on) of 64 HDR pixels
    rgb = np.exp(rng.uniform(-4, 4, (3, 64))).astype(np.float32)
    g["tonemap_in"] = rgb
    g["tonemap_out"] = np.stack([O.tonemap(rgb, op, True) for op in range(4)]).astype(np.float32)
    return g


def main():
    g = compute()
    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes): " + ", ".join(f"{k} {v.shape}" for k, v in g.items()))


if __name__ == "__main__":
    main()
