"""Builds libhvx.so (HIP kernels + C ABI for gfx950) and the CPU oracle.

Everything is compiled in-tree so the built libraries travel with the source:
  hellovulkan_amd/libhvx.so          the product (hipcc --offload-arch=gfx950)
  oracle/_build/libhvx_oracle.so     the CPU oracle (g++ -ffp-contract=off)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhvx.so")
ORACLE_SRC = os.path.join(ROOT, "oracle", "hvx_oracle.cpp")
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "libhvx_oracle.so")

HIP_SOURCES = ["cf_kernels.hip", "shade_kernels.hip", "ibl_kernels.hip", "hvx_api.hip", "batch_api.hip", "contract.hip", "mlp.hip", "post_kernels.hip"]
CPP_SOURCES = ["scene.cpp", "hdr.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    # bit-reproducible arithmetic: no fused multiply-add unless written out
    "-ffp-contract=off",
    # no packed-FP32 VALU instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall", "-pthread"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


def build_product(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        jobs.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
    for src in CPP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    objs = [os.path.join(BUILD, s.rsplit(".", 1)[0] + ".o") for s in HIP_SOURCES + CPP_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-pthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(ORACLE_LIB), exist_ok=True)
    tmp = ORACLE_LIB + ".tmp"
    _run(["g++", *CXX_FLAGS, "-shared", "-fno-fast-math", "-o", tmp, ORACLE_SRC], verbose)
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


TEST_KERNELS_SRC = [os.path.join(ROOT, "tests", "gpu_kernels", f) for f in ("fastmath_check.hip", "mfma_layout.hip", "mlp_encode.hip")]
TEST_KERNELS_LIB = os.path.join(ROOT, "tests", "_build", "libhvx_testkernels.so")


def build_test_kernels(verbose: bool = False) -> str:
    """Test-only GPU kernels (exhaustive arithmetic checks) used by tests/."""
    os.makedirs(os.path.dirname(TEST_KERNELS_LIB), exist_ok=True)
    tmp = TEST_KERNELS_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *TEST_KERNELS_SRC], verbose)
    os.replace(tmp, TEST_KERNELS_LIB)
    return TEST_KERNELS_LIB


FRAME_SRC = os.path.join(PKG, "host", "hvx_frame.cpp")
FRAME_BIN = os.path.join(PKG, "bin", "hvx_frame")


def build_frame_driver(verbose: bool = False) -> str:
    """The C++20 frame driver (include/hvx/pipelines.hpp over libhvx.so)."""
    os.makedirs(os.path.dirname(FRAME_BIN), exist_ok=True)
    tmp = FRAME_BIN + ".tmp"
    _run([HIPCC, "-std=c++20", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"), FRAME_SRC, "-L", PKG, "-lhvx",
          "-Wl,-rpath,$ORIGIN/..", "-o", tmp], verbose)
    os.replace(tmp, FRAME_BIN)
    return FRAME_BIN


def build(verbose: bool = False) -> None:
    with cf.ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_product, verbose), ex.submit(build_oracle, verbose),
                ex.submit(build_test_kernels, verbose)]
        for f in futs:
            f.result()
    build_frame_driver(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, ORACLE_LIB)
