"""Builds the in-tree native artefacts with hipcc/g++ (no cmake, no JIT cache).

Outputs (git-ignored, shipped with the tree):
  distembed_amd/libforce_embed.so   HIP kernels for gfx950 + C ABI + host code
  distembed_amd/bin/dist_embed      C++ driver with the reference CLI flags
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("FE_OFFLOAD_ARCH", "gfx950")

LIB = PKG / "libforce_embed.so"
DRIVER = PKG / "bin" / "dist_embed"

# Packed-FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no implicit contraction: every fused multiply-add in the kernels is an
    # explicit fmaf, so rounding does not depend on the unroll instantiation
    "-ffp-contract=off",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-D__HIP_PLATFORM_AMD__",
              "-Wall", "-Wno-unused-result"]

HIP_SOURCES = ["fe_kernels.hip"]
HOST_SOURCES = ["fe_host.cpp", "fe_plan.cpp", "fe_job.cpp", "fe_capi.cpp", "fe_engine.cpp"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def build(verbose: bool = False) -> Path:
    hipcc = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    gxx = shutil.which("g++") or "g++"
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []
    inc = [f"-I{INCLUDE}", f"-I{CSRC}", f"-I{ROCM / 'include'}"]
    for src in HIP_SOURCES:
        if not (CSRC / src).exists():
            continue
        obj = objdir / (src + ".o")
        _run([hipcc, *HIP_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in HOST_SOURCES:
        if not (CSRC / src).exists():
            continue
        obj = objdir / (src + ".o")
        _run([gxx, *HOST_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", *map(str, objs),
          "-o", str(tmp), f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, LIB)
    drv_src = CSRC / "dist_embed_main.cpp"
    if drv_src.exists():
        DRIVER.parent.mkdir(exist_ok=True)
        _run([gxx, "-O2", "-std=c++17", *inc, str(drv_src), "-o", str(DRIVER),
              f"-L{PKG}", "-lforce_embed", f"-Wl,-rpath,$ORIGIN/..",
              f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
