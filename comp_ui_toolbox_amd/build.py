"""Build libborn.so in-tree with hipcc for gfx950.

The reference built its MEX with nvcc for sm_61 and linked cudart/cublas
(+auxiliary/nvmex_linux.m:22, :40-41); this library needs only the HIP
runtime.  Packed-FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "libborn.so"
SOURCES = [PKG / "csrc" / "born_api.hip", PKG / "csrc" / "born_kernels.hip", PKG / "csrc" / "born_structured.hip",
           PKG / "csrc" / "born_split.hip", PKG / "csrc" / "born_split3d.hip",
           PKG / "csrc" / "born_structured3d.hip", PKG / "csrc" / "born_fewcol.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BORN_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}", f"-I{PKG / 'csrc'}",
]


def build(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, *map(str, SOURCES), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
