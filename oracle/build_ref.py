"""Builds the oracle's C loop nest (born_ref_loop.c) with gcc + OpenMP into
oracle/libborn_ref_loop.so.  TEST INFRASTRUCTURE ONLY.

The reference's own operator path is MATLAB plus a CUDA MEX; neither MATLAB
nor nvcc exists in this image, so there is no `oracle/_ref` build of it.
"""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "born_ref_loop.c"
LIB = HERE / "libborn_ref_loop.so"


def build() -> Path:
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-std=c11", str(SRC), "-o", str(LIB), "-lm"]
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    print(build())
