"""Builds the MEX shim (comp_ui_toolbox_amd/mex/combined_quick_gpu.cpp)
against the MEX API test double in this folder into
tests/mex_mock/libmexmock.so, linked to the in-tree libborn.so.  TEST
INFRASTRUCTURE ONLY."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
LIB = HERE / "libmexmock.so"


def build() -> Path:
    cmd = ["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Wextra", f"-I{HERE}", f"-I{ROOT / 'include'}",
           str(ROOT / "comp_ui_toolbox_amd" / "mex" / "combined_quick_gpu.cpp"), str(HERE / "mock_mex.cpp"),
           f"-L{ROOT / 'comp_ui_toolbox_amd'}", "-lborn", "-Wl,-rpath,$ORIGIN/../../comp_ui_toolbox_amd",
           "-o", str(LIB)]
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    print(build())
