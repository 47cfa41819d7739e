"""In-tree build of libjpgpu.so for gfx950 (hipcc)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "jpgpu.hip")
OUT = os.path.join(HERE, "libjpgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 VALU ops stay off (GPU box notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def build_library(verbose: bool = False) -> str:
    cmd = [HIPCC, *FLAGS, "-o", OUT + ".tmp", SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print(" ".join(cmd))
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))
