"""In-tree builds: libjpgpu.so for gfx950 (hipcc) and the Fortran host
program fortran/jp_host (flang: ISO_C_BINDING module fortran/jpgpu.f90 and
the md_run input readers fortran/jp_input.f90)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "jpgpu.hip")
OUT = os.path.join(HERE, "libjpgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 VALU instructions in device code
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


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FDIR = os.path.join(HERE, "fortran")
FHOST = os.path.join(FDIR, "jp_host")


def build_fortran(verbose: bool = False) -> str:
    """Fortran module + host program, linked against the in-tree library
    (runpath $ORIGIN/.., so the pair travels together)."""
    mdir = os.path.join(FDIR, "_build")
    os.makedirs(mdir, exist_ok=True)
    cmd = [FLANG, "-O2", "-J", mdir, os.path.join(FDIR, "jpgpu.f90"), os.path.join(FDIR, "jp_input.f90"),
           os.path.join(FDIR, "jp_host.f90"),
           "-L" + HERE, "-ljpgpu", "-Wl,-rpath,$ORIGIN/..", "-o", FHOST + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=mdir)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stderr[-4000:])
    os.replace(FHOST + ".tmp", FHOST)
    if verbose:
        print(" ".join(cmd))
    return FHOST


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_fortran(verbose=True))
