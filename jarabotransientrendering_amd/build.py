"""Build recipe for the HIP library (gfx950) — compiled in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "_lib")
LIB = os.path.join(LIBDIR, "libtransient_hip.so")

SOURCES = ["tr_kernels.hip", "tr_guide.hip", "tr_bvh.cpp", "tr_image_io.cpp", "tr_obj.cpp"]
HEADERS = ["tr_device.h", "tr_internal.h", "tr_guide.h", "tr_guide_host.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


CLI_SRC = os.path.join(PKG, "cli", "render_scene.cpp")
CLI = os.path.join(LIBDIR, "render_scene")
# the reference's command line (bunnykiller.cpp) over the same library
BK_SRC = os.path.join(PKG, "cli", "bunnykiller.cpp")
BK = os.path.join(LIBDIR, "bunnykiller")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed ({res.returncode}):\n{res.stderr[-4000:]}")


def build(verbose: bool = False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    _run([HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          *[os.path.join(CSRC, s) for s in SOURCES], "-o", tmp], verbose)
    os.replace(tmp, LIB)
    # the C++ host driver over the C ABI (plain g++; finds the library next to itself)
    for src, exe in ((CLI_SRC, CLI), (BK_SRC, BK)):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src,
              "-L", LIBDIR, "-ltransient_hip", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"], verbose)
        os.replace(exe + ".tmp", exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
