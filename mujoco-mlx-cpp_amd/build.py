"""Build the engine and its test companions in-tree.

  lib/libmjmlx.so      HIP engine + mjmlx C ABI (hipcc, gfx950)
  lib/libmjb.so        mjb C ABI over the engine (hipcc host compile)
  lib/libmjmlx_emu.so  test-only serial interpreter of the kernel source (g++)

Run ``python mujoco-mlx-cpp_amd/build.py`` or ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")
ARCH = os.environ.get("MJMLX_HIP_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are disabled for all device code of this project
HIP_FLAGS = ["-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]
HOST_SRC = ["mjcf.cpp", "hostmodel.cpp"]


def _run(cmd, cwd=CSRC):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmjmlx.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", f"-I{INC}", f"-I{CSRC}", "-o", out, "engine.hip", "mlp.hip", *HOST_SRC])
    return out


def build_prof():
    """Profiling build of the engine (per-stage cycle counters compiled in)."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmjmlx_prof.so")
    _run([HIPCC, *HIP_FLAGS, "-DMJK_PROFILE", "-shared", f"-I{INC}", f"-I{CSRC}", "-o", out, "engine.hip", "mlp.hip",
          *HOST_SRC])
    return out


def build_mjb():
    out = os.path.join(LIB, "libmjb.so")
    _run([HIPCC, "-std=c++17", "-O2", "-fPIC", "-shared", f"-I{INC}", f"-I{CSRC}", "-o", out, "mjb.cpp",
          f"-L{LIB}", "-lmjmlx", "-Wl,-rpath,$ORIGIN"])
    return out


def build_emu():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmjmlx_emu.so")
    # a*b+c contracted into fused multiply-adds, as hipcc does in device code
    _run(["g++", "-std=c++17", "-O2", "-mfma", "-ffp-contract=fast", "-fPIC", "-shared", f"-I{CSRC}", "-o", out,
          "emu.cpp", *HOST_SRC])
    return out


def build_all():
    build_engine()
    if os.path.exists(os.path.join(CSRC, "mjb.cpp")):
        build_mjb()
    build_emu()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "engine": build_engine, "mjb": build_mjb, "emu": build_emu, "prof": build_prof}[w]()
