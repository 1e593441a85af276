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
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")
ARCH = os.environ.get("MJMLX_HIP_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 instructions are disabled for all device code of this project
# FP32 division and square root by the hardware's reciprocal / square-root
# instructions (1-2 ulp, not correctly rounded) and FP32 denormals flushed to
# zero: fewer instructions per division, square root and reciprocal in the
# step kernels (quadruped -1.3 %, humanoid -2.7 % per control step); FP64 is
# unaffected (the Newton factor and the quaternion integration keep it).
HIP_FLAGS = ["-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero",
             "-Wno-unused-result", "-Wno-unused-value"]
HOST_SRC = ["mjcf.cpp", "hostmodel.cpp"]
# narrow-phase sets of the step kernel, one translation unit each (step_np.hip)
NP_SETS = (0, 1, 2, 4, 5, 6)
JOBS = max(1, min(16, os.cpu_count() or 1))


def _run(cmd, cwd=CSRC):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def _objects(tag, extra):
    """Compile the engine's translation units in parallel into _build/<tag>/;
    returns the object paths."""
    obj = os.path.join(PKG, "_build", tag)
    os.makedirs(obj, exist_ok=True)
    inc = [f"-I{INC}", f"-I{CSRC}"]
    jobs = [([HIPCC, *HIP_FLAGS, *extra, *inc, "-c", src, "-o", os.path.join(obj, src.rsplit(".", 1)[0] + ".o")],
             os.path.join(obj, src.rsplit(".", 1)[0] + ".o")) for src in ("engine.hip", "mlp.hip", *HOST_SRC)]
    # each set twice: for models of at most 32 dofs (MJK_BIG 0) and with every path
    jobs += [([HIPCC, *HIP_FLAGS, *extra, *inc, f"-DMJK_NP={k}", f"-DMJK_BIG={big}", "-c", "step_np.hip", "-o",
               os.path.join(obj, f"step_np{k}{sfx}.o")], os.path.join(obj, f"step_np{k}{sfx}.o"))
             for k in NP_SETS for big, sfx in ((0, ""), (1, "_big"))]
    # and with 32 lanes per env (two envs per wave), small models only
    jobs += [([HIPCC, *HIP_FLAGS, *extra, *inc, f"-DMJK_NP={k}", "-DMJK_BIG=0", "-DMJK_HALF=1", "-c", "step_np.hip",
               "-o", os.path.join(obj, f"step_np{k}_half.o")], os.path.join(obj, f"step_np{k}_half.o")) for k in NP_SETS]
    # the largest translation units first
    jobs.sort(key=lambda j: 0 if "step_np" in j[1] else 1)
    with ThreadPoolExecutor(JOBS) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    return [o for _, o in jobs]


def _link(out, objs):
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out, *objs])


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmjmlx.so")
    _link(out, _objects("engine", []))
    return out


def build_prof():
    """Profiling build of the engine (per-stage cycle counters compiled in)."""
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmjmlx_prof.so")
    _link(out, _objects("prof", ["-DMJK_PROFILE"]))
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
    # the stage profiles (bench.py --stages) always describe the current kernels
    build_prof()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "engine": build_engine, "mjb": build_mjb, "emu": build_emu, "prof": build_prof}[w]()
