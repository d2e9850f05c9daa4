"""Build libffm.so (gfx950 HIP kernels + C++ host runtime) in-tree.

    python fflame-rust_amd/build.py          # builds fflame-rust_amd/lib/libffm.so

Every .hip file is compiled for gfx950 only, with FMA contraction off (the
kernels state the reference's FP32 operation order) and packed-FP32 ops
disabled; the objects are linked with the host runtime into one shared
library whose exports are exactly include/ffm.h.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(OUT_DIR, "libffm.so")
OBJ_DIR = os.path.join(PKG, "_build")

HIP_SOURCES = ["iterate.hip", "stencils.hip", "tilesplat.hip", "softsplat.hip", "effects.hip",
               "noalloc.hip"]
HOST_SOURCES = ["runtime.cpp", "scene.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build(verbose=False):
    hipcc = _hipcc()
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, src + ".o")
        out = _run([hipcc, "-x", "hip", *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ_DIR, src + ".o")
        _run([hipcc, *COMMON, "--offload-arch=gfx950", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
