"""Builds the MI355X trainer library (libsplatgs.so) and the Node.js addon
in-tree, for gfx950.

    python splat.js_amd/build.py          # library + addon
    python splat.js_amd/build.py --lib    # library only

Flags that matter for results:
  -ffp-contract=off     no implicit multiply-add fusion: results round exactly
                        as the source spells them (explicit fmaf only)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                        no packed FP32 (v_pk_*_f32) instructions
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsplatgs.so")
ADDON = os.path.join(HERE, "addon", "gs_napi.node")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]

SOURCES = ["gs_project.hip", "gs_sort.hip", "gs_raster.hip", "gs_update.hip", "gs_loss.hip",
           "gs_refine.hip", "gs_ply.hip", "gs_cam.hip", "gs_fd.hip", "gs_trainer.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "argument unused" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_lib(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "-c", "-x", "hip", os.path.join(CSRC, src), "-o", obj] + HIP_FLAGS
        out = _run(cmd)
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_addon():
    src = os.path.join(HERE, "addon", "gs_napi.cc")
    if not os.path.exists(src):
        return None
    inc = "/usr/include/node"
    if not os.path.exists(os.path.join(inc, "node_api.h")):
        return None
    cxx = shutil.which("g++") or "g++"
    tmp = ADDON + ".tmp"
    _run([cxx, "-O2", "-fPIC", "-shared", "-std=c++17", "-I" + inc, "-I" + os.path.join(ROOT, "include"),
          "-DNAPI_VERSION=8", "-DNODE_GYP_MODULE_NAME=gs_napi", src, "-o", tmp, "-ldl"])
    os.replace(tmp, ADDON)
    return ADDON


def build(lib_only=False):
    build_lib()
    if not lib_only:
        build_addon()


if __name__ == "__main__":
    build(lib_only="--lib" in sys.argv)
    print("built", LIB)
