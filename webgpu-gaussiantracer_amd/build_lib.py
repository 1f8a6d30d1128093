"""Builds libgt.so (HIP device code for gfx950 + host C++).

Flags that the bit-exact parity of hit lists depends on:
  -ffp-contract=off                         no fused multiply-add anywhere
  -fhip-fp32-correctly-rounded-divide-sqrt  IEEE division / square root on the GPU
  -fno-gpu-flush-denormals-to-zero          f32 denormals kept, as on the host
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 instructions
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = [
    "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
    "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
]

# Per-file hipcc flags.  gt_forward.hip: leave wave-uniform branches
# unstructured.  The traversal loop's leaf/node branch is uniform (the node
# reference is read once per wave); structurized, it became flow blocks whose
# PHIs merged the register k-buffer with undefined values, and the register
# allocator then copied the 32-register buffer into and out of every leaf
# visit, and spilled 88 B per lane.  Unstructured: no copies, no spills,
# 4 % less time per frame, the same results.
FILE_FLAGS = {"gt_forward.hip": ["-mllvm", "-structurizecfg-skip-uniform-regions=true"]}

HIP_SOURCES = ["gt_forward.hip", "gt_backward.hip", "gt_build.hip", "gt_train.hip", "gt_image.hip", "gt_noalloc.hip"]
CXX_SOURCES = ["gt_scene.cpp", "gt_host_build.cpp", "gt_util.cpp", "gt_batch.cpp", "gt_autotune.cpp", "gt_png.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build_lib(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + FILE_FLAGS.get(s, []) + ["-c", os.path.join(CSRC, s), "-o", o])
    for s in CXX_SOURCES:
        o = os.path.join(BUILD, s + ".o")
        jobs.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    objs = [os.path.join(BUILD, s + ".o") for s in HIP_SOURCES + CXX_SOURCES]
    out = os.path.join(LIBDIR, "libgt.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs +
         ["-Wl,-soname,libgt.so"], verbose)
    os.replace(out + ".tmp", out)
    build_node_addon(verbose)
    return out


NODE_INCLUDE = "/usr/include/node"


def build_node_addon(verbose=False):
    """lib/gtrace.node: the Node-API addon of the JavaScript host (node/),
    linked against lib/libgt.so (found next to it at run time).  Skipped
    when the Node headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    out = os.path.join(LIBDIR, "gtrace.node")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + NODE_INCLUDE,
          os.path.join(PKG, "node", "gt_napi.cpp"), "-o", out + ".tmp", "-L" + LIBDIR, "-lgt",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build_lib(verbose="-v" in sys.argv))
