"""CPU emulation of the GPU scene build's tree kernels (no GPU).

The kernels of webgpu-gaussiantracer_amd/csrc/gt_build.hip that make the
tree (k_karras, k_treelet_roots, k_ploc, k_refit_top, k_gather_final,
k_collapse) are taken as text from the source, compiled for the CPU with a
few definitions (one thread per lane, a real barrier for __syncthreads, a
barrier-based __ballot), and run on seeded boxes by tests/build_emul/
harness.cpp.  The harness checks what the GPU tests cannot see directly:
the leaf permutation, that each proxy sits in exactly one leaf, that every
box contains its children, and no capacity overflow.  This test also
checks that the 4-wide tree the collapse emits has exactly the SAH cost
its dynamic programme predicted (an error in the slot bookkeeping would
break the equality).  The GPU tests check the same builder end to end
through renders.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD_SRC = os.path.join(ROOT, "webgpu-gaussiantracer_amd", "csrc", "gt_build.hip")
INTERNAL_H = os.path.join(ROOT, "webgpu-gaussiantracer_amd", "csrc", "gt_internal.h")
HARNESS = os.path.join(ROOT, "tests", "build_emul", "harness.cpp")

SHIMS = r'''
#include <stdint.h>
#include <math.h>
#include <string.h>
#include <stdio.h>
#include <stdlib.h>
#include <algorithm>
#include <vector>
#include <thread>
#include <pthread.h>
#include <cstring>
using std::min; using std::max;
struct float4 { float x, y, z, w; };
static inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __restrict__
#define __launch_bounds__(...)
#define __shared__ static
static thread_local struct { uint32_t x; } threadIdx;
static thread_local struct { uint32_t x; } blockIdx;
static struct { uint32_t x; } gridDim;
static struct { uint32_t x; } blockDim;
static pthread_barrier_t g_bar;
static inline void __syncthreads() { pthread_barrier_wait(&g_bar); }
static int g_flag[1024]; static uint32_t g_shfl[1024];
static inline uint64_t __ballot(int pred) {
  g_flag[threadIdx.x] = pred ? 1 : 0; __syncthreads();
  uint64_t m = 0; uint32_t w0 = threadIdx.x & ~63u;
  for (uint32_t l = 0; l < 64 && w0 + l < blockDim.x; ++l) if (g_flag[w0 + l]) m |= 1ull << l;
  __syncthreads(); return m; }
static inline int __shfl(int v, int src, int w) {
  g_shfl[threadIdx.x] = (uint32_t)v; __syncthreads();
  int r = (int)g_shfl[(threadIdx.x & ~63u) + src]; __syncthreads(); return r; }
template <typename T> static inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <typename T> static inline T atomicOr(T* p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
#define __hip_atomic_fetch_add(p, v, o, s) __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST)
static inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }
static inline int __clzll(long long x) { return __builtin_clzll((unsigned long long)x); }
static inline float __uint_as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t __float_as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
'''

RUNNER = r'''
template <typename F> static void launch(uint32_t grid, uint32_t block, F body) {
  gridDim.x = grid; blockDim.x = block;
  pthread_barrier_init(&g_bar, nullptr, block);
  std::vector<std::thread> th;
  for (uint32_t t = 0; t < block; ++t)
    th.emplace_back([&, t]() {
      threadIdx.x = t;
      for (uint32_t b = 0; b < grid; ++b) { blockIdx.x = b; body(); pthread_barrier_wait(&g_bar); }
    });
  for (auto& x : th) x.join();
  pthread_barrier_destroy(&g_bar);
}
'''


def _between(src, start, end):
    a = src.index(start)
    return src[a:src.index(end, a)]


def _kernels_inc():
    src = open(BUILD_SRC).read()
    defs = _between(src, "#define BUILD_BLOCK", "namespace {")
    body = _between(src, "struct BNode {", "__device__ __forceinline__ uint32_t wave_min_u32")
    body += _between(src, "__device__ __forceinline__ int kdelta", "/* Refit, phase 1: one thread per leaf")
    body += _between(src, "/* Refit, phase 3: one workgroup climbs", "/* Refit, proxies: records and leaf boxes")
    body += _between(src, "__global__ __launch_bounds__(BUILD_BLOCK) void k_need(", "inline uint32_t blocks_for")
    return SHIMS + f'#include "{INTERNAL_H}"\n' + defs + "\n" + body + RUNNER


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is absent")
    d = tmp_path_factory.mktemp("build_emul")
    (d / "kern.inc").write_text(_kernels_inc())
    shutil.copy(HARNESS, d / "harness.cpp")
    exe = d / "harness"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-o", str(exe), str(d / "harness.cpp")], check=True,
                   capture_output=True, text=True, timeout=300)
    return exe


def _boxes(path, n, seed):
    """Clustered boxes of very different sizes, as proxy boxes are."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-10, 10, (8, 3))
    c = centres[rng.integers(0, 8, n)] + rng.normal(0, 1.0, (n, 3))
    half = np.exp(rng.normal(-3.0, 0.8, (n, 3)))
    np.concatenate([c - half, c + half], axis=1).astype(np.float32).tofile(path)


@pytest.mark.parametrize("n", [2, 3, 9, 100, 256, 257, 300, 4097, 12000])
def test_emulated_device_build(harness, tmp_path, n):
    f = tmp_path / "boxes.bin"
    _boxes(f, n, seed=n)
    r = subprocess.run([str(harness), str(f)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    out = r.stdout
    m = re.search(r"missing (\d+), dup (\d+), bad (\d+), SAH\(4-wide, /root area\) ([0-9.]+)", out)
    assert m and m.group(1) == "0" and m.group(2) == "0" and m.group(3) == "0", out
    assert "ovf 0" in out, out
    d1 = re.search(r"root D1 \(4-wide DP cost\) ([0-9.]+) \(root leaf (\d)\)", out)
    sah, pred, root_leaf = float(m.group(4)), float(d1.group(1)), d1.group(2) == "1"
    # a wide root adds its own visit (C_T = 1 in units of the root's area) to the DP's cost
    assert abs((pred if root_leaf else pred - 1.0) - sah) <= 1e-4 * max(1.0, sah), out
