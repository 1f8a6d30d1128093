/*
 * gt_build.hip - scene construction on the GPU (gt_scene_create_device,
 * gt_scene_update_device).
 *
 * Replaces the host stages of the reference's acceleration-container build
 * for the Gaussian path:
 *   activations + record packing    src/GS_ply.ts:187-196, 225-235
 *   gaussianSceneToBvh              src/main.ts:30-111 (filter, proxy transforms)
 *   wasm TLAS build                 webrtx/src/wasm_bvh_builder.ts:260-377,
 *                                   webrtx/bvh/src/lib.rs:253-347, 482-505
 * which the reference runs on the CPU (and which runs out of memory on
 * large scenes, README.md:50-51).  Everything stays in HBM:
 *
 *   k_prep     one thread per Gaussian: activations in double, the 48-B
 *              geometry and 192-B SH records, proxy transform + inverse and
 *              padded world box (gt_prep.h, the same code as the host
 *              builder, so the records are bit-identical), and per block
 *              the proxy count and centroid bounds (wave + LDS reductions)
 *   k_bounds   one workgroup folds the block partials into the totals
 *   k_keys     64-bit key = 30-bit Morton code of the box centroid << 32 |
 *              Gaussian id (unique keys; non-proxies sort last)
 *   radix sort rocPRIM radix_sort_keys
 *   k_gather   64-B proxy records in key order (BVH leaf order)
 *   k_karras   binary radix tree over the sorted keys (T. Karras, "Maximizing
 *              parallelism in the construction of BVHs, octrees, and k-d
 *              trees", HPG 2012), one thread per inner node
 *   k_treelet_roots / k_ploc
 *              the maximal subtrees of at most 256 leaves are rebuilt by
 *              locally-ordered clustering (PLOC), one workgroup per subtree
 *              in LDS, with the SAH leaf decision as nodes are made (the
 *              tree's SAH cost falls by 10 % on a dense cluster, 6 % on the
 *              64-cluster scene); leaves are renumbered so every subtree's
 *              proxies stay contiguous
 *   k_refit_top  the radix-tree nodes above the treelets: bottom-up boxes
 *              with one arrival counter per node (the second child to arrive
 *              continues) and the SAH-optimal leaf decision
 *              cost = min(C_I n A, C_T A + cost_L + cost_R) for n <= 8, in
 *              one workgroup (no device-scope fences)
 *   k_refit_local / k_refit_mid / k_refit_top (GT_TREELET_LEAVES 0: plain
 *              radix tree) the same refit over the whole tree: per block of
 *              256 leaves with LDS counters, then per group of 4096 leaves,
 *              then the nodes that span groups in one workgroup
 *   k_gather_final  64-B proxy records and leaf boxes in the final leaf order
 *   k_collapse top-down, one launch per level (the level sizes stay on
 *              the device; the host reads them every 8 levels): each
 *              4-wide node opens the inner child of largest surface area
 *              until it has four
 *              children (the rule of the host builder); subtrees marked as
 *              leaves become leaf refs over their contiguous proxy range
 *   k_need     bottom-up worst-case traversal stack per node
 *
 * The node and proxy formats are those of the host builder (gt_internal.h),
 * so the render and backward kernels run unchanged.  Which hits a ray finds
 * does not depend on the tree (conservative boxes, (t, id) key order), so a
 * device-built scene renders exactly what the host-built one renders.
 */
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "gt_prep.h"
#include "gt_scene_impl.h"

#define BUILD_BLOCK 256
#define BIN_LEAF 0x80000000u /* binary ref: leaf = sorted proxy index | BIN_LEAF */
#define BIN_NONE 0xffffffffu

/* counters (u32); CTL_LVL + L = fronts (4-wide nodes) of collapse level L */
enum { CTL_NPROX = 0, CTL_CMIN = 1, CTL_CMAX = 4, CTL_NODES = 7, CTL_OVF = 9, CTL_ITEMS = 10, CTL_STALE = 11,
       CTL_PNODES = 12, CTL_BAD = 13 /* first invalid Gaussian (gt_prep_valid, gt_prep_proxy_finite), ~0 if none */, CTL_LVL = 16,
       CTL_WORDS = CTL_LVL + GT_STACK_MAX + 4 };
#define COLLAPSE_BATCH 8 /* collapse levels launched between host reads of the level counts */
#define MID_LEAVES (16 * BUILD_BLOCK) /* leaves per refit phase-2 workgroup (LDS counters) */
#ifndef GT_TREELET_LEAVES
#define GT_TREELET_LEAVES 256 /* PLOC treelets: Morton subtrees of at most this many leaves (0: none) */
#endif
#define PLOC_RADIUS 16        /* nearest-neighbour window, clusters on each side */
#define PART_WORDS 8 /* k_prep partial per block: count, centroid min[3], max[3], pad */

namespace {

struct BNode {              /* inner node of the binary tree (80 B) */
  float lo[3], hi[3];
  uint32_t left, right;     /* binary refs */
  uint32_t parent;          /* BIN_NONE for the root */
  uint32_t first, count;    /* sorted proxy range */
  uint32_t leaf;            /* 1: this subtree becomes one leaf (count <= GT_LEAF_MAX) */
  float D[4];               /* least cost of the subtree as at most 1..4 wide-node slots */
  uint8_t choice[4];        /* D[0]: 0 leaf, k wide node with k slots left; D[j>0]: 0 one slot, k split */
  float bcost;              /* binary SAH cost with SAH-optimal leaves */
  uint32_t _pad[2];
};

/* The binary tree's leaves first: a node becomes a leaf where
 * C_I cnt A <= C_T A + cost_L + cost_R (cnt <= 8), the SAH-optimal binary
 * decision (bcost).  Then the least-SAH-cost 4-wide collapse over those
 * leaves (the host builder's Collapser, GT_COLLAPSE_DP): from the children's
 * D(., 1..4) (a single proxy costs C_I A in any number of slots), node N's
 *   D(N, 1) = min(C_I cnt A [cnt <= 8], C_T A + min_k D(l, k) + D(r, 4 - k))
 *   D(N, j) = min(D(N, 1), min_{0<k<j} D(l, k) + D(r, j - k)),  j = 2..4.
 * Keeping the binary leaves matters: a collapse free to split them makes
 * many small leaves (twice the nodes at 1M proxies) whose stack traffic the
 * cost model does not see, and the frame got no faster. */
__device__ __forceinline__ void dp_node(const float* dl, const float* dr, float bl, float br, uint32_t count, float A,
                                        float* D, uint8_t* ch, uint32_t* leaf, float* bcost) {
  const float leafc = count <= GT_LEAF_MAX ? ((float)GT_SAH_ISECT * (float)count + (float)GT_SAH_LEAF) * A : INFINITY;
  const float bsplit = 1.0f * A + bl + br;
  if (leafc <= bsplit) { /* a binary leaf: one slot, never split */
    *bcost = leafc;
    D[0] = D[1] = D[2] = D[3] = leafc;
    ch[0] = ch[1] = ch[2] = ch[3] = 0;
    *leaf = 1u;
    return;
  }
  *bcost = bsplit;
  float dist4 = INFINITY;
  uint8_t k4 = 1;
#pragma unroll
  for (int k = 1; k <= 3; ++k) {
    const float v = dl[k - 1] + dr[3 - k];
    if (v < dist4) { dist4 = v; k4 = (uint8_t)k; }
  }
  const float wide = 1.0f * A + dist4;
  if (leafc <= wide) { D[0] = leafc; ch[0] = 0; *leaf = 1u; }
  else { D[0] = wide; ch[0] = k4; *leaf = 0u; }
#pragma unroll
  for (int j = 2; j <= 4; ++j) {
    float best = D[0];
    uint8_t c = 0;
    for (int k = 1; k < j; ++k) {
      const float v = dl[k - 1] + dr[j - k - 1];
      if (v < best) { best = v; c = (uint8_t)k; }
    }
    D[j - 1] = best;
    ch[j - 1] = c;
  }
}

struct Front {              /* collapse work item: binary ref -> 4-wide node index */
  uint32_t bref, node4;
};

__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float ord2f(uint32_t k) {
  const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

__device__ __forceinline__ uint32_t dev_leaf_ref(uint32_t first, uint32_t count) {
  return GT_LEAF_BIT | ((count - 1u) << GT_LEAF_COUNT_SHIFT) | first;
}

/* surface area as the host builder computes it (Box::area) */
__device__ __forceinline__ double box_area(const float* lo, const float* hi) {
  const double dx = fmaxf(0.0f, hi[0] - lo[0]), dy = fmaxf(0.0f, hi[1] - lo[1]), dz = fmaxf(0.0f, hi[2] - lo[2]);
  return 2.0 * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

__global__ void k_init(unsigned int* ctl) {
  for (uint32_t i = threadIdx.x; i < CTL_WORDS; i += blockDim.x) {
    uint32_t v = 0;
    if (i >= CTL_CMIN && i < CTL_CMIN + 3) v = 0xffffffffu;
    if (i == CTL_BAD) v = 0xffffffffu;
    if (i == CTL_NODES) v = 1u;   /* node 0 = root */
    if (i == CTL_LVL) v = 1u;     /* level 0 = the root's front */
    ctl[i] = v;
  }
}

__global__ void k_fill(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  if (i < n) p[i] = v;
}

/* box: 2 float4 per Gaussian {lo.xyz, has_proxy} {hi.xyz, 0} */
__global__ __launch_bounds__(BUILD_BLOCK) void k_prep(const float* __restrict__ raw, uint32_t n, double alpha_min,
                                                      float* __restrict__ geom, float* __restrict__ sh,
                                                      float* __restrict__ o2w, float* __restrict__ w2o,
                                                      float4* __restrict__ box, uint32_t* __restrict__ part,
                                                      unsigned int* __restrict__ bad) {
  const uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  bool has = false;
  float c[3] = {0.0f, 0.0f, 0.0f};
  if (i < n) {
    float r[GT_RAW_FLOATS];
    const float* src = raw + (size_t)i * GT_RAW_FLOATS;
#pragma unroll
    for (int j = 0; j < GT_RAW_FLOATS; ++j) r[j] = src[j];
    double act[8];
    float g[12], s[48];
    gt_prep_activate(r, act, g, s);
    if (!gt_prep_valid(r, act)) atomicMin(bad, i);
    float4* G = reinterpret_cast<float4*>(geom) + 3 * (size_t)i;
    G[0] = make_float4(g[0], g[1], g[2], g[3]);
    G[1] = make_float4(g[4], g[5], g[6], g[7]);
    G[2] = make_float4(g[8], g[9], g[10], g[11]);
    float4* SH = reinterpret_cast<float4*>(sh) + 12 * (size_t)i;
#pragma unroll
    for (int q = 0; q < 12; ++q) SH[q] = make_float4(s[4 * q], s[4 * q + 1], s[4 * q + 2], s[4 * q + 3]);
    float m[12], inv[12];
    has = gt_prep_proxy(act, r, alpha_min, m, inv) != 0;
    float4* B = box + 2 * (size_t)i;
    if (has) {
      float4* M = reinterpret_cast<float4*>(o2w) + 3 * (size_t)i;
      float4* V = reinterpret_cast<float4*>(w2o) + 3 * (size_t)i;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        M[q] = make_float4(m[4 * q], m[4 * q + 1], m[4 * q + 2], m[4 * q + 3]);
        V[q] = make_float4(inv[4 * q], inv[4 * q + 1], inv[4 * q + 2], inv[4 * q + 3]);
      }
      float lo[3], hi[3];
      gt_prep_box(m, lo, hi);
      if (!gt_prep_proxy_finite(m, inv, lo, hi)) atomicMin(bad, i);
      B[0] = make_float4(lo[0], lo[1], lo[2], 1.0f);
      B[1] = make_float4(hi[0], hi[1], hi[2], 0.0f);
#pragma unroll
      for (int j = 0; j < 3; ++j) c[j] = 0.5f * (lo[j] + hi[j]);
    } else {
      B[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      B[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
  }
  const uint64_t m = __ballot(has);
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t mn[3], mx[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    mn[j] = wave_min_u32(has ? f2ord(c[j]) : 0xffffffffu);
    mx[j] = wave_max_u32(has ? f2ord(c[j]) : 0u);
  }
  /* one partial per block, no atomics: 7 same-line atomics per wave took
   * 1.26 ms at 1M Gaussians, serialised at one L2 channel */
  __shared__ uint32_t s_part[BUILD_BLOCK / 64][PART_WORDS];
  const uint32_t wv = threadIdx.x >> 6;
  if (lane == 0) {
    s_part[wv][0] = (uint32_t)__popcll(m);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      s_part[wv][1 + j] = mn[j];
      s_part[wv][4 + j] = mx[j];
    }
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    const uint32_t j = threadIdx.x;
    uint32_t v = s_part[0][j];
#pragma unroll
    for (uint32_t w = 1; w < BUILD_BLOCK / 64; ++w) {
      const uint32_t x = s_part[w][j];
      v = j == 0 ? v + x : (j < 4 ? min(v, x) : max(v, x));
    }
    part[(size_t)blockIdx.x * PART_WORDS + j] = v;
  }
}

/* proxy count and centroid bounds of k_prep's per-block partials -> ctl */
__global__ __launch_bounds__(1024) void k_bounds(const uint32_t* __restrict__ part, uint32_t nblocks,
                                                 unsigned int* __restrict__ ctl) {
  uint32_t v[7] = {0u, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
  for (uint32_t b = threadIdx.x; b < nblocks; b += 1024) {
    const uint32_t* P = part + (size_t)b * PART_WORDS;
    v[0] += P[0];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      v[1 + j] = min(v[1 + j], P[1 + j]);
      v[4 + j] = max(v[4 + j], P[4 + j]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v[0] += (uint32_t)__shfl_xor((int)v[0], o, 64);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    v[1 + j] = wave_min_u32(v[1 + j]);
    v[4 + j] = wave_max_u32(v[4 + j]);
  }
  __shared__ uint32_t s_red[1024 / 64][PART_WORDS];
  const uint32_t wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0) {
#pragma unroll
    for (int j = 0; j < 7; ++j) s_red[wv][j] = v[j];
  }
  __syncthreads();
  if (threadIdx.x < 7) {
    const uint32_t j = threadIdx.x;
    uint32_t r = s_red[0][j];
    for (uint32_t w = 1; w < 1024 / 64; ++w) {
      const uint32_t x = s_red[w][j];
      r = j == 0 ? r + x : (j < 4 ? min(r, x) : max(r, x));
    }
    ctl[CTL_NPROX + j] = r; /* NPROX, CMIN[3], CMAX[3] are words 0..6 */
  }
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ __launch_bounds__(BUILD_BLOCK) void k_keys(const float4* __restrict__ box, uint32_t n,
                                                      const unsigned int* __restrict__ ctl,
                                                      unsigned long long* __restrict__ keys) {
  const uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  if (i >= n) return;
  const float4 b0 = box[2 * (size_t)i], b1 = box[2 * (size_t)i + 1];
  if (b0.w == 0.0f) {
    keys[i] = ~0ull;
    return;
  }
  const float lo[3] = {b0.x, b0.y, b0.z}, hi[3] = {b1.x, b1.y, b1.z};
  uint32_t q[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const float cmin = ord2f(ctl[CTL_CMIN + j]), cmax = ord2f(ctl[CTL_CMAX + j]);
    const float cj = 0.5f * (lo[j] + hi[j]);
    const float ext = cmax - cmin;
    const float u = ext > 0.0f ? (cj - cmin) / ext : 0.0f;
    int v = (int)(u * 1024.0f);
    v = v < 0 ? 0 : (v > 1023 ? 1023 : v);
    q[j] = (uint32_t)v;
  }
  const uint32_t morton = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
  keys[i] = ((unsigned long long)morton << 32) | i;
}

/* proxies in key order; front[0] = the root work item */
__global__ __launch_bounds__(BUILD_BLOCK) void k_gather(const unsigned long long* __restrict__ keys, uint32_t np,
                                                        const float4* __restrict__ w2o, const float4* __restrict__ box,
                                                        float4* __restrict__ prox, float4* __restrict__ lbox,
                                                        Front* __restrict__ front) {
  const uint32_t p = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  if (p == 0) front[0] = Front{np == 1 ? BIN_LEAF : 0u, 0u};
  if (p >= np) return;
  const uint32_t id = (uint32_t)keys[p];
  const float4* V = w2o + 3 * (size_t)id;
  float4* P = prox + 4 * (size_t)p;
  P[0] = V[0];
  P[1] = V[1];
  P[2] = V[2];
  P[3] = make_float4(__uint_as_float(id), 0.0f, 0.0f, 0.0f);
  lbox[2 * (size_t)p] = box[2 * (size_t)id];
  lbox[2 * (size_t)p + 1] = box[2 * (size_t)id + 1];
}

__device__ __forceinline__ int kdelta(const unsigned long long* __restrict__ k, int np, int i, int j) {
  if (j < 0 || j >= np) return -1;
  return __clzll((long long)(k[i] ^ k[j])); /* keys are unique: the xor is never 0 */
}

/* inner node i of the radix tree over np >= 2 sorted unique keys */
__global__ __launch_bounds__(BUILD_BLOCK) void k_karras(const unsigned long long* __restrict__ k, int np,
                                                        BNode* __restrict__ bn, uint32_t* __restrict__ lparent) {
  const int i = (int)(blockIdx.x * BUILD_BLOCK + threadIdx.x);
  if (i >= np - 1) return;
  const int d = kdelta(k, np, i, i + 1) > kdelta(k, np, i, i - 1) ? 1 : -1;
  const int dmin = kdelta(k, np, i, i - d);
  int lmax = 2;
  while (kdelta(k, np, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (kdelta(k, np, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = kdelta(k, np, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (kdelta(k, np, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? d : 0);
  const int lo = min(i, j), hi = max(i, j);
  const uint32_t left = lo == gamma ? (BIN_LEAF | (uint32_t)gamma) : (uint32_t)gamma;
  const uint32_t right = hi == gamma + 1 ? (BIN_LEAF | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
  bn[i].left = left;
  bn[i].right = right;
  bn[i].first = (uint32_t)lo;
  bn[i].count = (uint32_t)(hi - lo + 1);
  if (left & BIN_LEAF) lparent[gamma] = (uint32_t)i; else bn[gamma].parent = (uint32_t)i;
  if (right & BIN_LEAF) lparent[gamma + 1] = (uint32_t)i; else bn[gamma + 1].parent = (uint32_t)i;
  if (i == 0) bn[0].parent = BIN_NONE;
}

__device__ __forceinline__ void child_data(const BNode* bn, const float4* lbox, uint32_t ref, float* lo, float* hi,
                                           float* D, float* bcost) {
  if (ref & BIN_LEAF) {
    const uint32_t p = ref & ~BIN_LEAF;
    const float4 a = lbox[2 * (size_t)p], b = lbox[2 * (size_t)p + 1];
    lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
    hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
    const float c = ((float)GT_SAH_ISECT + (float)GT_SAH_LEAF) * (float)box_area(lo, hi);
    D[0] = D[1] = D[2] = D[3] = c;
    *bcost = c;
  } else {
    const BNode& c = bn[ref];
    for (int j = 0; j < 3; ++j) { lo[j] = c.lo[j]; hi[j] = c.hi[j]; }
    for (int j = 0; j < 4; ++j) D[j] = c.D[j];
    *bcost = c.bcost;
  }
}

/* Completes inner node N from its two children: box union and the
 * collapse costs and choices (dp_node). */
__device__ __forceinline__ void refit_node(BNode* __restrict__ bn, const float4* __restrict__ lbox, BNode* N) {
  float llo[3], lhi[3], rlo[3], rhi[3], dl[4], dr[4], bl, br;
  child_data(bn, lbox, N->left, llo, lhi, dl, &bl);
  child_data(bn, lbox, N->right, rlo, rhi, dr, &br);
  float lo[3], hi[3];
  for (int j = 0; j < 3; ++j) {
    lo[j] = fminf(llo[j], rlo[j]);
    hi[j] = fmaxf(lhi[j], rhi[j]);
    N->lo[j] = lo[j];
    N->hi[j] = hi[j];
  }
  float D[4], bc;
  uint8_t ch[4];
  uint32_t leaf;
  dp_node(dl, dr, bl, br, N->count, (float)box_area(lo, hi), D, ch, &leaf, &bc);
  for (int j = 0; j < 4; ++j) { N->D[j] = D[j]; N->choice[j] = ch[j]; }
  N->leaf = leaf;
  N->bcost = bc;
}

/* Refit, phase 1: one thread per leaf climbs while it is the second child to
 * arrive, through the nodes whose leaves all belong to this block (an inner
 * node's index is an end of its range, so such a node's index is in the
 * block's range too: its arrival counter lives in LDS and the handshake is
 * workgroup-scoped).  Where the next node spans blocks the climb stops and
 * the completed child is handed to phase 2, in the list of its group of
 * MID_LEAVES leaves.  The handed-over subtrees are disjoint and lie in the
 * block, so a block hands over at most BUILD_BLOCK of them and a group at
 * most MID_LEAVES; one global atomic per block reserves the block's slots. */
__global__ __launch_bounds__(BUILD_BLOCK) void k_refit_local(uint32_t np, BNode* __restrict__ bn,
                                                             const float4* __restrict__ lbox,
                                                             const uint32_t* __restrict__ lparent,
                                                             uint32_t* __restrict__ items,
                                                             uint32_t* __restrict__ mid_count) {
  __shared__ uint32_t s_visit[BUILD_BLOCK];
  __shared__ uint32_t s_out[BUILD_BLOCK];
  __shared__ uint32_t s_n, s_base;
  s_visit[threadIdx.x] = 0u;
  if (threadIdx.x == 0) s_n = 0u;
  __syncthreads();
  const uint32_t p = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  const uint32_t blo = blockIdx.x * BUILD_BLOCK, bhi = min(blo + BUILD_BLOCK, np);
  if (p < np) {
    uint32_t child = BIN_LEAF | p, node = lparent[p];
    while (node != BIN_NONE) {
      BNode* N = bn + node;
      if (!(N->first >= blo && N->first + N->count <= bhi)) { /* spans blocks: phase 2 */
        s_out[atomicAdd(&s_n, 1u)] = child;
        break;
      }
      if (__hip_atomic_fetch_add(&s_visit[node - blo], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
        break;
      refit_node(bn, lbox, N);
      child = node;
      node = N->parent;
    }
  }
  __syncthreads();
  const uint32_t m = s_n;
  if (m == 0u) return;
  const uint32_t g = blockIdx.x / (MID_LEAVES / BUILD_BLOCK);
  if (threadIdx.x == 0) s_base = atomicAdd(mid_count + g, m);
  __syncthreads();
  if (threadIdx.x < m) items[(size_t)g * MID_LEAVES + s_base + threadIdx.x] = s_out[threadIdx.x];
}

/* Refit, phase 2: one workgroup per group of MID_LEAVES leaves climbs from
 * the group's phase-1 items through the nodes that lie in the group (LDS
 * arrival counters, workgroup scope; phase 1's writes are visible across the
 * kernel boundary).  Children of nodes that span groups go to phase 3 (at
 * most MID_LEAVES per group, by the same disjointness). */
__global__ __launch_bounds__(1024) void k_refit_mid(uint32_t np, BNode* __restrict__ bn,
                                                    const float4* __restrict__ lbox,
                                                    const uint32_t* __restrict__ lparent,
                                                    const uint32_t* __restrict__ items,
                                                    const uint32_t* __restrict__ mid_count,
                                                    uint32_t* __restrict__ top_items, unsigned int* ctl) {
  __shared__ uint32_t s_visit[MID_LEAVES];
  __shared__ uint32_t s_out[MID_LEAVES];
  __shared__ uint32_t s_n, s_base;
  for (uint32_t i = threadIdx.x; i < MID_LEAVES; i += 1024u) s_visit[i] = 0u;
  if (threadIdx.x == 0) s_n = 0u;
  __syncthreads();
  const uint32_t glo = blockIdx.x * MID_LEAVES, ghi = min(glo + MID_LEAVES, np);
  const uint32_t cnt = mid_count[blockIdx.x];
  const uint32_t* in = items + (size_t)blockIdx.x * MID_LEAVES;
  for (uint32_t it = threadIdx.x; it < cnt; it += 1024u) {
    uint32_t child = in[it];
    uint32_t node = (child & BIN_LEAF) ? lparent[child & ~BIN_LEAF] : bn[child].parent;
    while (node != BIN_NONE) {
      BNode* N = bn + node;
      if (!(N->first >= glo && N->first + N->count <= ghi)) { /* spans groups: phase 3 */
        s_out[atomicAdd(&s_n, 1u)] = child;
        break;
      }
      if (__hip_atomic_fetch_add(&s_visit[node - glo], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u)
        break;
      refit_node(bn, lbox, N);
      child = node;
      node = N->parent;
    }
  }
  __syncthreads();
  const uint32_t m = s_n;
  if (m == 0u) return;
  if (threadIdx.x == 0) s_base = atomicAdd(ctl + CTL_ITEMS, m);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < m; i += 1024u) top_items[s_base + i] = s_out[i];
}

/* Refit, phase 3: one workgroup climbs from every phase-2 item through the
 * nodes that span groups (all that phases 1 and 2 wrote is visible: kernel
 * boundaries), with workgroup-scoped arrival counters in `visit`. */
__global__ __launch_bounds__(1024) void k_refit_top(BNode* __restrict__ bn, const float4* __restrict__ lbox,
                                                    const uint32_t* __restrict__ lparent,
                                                    const uint32_t* __restrict__ items, const unsigned int* ctl,
                                                    uint32_t* visit) {
  const uint32_t n_items = ctl[CTL_ITEMS];
  for (uint32_t it = threadIdx.x; it < n_items; it += 1024u) {
    const uint32_t c = items[it];
    uint32_t node = (c & BIN_LEAF) ? lparent[c & ~BIN_LEAF] : bn[c].parent;
    while (node != BIN_NONE) {
      if (__hip_atomic_fetch_add(visit + node, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0u) break;
      BNode* N = bn + node;
      refit_node(bn, lbox, N);
      node = N->parent;
    }
  }
}

__device__ __forceinline__ bool openable(const BNode* bn, uint32_t ref) {
  return !(ref & BIN_LEAF) && bn[ref].leaf == 0u;
}

/* one 4-wide node per work item; cap = capacity of nodes[] and out[] */
/* Collapse level L: the input count is ctl[CTL_LVL + L] (written by the
 * launch of level L - 1), so the host launches levels without reading it;
 * every inner child becomes a 4-wide node of level L + 1, allocated from
 * CTL_NODES, which makes each level's nodes one contiguous index range. */
__device__ void collapse_one(const Front e, Front* __restrict__ out, unsigned int* ctl, uint32_t level,
                             const BNode* __restrict__ bn, const float4* __restrict__ lbox,
                             gt_node4* __restrict__ nodes, uint32_t cap) {
  uint32_t ch[4];
  int m;
  if (!openable(bn, e.bref)) {
    ch[0] = e.bref;
    m = 1;
  } else {
    /* the slots chosen by the collapse DP, left to right: a stack of
     * (subtree, slots) pairs whose slot counts sum to 4 */
    const BNode& R = bn[e.bref];
    uint32_t sref[4];
    int sj[4], sp = 0;
    sref[sp] = R.right; sj[sp++] = 4 - R.choice[0];
    sref[sp] = R.left; sj[sp++] = R.choice[0];
    m = 0;
    while (sp > 0) {
      --sp;
      const uint32_t r = sref[sp];
      const int j = sj[sp];
      const int c = (j == 1 || (r & BIN_LEAF)) ? 0 : bn[r].choice[j - 1];
      if (c == 0 || sp + 2 > 4) {
        ch[m++] = r;
        continue;
      }
      sref[sp] = bn[r].right; sj[sp++] = j - c;
      sref[sp] = bn[r].left; sj[sp++] = c;
    }
  }
  gt_node4 nd;
  for (int c = 0; c < 4; ++c) {
    nd.lo_x[c] = nd.lo_y[c] = nd.lo_z[c] = INFINITY; /* empty slots never hit */
    nd.hi_x[c] = nd.hi_y[c] = nd.hi_z[c] = -INFINITY;
    nd.child[c] = GT_NODE_EMPTY;
    nd._pad[c] = 0u;
  }
  for (int c = 0; c < m; ++c) {
    const uint32_t r = ch[c];
    float lo[3], hi[3];
    uint32_t ref;
    if (r & BIN_LEAF) {
      const uint32_t p = r & ~BIN_LEAF;
      const float4 a = lbox[2 * (size_t)p], b = lbox[2 * (size_t)p + 1];
      lo[0] = a.x; lo[1] = a.y; lo[2] = a.z;
      hi[0] = b.x; hi[1] = b.y; hi[2] = b.z;
      ref = dev_leaf_ref(p, 1u);
    } else {
      const BNode& b = bn[r];
      for (int j = 0; j < 3; ++j) { lo[j] = b.lo[j]; hi[j] = b.hi[j]; }
      if (b.leaf) {
        ref = dev_leaf_ref(b.first, b.count);
      } else {
        const uint32_t j = atomicAdd(ctl + CTL_NODES, 1u);
        const uint32_t slot = atomicAdd(ctl + CTL_LVL + level + 1u, 1u);
        if (j < cap && slot < cap) {
          out[slot] = Front{r, j};
          ref = j;
        } else {
          atomicOr(ctl + CTL_OVF, 1u);
          ref = GT_NODE_EMPTY;
        }
      }
    }
    nd.lo_x[c] = lo[0]; nd.hi_x[c] = hi[0];
    nd.lo_y[c] = lo[1]; nd.hi_y[c] = hi[1];
    nd.lo_z[c] = lo[2]; nd.hi_z[c] = hi[2];
    nd.child[c] = ref;
  }
  if (e.node4 < cap) nodes[e.node4] = nd;
}

__global__ __launch_bounds__(BUILD_BLOCK) void k_collapse(const Front* __restrict__ in, uint32_t level,
                                                          Front* __restrict__ out, unsigned int* ctl,
                                                          const BNode* __restrict__ bn,
                                                          const float4* __restrict__ lbox,
                                                          gt_node4* __restrict__ nodes, uint32_t cap) {
  const uint32_t n_in = min(ctl[CTL_LVL + level], cap);
  for (uint32_t t = blockIdx.x * BUILD_BLOCK + threadIdx.x; t < n_in; t += gridDim.x * BUILD_BLOCK)
    collapse_one(in[t], out, ctl, level, bn, lbox, nodes, cap);
}

/* Treelets: the maximal radix-tree subtrees of at most GT_TREELET_LEAVES
 * leaves (a child of a larger node, or the root of a small tree).  They
 * partition the leaves; each is rebuilt by k_ploc and the nodes above them
 * keep the radix tree's structure. */
__global__ __launch_bounds__(BUILD_BLOCK) void k_treelet_roots(uint32_t np, const BNode* __restrict__ bn,
                                                               uint32_t* __restrict__ items, unsigned int* ctl) {
  const uint32_t i = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  bool e0 = false, e1 = false;
  uint32_t r0 = 0, r1 = 0;
  if (i < np - 1) {
    const BNode& N = bn[i];
    if (N.count > GT_TREELET_LEAVES) {
      r0 = N.left;
      r1 = N.right;
      e0 = ((r0 & BIN_LEAF) ? 1u : bn[r0].count) <= GT_TREELET_LEAVES;
      e1 = ((r1 & BIN_LEAF) ? 1u : bn[r1].count) <= GT_TREELET_LEAVES;
    } else if (i == 0) { /* the whole tree is one treelet */
      r0 = 0;
      e0 = true;
    }
  }
  const uint64_t m0 = __ballot(e0), m1 = __ballot(e1);
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t base = 0;
  if (lane == 0 && (m0 | m1)) base = atomicAdd(ctl + CTL_ITEMS, (unsigned int)(__popcll(m0) + __popcll(m1)));
  base = (uint32_t)__shfl((int)base, 0, 64);
  if (e0) items[base + __popcll(m0 & below)] = r0;
  if (e1) items[base + __popcll(m0) + __popcll(m1 & below)] = r1;
}

/* PLOC (D. Meister, J. Bittner, "Parallel Locally-Ordered Clustering for
 * Bounding Volume Hierarchy Construction", IEEE TVCG 2018) inside one
 * treelet per workgroup, in LDS: every cluster finds the cluster within
 * PLOC_RADIUS positions whose union box has the least surface area (ties to
 * the lower position, which makes the smallest pair mutual, so every round
 * merges); mutual pairs merge into a new node in the lower position; the
 * list is compacted in order.  A node's box, count and SAH leaf decision
 * are known when it is made.  Then positions are assigned top-down so each
 * subtree's leaves are contiguous: perm[final position] = radix-tree
 * position.  The last merge reuses the treelet root's index (its parent's
 * reference stays valid); the others take slots after the radix tree's
 * nodes. */
#define PLOC_N GT_TREELET_LEAVES
#define PLOC_NODE_LEAF 0x80000000u /* local refs: leaf j = flag | j, node l = l */
__global__ __launch_bounds__(PLOC_N) void k_ploc(uint32_t np, BNode* __restrict__ bn,
                                                 const float4* __restrict__ lbox,
                                                 const uint32_t* __restrict__ items, unsigned int* ctl,
                                                 uint32_t* __restrict__ perm, uint32_t bn_cap) {
  /* the cluster list, compacted in place each round */
  __shared__ float s_lo[3][PLOC_N], s_hi[3][PLOC_N];
  __shared__ float s_D[4][PLOC_N]; /* collapse costs of the clusters (dp_node) */
  __shared__ float s_bc[PLOC_N];   /* binary costs of the clusters */
  __shared__ uint32_t s_ref[PLOC_N], s_cnt[PLOC_N];
  __shared__ uint32_t s_nn[PLOC_N];
  __shared__ float s_blo[3][PLOC_N], s_bhi[3][PLOC_N]; /* made nodes, by local slot */
  __shared__ float s_bD[4][PLOC_N], s_bbc[PLOC_N];
  __shared__ uint32_t s_bch[PLOC_N]; /* the four choices, one byte each */
  __shared__ uint32_t s_bl[PLOC_N], s_br[PLOC_N], s_bcnt[PLOC_N], s_bleaf[PLOC_N], s_bfirst[PLOC_N];
  __shared__ uint32_t s_itend[PLOC_N], s_leafpos[PLOC_N];
  __shared__ uint32_t s_wsum[PLOC_N / 64];
  __shared__ uint32_t s_made, s_base, s_fail;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
  const uint32_t n_items = ctl[CTL_ITEMS];
  for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
    const uint32_t R = items[it];
    if (R & BIN_LEAF) { /* one-leaf treelet: stays in place */
      if (tid == 0) perm[R & ~BIN_LEAF] = R & ~BIN_LEAF;
      continue;
    }
    const uint32_t first = bn[R].first, count = bn[R].count;
    __syncthreads(); /* the previous treelet's readers of the shared state are done */
    if (tid == 0) {
      s_made = 0u;
      s_fail = 0u;
      s_base = count > 2u ? atomicAdd(ctl + CTL_PNODES, count - 2u) : 0u;
    }
    if (tid < count) {
      const float4 a = lbox[2 * (size_t)(first + tid)], b = lbox[2 * (size_t)(first + tid) + 1];
      s_lo[0][tid] = a.x; s_lo[1][tid] = a.y; s_lo[2][tid] = a.z;
      s_hi[0][tid] = b.x; s_hi[1][tid] = b.y; s_hi[2][tid] = b.z;
      const float lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
      const float c = ((float)GT_SAH_ISECT + (float)GT_SAH_LEAF) * (float)box_area(lo, hi);
#pragma unroll
      for (int q = 0; q < 4; ++q) s_D[q][tid] = c;
      s_bc[tid] = c;
      s_ref[tid] = PLOC_NODE_LEAF | tid;
      s_cnt[tid] = 1u;
    }
    __syncthreads();
    const uint32_t base = s_base;
    const bool fits = (np - 1u) + base + (count > 2u ? count - 2u : 0u) <= bn_cap;
    uint32_t n = count, nit = 0;
    while (n > 1u && fits) {
      /* nearest neighbour in the window */
      if (tid < n) {
        const float alo[3] = {s_lo[0][tid], s_lo[1][tid], s_lo[2][tid]};
        const float ahi[3] = {s_hi[0][tid], s_hi[1][tid], s_hi[2][tid]};
        const uint32_t jlo = tid > PLOC_RADIUS ? tid - PLOC_RADIUS : 0u;
        const uint32_t jhi = min(tid + PLOC_RADIUS, n - 1u);
        float best = INFINITY;
        uint32_t bj = tid;
        for (uint32_t j = jlo; j <= jhi; ++j) {
          if (j == tid) continue;
          const float dx = fmaxf(ahi[0], s_hi[0][j]) - fminf(alo[0], s_lo[0][j]);
          const float dy = fmaxf(ahi[1], s_hi[1][j]) - fminf(alo[1], s_lo[1][j]);
          const float dz = fmaxf(ahi[2], s_hi[2][j]) - fminf(alo[2], s_lo[2][j]);
          const float a = dx * dy + dy * dz + dz * dx;
          if (a < best || bj == tid) { best = a; bj = j; }
        }
        s_nn[tid] = bj;
      }
      __syncthreads();
      /* this cluster and, for a merge, its partner, read into registers
       * before the barrier that precedes the in-place compaction */
      bool keep = false, make = false;
      float mlo[3] = {0.f, 0.f, 0.f}, mhi[3] = {0.f, 0.f, 0.f}, mD[4] = {0.f, 0.f, 0.f, 0.f}, mbc = 0.f;
      float plo[3] = {0.f, 0.f, 0.f}, phi[3] = {0.f, 0.f, 0.f}, pD[4] = {0.f, 0.f, 0.f, 0.f}, pbc = 0.f;
      uint32_t mref = 0u, mcnt = 0u, pref = 0u, pcnt = 0u;
      if (tid < n) {
        const uint32_t j = s_nn[tid];
        const bool mutual = j != tid && s_nn[j] == tid;
        keep = !mutual || tid < j;
        make = mutual && tid < j;
        if (keep) {
#pragma unroll
          for (int q = 0; q < 3; ++q) { mlo[q] = s_lo[q][tid]; mhi[q] = s_hi[q][tid]; }
#pragma unroll
          for (int q = 0; q < 4; ++q) mD[q] = s_D[q][tid];
          mbc = s_bc[tid];
          mref = s_ref[tid];
          mcnt = s_cnt[tid];
        }
        if (make) {
#pragma unroll
          for (int q = 0; q < 3; ++q) { plo[q] = s_lo[q][j]; phi[q] = s_hi[q][j]; }
#pragma unroll
          for (int q = 0; q < 4; ++q) pD[q] = s_D[q][j];
          pbc = s_bc[j];
          pref = s_ref[j];
          pcnt = s_cnt[j];
        }
      }
      /* exclusive scan of keep over the block */
      const uint64_t km = __ballot(keep);
      if (lane == 0) s_wsum[wv] = (uint32_t)__popcll(km);
      __syncthreads();
      uint32_t pos = (uint32_t)__popcll(km & ((1ull << lane) - 1ull)), total = 0;
      for (uint32_t w = 0; w < PLOC_N / 64; ++w) {
        if (w < wv) pos += s_wsum[w];
        total += s_wsum[w];
      }
      if (make) {
        const uint32_t l = atomicAdd(&s_made, 1u);
        float lo[3], hi[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          lo[q] = fminf(mlo[q], plo[q]);
          hi[q] = fmaxf(mhi[q], phi[q]);
        }
        const uint32_t cnt = mcnt + pcnt;
        float D[4], bc;
        uint8_t chs[4];
        uint32_t leaf;
        dp_node(mD, pD, mbc, pbc, cnt, (float)box_area(lo, hi), D, chs, &leaf, &bc);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          s_blo[q][l] = lo[q]; s_bhi[q][l] = hi[q];
          s_lo[q][pos] = lo[q]; s_hi[q][pos] = hi[q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) { s_bD[q][l] = D[q]; s_D[q][pos] = D[q]; }
        s_bbc[l] = bc;
        s_bc[pos] = bc;
        s_bch[l] = (uint32_t)chs[0] | ((uint32_t)chs[1] << 8) | ((uint32_t)chs[2] << 16) | ((uint32_t)chs[3] << 24);
        s_bl[l] = mref;
        s_br[l] = pref;
        s_bcnt[l] = cnt;
        s_bleaf[l] = leaf;
        s_ref[pos] = l;
        s_cnt[pos] = cnt;
      } else if (keep) {
#pragma unroll
        for (int q = 0; q < 3; ++q) { s_lo[q][pos] = mlo[q]; s_hi[q][pos] = mhi[q]; }
#pragma unroll
        for (int q = 0; q < 4; ++q) s_D[q][pos] = mD[q];
        s_bc[pos] = mbc;
        s_ref[pos] = mref;
        s_cnt[pos] = mcnt;
      }
      __syncthreads();
      if (total >= n) { /* no merge: only with NaN boxes */
        if (tid == 0) s_fail = 1u;
        break;
      }
      if (tid == 0) s_itend[nit] = s_made;
      n = total;
      ++nit;
    }
    __syncthreads();
    if (!fits || s_fail) {
      if (tid == 0) atomicOr(ctl + CTL_OVF, 1u);
      continue;
    }
    /* top-down leaf positions: the root (slot count - 2) starts at 0 */
    if (tid == 0) s_bfirst[count - 2u] = 0u;
    __syncthreads();
    for (uint32_t t = nit; t-- > 0;) {
      const uint32_t a = t ? s_itend[t - 1] : 0u, b = s_itend[t];
      for (uint32_t l = a + tid; l < b; l += PLOC_N) {
        const uint32_t f = s_bfirst[l], L = s_bl[l], Rr = s_br[l];
        const uint32_t cl = (L & PLOC_NODE_LEAF) ? 1u : s_bcnt[L];
        if (L & PLOC_NODE_LEAF) s_leafpos[L & ~PLOC_NODE_LEAF] = f; else s_bfirst[L] = f;
        if (Rr & PLOC_NODE_LEAF) s_leafpos[Rr & ~PLOC_NODE_LEAF] = f + cl; else s_bfirst[Rr] = f + cl;
      }
      __syncthreads();
    }
    /* write the nodes and the permutation */
    for (uint32_t l = tid; l + 1u < count; l += PLOC_N) {
      const uint32_t g = (l == count - 2u) ? R : (np - 1u) + base + l;
      BNode* B = bn + g;
#pragma unroll
      for (int q = 0; q < 3; ++q) { B->lo[q] = s_blo[q][l]; B->hi[q] = s_bhi[q][l]; }
      const uint32_t L = s_bl[l], Rr = s_br[l];
      B->left = (L & PLOC_NODE_LEAF) ? (BIN_LEAF | (first + s_leafpos[L & ~PLOC_NODE_LEAF]))
                                     : (L == count - 2u ? R : (np - 1u) + base + L);
      B->right = (Rr & PLOC_NODE_LEAF) ? (BIN_LEAF | (first + s_leafpos[Rr & ~PLOC_NODE_LEAF]))
                                       : (Rr == count - 2u ? R : (np - 1u) + base + Rr);
      if (l != count - 2u) B->parent = BIN_NONE; /* the treelet root keeps its radix-tree parent */
      B->first = first + s_bfirst[l];
      B->count = s_bcnt[l];
      B->leaf = s_bleaf[l];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        B->D[q] = s_bD[q][l];
        B->choice[q] = (uint8_t)(s_bch[l] >> (8 * q));
      }
      B->bcost = s_bbc[l];
    }
    if (tid < count) {
      const uint32_t lp = s_leafpos[tid];
      if (lp < count) perm[first + lp] = first + tid;
      else atomicOr(ctl + CTL_OVF, 1u);
    }
  }
}

/* proxies in final leaf order: position p holds the proxy at radix-tree
 * position perm[p] (identity without treelets); front[0] = the root */
__global__ __launch_bounds__(BUILD_BLOCK) void k_gather_final(const unsigned long long* __restrict__ keys,
                                                              uint32_t np, const uint32_t* __restrict__ perm,
                                                              const float4* __restrict__ w2o,
                                                              const float4* __restrict__ box,
                                                              float4* __restrict__ prox, float4* __restrict__ lbox,
                                                              Front* __restrict__ front) {
  const uint32_t p = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  if (p == 0) front[0] = Front{np == 1 ? BIN_LEAF : 0u, 0u};
  if (p >= np) return;
  const uint32_t q = perm ? perm[p] : p;
  const uint32_t id = (uint32_t)keys[q < np ? q : p];
  const float4* V = w2o + 3 * (size_t)id;
  float4* P = prox + 4 * (size_t)p;
  P[0] = V[0];
  P[1] = V[1];
  P[2] = V[2];
  P[3] = make_float4(__uint_as_float(id), 0.0f, 0.0f, 0.0f);
  lbox[2 * (size_t)p] = box[2 * (size_t)id];
  lbox[2 * (size_t)p + 1] = box[2 * (size_t)id + 1];
}

/* Refit, proxies: records and leaf boxes in leaf order from k_prep's output.
 * The tree stays valid only for the same proxy set: a proxy whose Gaussian
 * lost its proxy, or a different proxy count, marks the scene stale. */
__global__ __launch_bounds__(BUILD_BLOCK) void k_refit_prox(uint32_t np, float4* __restrict__ prox,
                                                            const float4* __restrict__ w2o,
                                                            const float4* __restrict__ box,
                                                            float4* __restrict__ lbox, unsigned int* ctl) {
  const uint32_t p = blockIdx.x * BUILD_BLOCK + threadIdx.x;
  if (p == 0 && ctl[CTL_NPROX] != np) atomicOr(ctl + CTL_STALE, 1u);
  if (p >= np) return;
  float4* P = prox + 4 * (size_t)p;
  const uint32_t id = __float_as_uint(P[3].x);
  const float4 b0 = box[2 * (size_t)id], b1 = box[2 * (size_t)id + 1];
  if (b0.w == 0.0f) {
    atomicOr(ctl + CTL_STALE, 1u);
    return;
  }
  const float4* V = w2o + 3 * (size_t)id;
  P[0] = V[0];
  P[1] = V[1];
  P[2] = V[2];
  lbox[2 * (size_t)p] = b0;
  lbox[2 * (size_t)p + 1] = b1;
}

/* Refit, nodes of one level (deepest level first): each child slot's box is
 * the union of its leaf's proxy boxes or of its child node's four slots
 * (empty slots hold +inf / -inf and drop out of the union). */
__global__ __launch_bounds__(BUILD_BLOCK) void k_refit4(gt_node4* __restrict__ nodes,
                                                        const uint32_t* __restrict__ order, uint32_t first,
                                                        uint32_t last, const float4* __restrict__ lbox) {
  const uint32_t t = first + blockIdx.x * BUILD_BLOCK + threadIdx.x;
  if (t >= last) return;
  const uint32_t j = order ? order[t] : t;
  gt_node4 nd = nodes[j];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const uint32_t r = nd.child[c];
    if (r == GT_NODE_EMPTY) continue;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (r & GT_LEAF_BIT) {
      const uint32_t f = r & GT_LEAF_FIRST_MASK, cnt = ((r >> GT_LEAF_COUNT_SHIFT) & GT_LEAF_COUNT_MASK) + 1u;
      for (uint32_t q = f; q < f + cnt; ++q) {
        const float4 a = lbox[2 * (size_t)q], b = lbox[2 * (size_t)q + 1];
        lo[0] = fminf(lo[0], a.x); lo[1] = fminf(lo[1], a.y); lo[2] = fminf(lo[2], a.z);
        hi[0] = fmaxf(hi[0], b.x); hi[1] = fmaxf(hi[1], b.y); hi[2] = fmaxf(hi[2], b.z);
      }
    } else {
      const gt_node4& k = nodes[r];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        lo[0] = fminf(lo[0], k.lo_x[q]); lo[1] = fminf(lo[1], k.lo_y[q]); lo[2] = fminf(lo[2], k.lo_z[q]);
        hi[0] = fmaxf(hi[0], k.hi_x[q]); hi[1] = fmaxf(hi[1], k.hi_y[q]); hi[2] = fmaxf(hi[2], k.hi_z[q]);
      }
    }
    nd.lo_x[c] = lo[0]; nd.lo_y[c] = lo[1]; nd.lo_z[c] = lo[2];
    nd.hi_x[c] = hi[0]; nd.hi_y[c] = hi[1]; nd.hi_z[c] = hi[2];
  }
  nodes[j] = nd;
}

/* worst-case stack of the subtree: max(m, (m - 1) + max need of inner children) */
__global__ __launch_bounds__(BUILD_BLOCK) void k_need(const gt_node4* __restrict__ nodes, uint32_t first,
                                                      uint32_t last, uint32_t* __restrict__ need) {
  const uint32_t j = first + blockIdx.x * BUILD_BLOCK + threadIdx.x;
  if (j >= last) return;
  const gt_node4 nd = nodes[j];
  uint32_t m = 0, cmax = 0;
  for (int c = 0; c < 4; ++c) {
    const uint32_t r = nd.child[c];
    if (r == GT_NODE_EMPTY) continue;
    ++m;
    if (!(r & GT_LEAF_BIT)) cmax = max(cmax, need[r]);
  }
  need[j] = max(m, (m - 1u) + cmax);
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + BUILD_BLOCK - 1) / BUILD_BLOCK); }

}  // namespace

struct gt_devbuild {
  DevBuf o2w, w2o, box, part, keys_a, keys_b, sort_tmp, lbox, lbox2, perm, bnode, lparent, visit, items, mid_count,
      top_items, front_a, front_b, need, ctl;
  unsigned int* host_ctl = nullptr; /* pinned, CTL_WORDS */
  ~gt_devbuild() {
    if (host_ctl) (void)hipHostFree(host_ctl);
  }
};

void gt_devbuild_destroy(gt_devbuild* w) { delete w; }

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      gt_set_error("%s failed: %s", #expr, hipGetErrorString(e_));                 \
      return GT_E_HIP;                                                             \
    }                                                                              \
  } while (0)
#define GT_TRY(expr)                \
  do {                              \
    const int rc_ = (expr);         \
    if (rc_ != GT_OK) return rc_;   \
  } while (0)

static int read_ctl(gt_devbuild* w, hipStream_t st) {
  HIP_TRY(hipMemcpyAsync(w->host_ctl, w->ctl.p, CTL_WORDS * sizeof(unsigned int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return GT_OK;
}

/* GT_E_INVALID naming the first Gaussian k_prep found invalid (gt_prep_valid, gt_prep_proxy_finite) */
static int check_bad(const gt_devbuild* w) {
  const uint32_t bad = w->host_ctl[CTL_BAD];
  if (bad == 0xffffffffu) return GT_OK;
  gt_set_error("Gaussian %u: non-finite parameter, zero quaternion, a scale that is 0 or inf in f32, or a proxy "
               "transform or box that is not finite in f32", bad);
  return GT_E_INVALID;
}

int gt_device_build(gt_scene* s, hipStream_t st) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!s->db) {
    s->db = new (std::nothrow) gt_devbuild();
    if (!s->db) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  }
  gt_devbuild* w = s->db;
  if (!w->host_ctl) HIP_TRY(hipHostMalloc((void**)&w->host_ctl, CTL_WORDS * sizeof(unsigned int)));
  const uint32_t n = s->n;
  GT_TRY(gt_scene_static_upload(s));
  GT_TRY(gt_devbuf_reserve(s->geom, (size_t)n * 48));
  GT_TRY(gt_devbuf_reserve(s->sh, (size_t)n * 192));
  GT_TRY(gt_devbuf_reserve(w->o2w, (size_t)n * 48));
  GT_TRY(gt_devbuf_reserve(w->w2o, (size_t)n * 48));
  GT_TRY(gt_devbuf_reserve(w->box, (size_t)n * 32));
  GT_TRY(gt_devbuf_reserve(w->part, (size_t)blocks_for(n) * PART_WORDS * 4));
  GT_TRY(gt_devbuf_reserve(w->keys_a, (size_t)n * 8));
  GT_TRY(gt_devbuf_reserve(w->keys_b, (size_t)n * 8));
  GT_TRY(gt_devbuf_reserve(w->ctl, CTL_WORDS * sizeof(unsigned int)));
  unsigned int* ctl = (unsigned int*)w->ctl.p;
  k_init<<<1, BUILD_BLOCK, 0, st>>>(ctl);
  HIP_TRY(hipGetLastError());
  if (n > 0) {
    k_prep<<<blocks_for(n), BUILD_BLOCK, 0, st>>>((const float*)s->raw_dev.p, n, s->alpha_min, (float*)s->geom.p,
                                                  (float*)s->sh.p, (float*)w->o2w.p, (float*)w->w2o.p,
                                                  (float4*)w->box.p, (uint32_t*)w->part.p, ctl + CTL_BAD);
    HIP_TRY(hipGetLastError());
    k_bounds<<<1, 1024, 0, st>>>((const uint32_t*)w->part.p, blocks_for(n), ctl);
    HIP_TRY(hipGetLastError());
    k_keys<<<blocks_for(n), BUILD_BLOCK, 0, st>>>((const float4*)w->box.p, n, ctl,
                                                  (unsigned long long*)w->keys_a.p);
    HIP_TRY(hipGetLastError());
    size_t tmp_bytes = 0;
    HIP_TRY(rocprim::radix_sort_keys(nullptr, tmp_bytes, (unsigned long long*)w->keys_a.p,
                                     (unsigned long long*)w->keys_b.p, n, 0, 64, st));
    GT_TRY(gt_devbuf_reserve(w->sort_tmp, tmp_bytes));
    HIP_TRY(rocprim::radix_sort_keys(w->sort_tmp.p, tmp_bytes, (unsigned long long*)w->keys_a.p,
                                     (unsigned long long*)w->keys_b.p, n, 0, 64, st));
  }
  GT_TRY(read_ctl(w, st));
  GT_TRY(check_bad(w));
  const uint32_t np = w->host_ctl[CTL_NPROX];
  if (np > n || np > GT_MAX_PROXIES) {
    gt_set_error("device build: %u proxies for %u Gaussians (limit %u)", np, n, GT_MAX_PROXIES);
    return GT_E_LIMIT;
  }
  /* every wide node but a one-leaf root has >= 2 children (the collapse DP
   * splits a node's slots between both of its binary children), and the
   * leaves hold >= 1 proxy each, so nodes <= np - 1 (1 for np <= 1) */
  const uint32_t cap = np > 1u ? np : 1u;
  GT_TRY(gt_devbuf_reserve(s->prox, (size_t)np * 64));
  GT_TRY(gt_devbuf_reserve(s->nodes, (size_t)cap * sizeof(gt_node4)));
  uint32_t root = GT_NODE_EMPTY, n_nodes = 0, need = 0, depth = 0;
  if (np > 0) {
    GT_TRY(gt_devbuf_reserve(w->lbox, (size_t)np * 32));
    /* radix-tree nodes, then the treelets' new nodes (fewer than np) */
    const uint32_t bn_cap = 2u * np;
    GT_TRY(gt_devbuf_reserve(w->bnode, (size_t)bn_cap * sizeof(BNode)));
    const bool treelets = GT_TREELET_LEAVES > 0 && np >= 2;
    if (treelets) {
      GT_TRY(gt_devbuf_reserve(w->perm, (size_t)np * 4));
      GT_TRY(gt_devbuf_reserve(w->lbox2, (size_t)np * 32));
    }
    GT_TRY(gt_devbuf_reserve(w->lparent, (size_t)np * 4));
    GT_TRY(gt_devbuf_reserve(w->visit, (size_t)np * 4));
    const uint32_t n_groups = (uint32_t)((np + MID_LEAVES - 1) / MID_LEAVES);
    GT_TRY(gt_devbuf_reserve(w->items, (size_t)n_groups * MID_LEAVES * 4));
    GT_TRY(gt_devbuf_reserve(w->mid_count, (size_t)n_groups * 4));
    GT_TRY(gt_devbuf_reserve(w->top_items, (size_t)np * 4));
    GT_TRY(gt_devbuf_reserve(w->front_a, (size_t)cap * sizeof(Front)));
    GT_TRY(gt_devbuf_reserve(w->front_b, (size_t)cap * sizeof(Front)));
    GT_TRY(gt_devbuf_reserve(w->need, (size_t)cap * 4));
    BNode* bn = (BNode*)w->bnode.p;
    const float4* lbox = (const float4*)w->lbox.p;
    k_gather<<<blocks_for(np), BUILD_BLOCK, 0, st>>>((const unsigned long long*)w->keys_b.p, np,
                                                     (const float4*)w->w2o.p, (const float4*)w->box.p,
                                                     (float4*)s->prox.p, (float4*)w->lbox.p, (Front*)w->front_a.p);
    HIP_TRY(hipGetLastError());
    if (np >= 2) {
      k_fill<<<blocks_for(np), BUILD_BLOCK, 0, st>>>((uint32_t*)w->visit.p, np, 0u);
      HIP_TRY(hipGetLastError());
      k_fill<<<blocks_for(n_groups), BUILD_BLOCK, 0, st>>>((uint32_t*)w->mid_count.p, n_groups, 0u);
      HIP_TRY(hipGetLastError());
      k_karras<<<blocks_for(np - 1), BUILD_BLOCK, 0, st>>>((const unsigned long long*)w->keys_b.p, (int)np, bn,
                                                           (uint32_t*)w->lparent.p);
      HIP_TRY(hipGetLastError());
      if (treelets) {
        /* treelets rebuilt by PLOC, then the radix-tree nodes above them */
        k_treelet_roots<<<blocks_for(np - 1), BUILD_BLOCK, 0, st>>>(np, bn, (uint32_t*)w->items.p, ctl);
        HIP_TRY(hipGetLastError());
        k_ploc<<<std::min(np, 4096u), PLOC_N, 0, st>>>(np, bn, lbox, (const uint32_t*)w->items.p, ctl,
                                                       (uint32_t*)w->perm.p, bn_cap);
        HIP_TRY(hipGetLastError());
        k_refit_top<<<1, 1024, 0, st>>>(bn, lbox, (const uint32_t*)w->lparent.p, (const uint32_t*)w->items.p, ctl,
                                        (uint32_t*)w->visit.p);
        HIP_TRY(hipGetLastError());
        k_gather_final<<<blocks_for(np), BUILD_BLOCK, 0, st>>>((const unsigned long long*)w->keys_b.p, np,
                                                               (const uint32_t*)w->perm.p, (const float4*)w->w2o.p,
                                                               (const float4*)w->box.p, (float4*)s->prox.p,
                                                               (float4*)w->lbox2.p, (Front*)w->front_a.p);
        HIP_TRY(hipGetLastError());
        lbox = (const float4*)w->lbox2.p;
      } else {
        k_refit_local<<<blocks_for(np), BUILD_BLOCK, 0, st>>>(np, bn, lbox, (const uint32_t*)w->lparent.p,
                                                              (uint32_t*)w->items.p, (uint32_t*)w->mid_count.p);
        HIP_TRY(hipGetLastError());
        k_refit_mid<<<n_groups, 1024, 0, st>>>(np, bn, lbox, (const uint32_t*)w->lparent.p,
                                               (const uint32_t*)w->items.p, (const uint32_t*)w->mid_count.p,
                                               (uint32_t*)w->top_items.p, ctl);
        HIP_TRY(hipGetLastError());
        k_refit_top<<<1, 1024, 0, st>>>(bn, lbox, (const uint32_t*)w->lparent.p, (const uint32_t*)w->top_items.p,
                                        ctl, (uint32_t*)w->visit.p);
        HIP_TRY(hipGetLastError());
      }
    }
    /* top-down collapse, one launch per level, COLLAPSE_BATCH levels between
     * host reads; level L's nodes are [level[L], level[L + 1]) because a
     * level's nodes are all allocated by the launch of the level above */
    const uint32_t cgrid = std::min(blocks_for(cap), 1024u);
    std::vector<uint32_t> level{0u, 1u};
    uint32_t launched = 0;
    for (;;) {
      for (uint32_t b = 0; b < COLLAPSE_BATCH && launched <= GT_STACK_MAX; ++b, ++launched) {
        Front* fin = (Front*)((launched & 1u) ? w->front_b.p : w->front_a.p);
        Front* fout = (Front*)((launched & 1u) ? w->front_a.p : w->front_b.p);
        k_collapse<<<cgrid, BUILD_BLOCK, 0, st>>>(fin, launched, fout, ctl, bn, lbox, (gt_node4*)s->nodes.p, cap);
        HIP_TRY(hipGetLastError());
      }
      GT_TRY(read_ctl(w, st));
      if (w->host_ctl[CTL_OVF]) {
        gt_set_error("device build: more than %u BVH nodes, or non-finite proxy boxes", cap);
        return GT_E_LIMIT;
      }
      level.assign({0u, 1u});
      uint32_t L = 1;
      while (L <= launched && w->host_ctl[CTL_LVL + L] > 0u) {
        level.push_back(level.back() + w->host_ctl[CTL_LVL + L]);
        ++L;
      }
      if (L <= launched) { /* level L is empty: done */
        level.push_back(level.back());
        break;
      }
      if (launched > GT_STACK_MAX) {
        gt_set_error("device build: BVH deeper than %u levels", GT_STACK_MAX);
        return GT_E_LIMIT;
      }
    }
    n_nodes = level.back();
    depth = (uint32_t)level.size() - 2u;
    s->rlevels.assign(level.begin(), level.end() - 1);
    uint32_t* needp = (uint32_t*)w->need.p;
    for (size_t L = level.size() - 2; L-- > 0;) {
      const uint32_t a = level[L], b = level[L + 1];
      if (b > a) {
        k_need<<<blocks_for(b - a), BUILD_BLOCK, 0, st>>>((const gt_node4*)s->nodes.p, a, b, needp);
        HIP_TRY(hipGetLastError());
      }
    }
    HIP_TRY(hipMemcpyAsync(w->host_ctl, needp, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    need = w->host_ctl[0];
    root = 0;
  }
  if (need > GT_STACK_MAX) {
    gt_set_error("BVH needs a traversal stack of %u entries (limit %u)", need, GT_STACK_MAX);
    return GT_E_LIMIT;
  }
  s->dev.geom = (const float*)s->geom.p;
  s->dev.sh = (const float*)s->sh.p;
  s->dev.prox = (const float*)s->prox.p;
  s->dev.nodes = (const gt_node4*)s->nodes.p;
  s->dev.faces = (const float*)s->faces.p;
  s->dev.root = root;
  s->dev.n = n;
  s->dev.n_prox = np;
  s->dev.n_nodes = n_nodes;
  s->n_nodes = n_nodes;
  s->max_stack = need;
  s->depth = depth;
  if (np == 0) s->rlevels.clear();
  s->ridentity = true;
  s->rsched = true;
  s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return GT_OK;
}

/* Refit schedule of a tree built elsewhere (the host builder numbers nodes
 * depth first): breadth-first levels from the root over a host copy. */
static int refit_schedule(gt_scene* s) {
  const uint32_t nn = s->dev.n_nodes;
  std::vector<gt_node4> nodes(nn);
  if (nn) HIP_TRY(hipMemcpy(nodes.data(), s->nodes.p, (size_t)nn * sizeof(gt_node4), hipMemcpyDeviceToHost));
  std::vector<uint32_t> order;
  order.reserve(nn);
  s->rlevels.assign(1, 0u);
  if (nn && s->dev.root != GT_NODE_EMPTY) {
    order.push_back(s->dev.root);
    size_t a = 0;
    while (a < order.size()) {
      const size_t b = order.size();
      for (size_t t = a; t < b; ++t)
        for (int c = 0; c < 4; ++c) {
          const uint32_t r = nodes[order[t]].child[c];
          if (r == GT_NODE_EMPTY || (r & GT_LEAF_BIT)) continue;
          if (r >= nn || order.size() >= nn) {
            gt_set_error("refit: malformed BVH (child %u of %u nodes)", r, nn);
            return GT_E_INVALID;
          }
          order.push_back(r);
        }
      s->rlevels.push_back((uint32_t)b);
      a = b;
    }
  }
  GT_TRY(gt_devbuf_reserve(s->rorder, std::max<size_t>(order.size(), 1) * 4));
  if (!order.empty()) HIP_TRY(hipMemcpy(s->rorder.p, order.data(), order.size() * 4, hipMemcpyHostToDevice));
  s->ridentity = false;
  s->rsched = true;
  return GT_OK;
}

int gt_device_refit(gt_scene* s, const float* raw_dev, hipStream_t st, uint32_t* rebuilt) {
  const auto t0 = std::chrono::steady_clock::now();
  *rebuilt = 0;
  if (!s->db) {
    s->db = new (std::nothrow) gt_devbuild();
    if (!s->db) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  }
  gt_devbuild* w = s->db;
  if (!w->host_ctl) HIP_TRY(hipHostMalloc((void**)&w->host_ctl, CTL_WORDS * sizeof(unsigned int)));
  const uint32_t n = s->n, np = s->dev.n_prox;
  if (!s->rsched) GT_TRY(refit_schedule(s));
  const size_t raw_bytes = (size_t)n * GT_RAW_FLOATS * sizeof(float);
  GT_TRY(gt_devbuf_reserve(s->raw_dev, raw_bytes));
  if (raw_bytes && raw_dev != s->raw_dev.p)
    HIP_TRY(hipMemcpyAsync(s->raw_dev.p, raw_dev, raw_bytes, hipMemcpyDeviceToDevice, st));
  GT_TRY(gt_devbuf_reserve(s->geom, (size_t)n * 48));
  GT_TRY(gt_devbuf_reserve(s->sh, (size_t)n * 192));
  GT_TRY(gt_devbuf_reserve(w->o2w, (size_t)n * 48));
  GT_TRY(gt_devbuf_reserve(w->w2o, (size_t)n * 48));
  GT_TRY(gt_devbuf_reserve(w->box, (size_t)n * 32));
  GT_TRY(gt_devbuf_reserve(w->part, (size_t)blocks_for(n) * PART_WORDS * 4));
  GT_TRY(gt_devbuf_reserve(w->lbox, (size_t)np * 32));
  GT_TRY(gt_devbuf_reserve(w->ctl, CTL_WORDS * sizeof(unsigned int)));
  unsigned int* ctl = (unsigned int*)w->ctl.p;
  k_init<<<1, BUILD_BLOCK, 0, st>>>(ctl);
  HIP_TRY(hipGetLastError());
  if (n > 0) {
    k_prep<<<blocks_for(n), BUILD_BLOCK, 0, st>>>((const float*)s->raw_dev.p, n, s->alpha_min, (float*)s->geom.p,
                                                  (float*)s->sh.p, (float*)w->o2w.p, (float*)w->w2o.p,
                                                  (float4*)w->box.p, (uint32_t*)w->part.p, ctl + CTL_BAD);
    HIP_TRY(hipGetLastError());
    k_bounds<<<1, 1024, 0, st>>>((const uint32_t*)w->part.p, blocks_for(n), ctl);
    HIP_TRY(hipGetLastError());
  }
  k_refit_prox<<<std::max(blocks_for(np), 1u), BUILD_BLOCK, 0, st>>>(np, (float4*)s->prox.p,
                                                                    (const float4*)w->w2o.p,
                                                                    (const float4*)w->box.p,
                                                                    (float4*)w->lbox.p, ctl);
  HIP_TRY(hipGetLastError());
  const uint32_t* order = s->ridentity ? nullptr : (const uint32_t*)s->rorder.p;
  for (size_t L = s->rlevels.size(); L-- > 1;) {
    const uint32_t a = s->rlevels[L - 1], b = s->rlevels[L];
    if (b > a) {
      k_refit4<<<blocks_for(b - a), BUILD_BLOCK, 0, st>>>((gt_node4*)s->nodes.p, order, a, b,
                                                          (const float4*)w->lbox.p);
      HIP_TRY(hipGetLastError());
    }
  }
  GT_TRY(read_ctl(w, st));
  GT_TRY(check_bad(w));
  s->device_built = true;
  s->h = HostScene{};
  if (w->host_ctl[CTL_STALE]) {
    *rebuilt = 1;
    return gt_device_build(s, st);
  }
  s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return GT_OK;
}

int gt_device_host_data(const gt_scene* s, float* packed, uint32_t* proxy_ids, float* o2w, float* w2o) {
  const uint32_t n = s->n, np = s->dev.n_prox;
  HIP_TRY(hipDeviceSynchronize());
  if (packed) {
    std::vector<float> geom((size_t)n * 12), sh((size_t)n * 48);
    if (n) {
      HIP_TRY(hipMemcpy(geom.data(), s->geom.p, geom.size() * 4, hipMemcpyDeviceToHost));
      HIP_TRY(hipMemcpy(sh.data(), s->sh.p, sh.size() * 4, hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < n; ++i) {
      float* o = packed + (size_t)i * 80;
      const float* g = &geom[(size_t)i * 12];
      std::memset(o, 0, 80 * sizeof(float));
      o[0] = g[0]; o[1] = g[1]; o[2] = g[2];
      o[4] = g[4]; o[5] = g[5]; o[6] = g[6];
      o[8] = g[8]; o[9] = g[9]; o[10] = g[10]; o[11] = g[11];
      o[12] = g[3];
      for (int j = 0; j < 16; ++j)
        for (int c = 0; c < 3; ++c) o[16 + 4 * j + c] = sh[(size_t)i * 48 + 3 * j + c];
    }
  }
  if (proxy_ids || o2w || w2o) {
    std::vector<float> prox((size_t)np * 16), m((size_t)n * 12);
    if (np) HIP_TRY(hipMemcpy(prox.data(), s->prox.p, prox.size() * 4, hipMemcpyDeviceToHost));
    if (n && o2w) HIP_TRY(hipMemcpy(m.data(), s->db->o2w.p, m.size() * 4, hipMemcpyDeviceToHost));
    std::vector<std::pair<uint32_t, uint32_t>> order(np); /* (gid, sorted position) */
    for (uint32_t p = 0; p < np; ++p) {
      uint32_t gid;
      std::memcpy(&gid, &prox[(size_t)p * 16 + 12], 4);
      order[p] = {gid, p};
    }
    std::sort(order.begin(), order.end());
    for (uint32_t q = 0; q < np; ++q) {
      const uint32_t gid = order[q].first, p = order[q].second;
      if (proxy_ids) proxy_ids[q] = gid;
      if (w2o) std::memcpy(w2o + (size_t)q * 12, &prox[(size_t)p * 16], 12 * sizeof(float));
      if (o2w) std::memcpy(o2w + (size_t)q * 12, &m[(size_t)gid * 12], 12 * sizeof(float));
    }
  }
  return GT_OK;
}
