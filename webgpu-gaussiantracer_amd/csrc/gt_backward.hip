/*
 * gt_backward.hip - hand-written adjoint of the forward Gaussian ray tracer.
 *
 * The reference has no backward pass (README.md:68-69 lists training as a
 * plan).  Per ray, the forward output is
 *   L = sum_i T_i alpha_i c_i,  T_1 = 1,  T_{i+1} = T_i (1 - alpha_i)
 * over the composited hits i (GS_shader.ts:258-268) with
 *   alpha_i = response_i * opacity_i           (computeResponse, 111-136)
 *   c_i     = max(0, 0.5 + sum_j Y_j(dir_i) sh_ij), dir_i = normalize(mu_i - o)
 * Given G = dLoss/dL per pixel, the adjoints are
 *   dLoss/dc_i     = G T_i alpha_i
 *   dLoss/dalpha_i = G . (T_i c_i - R_i / (1 - alpha_i)),  R_i = sum_{j>i} T_j alpha_j c_j
 * computed in two passes over the ray's checkpointed hit list, in FP64:
 * the first accumulates L, the second walks the hits front to back with
 * R_i = L - sum_{j<=i} T_j alpha_j c_j.  The hit lists (and which entries
 * were composited) are held fixed, as the forward's discontinuities in
 * them have no derivative.  Chain rule to the raw PLY parameters:
 *   mean, log-scale (scale = exp), raw quaternion (normalised), opacity
 *   logit (sigmoid), SH coefficients (f_dc, f_rest),
 * through the whitened-frame response exp(-|a x b|^2 / |b|^2),
 * a = S^-1 A (o - mu), b = S^-1 A d (A = transpose of the quaternion's
 * rotation, as the shader's mat3), and through dir_i for the SH.
 * Per-Gaussian gradients are accumulated with FP32 atomics in the
 * 59-float raw layout of the scene's input.
 */
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/gt.h"
#include "../../include/gt_fp32.h"
#include "gt_internal.h"

/* gt_scene.cpp */
void* gt_scene_bwd_scratch(gt_scene* s, int which, size_t bytes);
uint32_t* gt_scene_rc_cap(gt_scene* s);
uint32_t* gt_scene_rects(gt_scene* s, void* stream, size_t extra_words, size_t* words);
const gt_dev_scene* gt_scene_dev(const gt_scene* s);
const gt_tuning* gt_scene_tuning(const gt_scene* s);
int gt_validate_uniforms(const gt_uniforms* u, uint32_t width, uint32_t height);
void gt_fill_dev_frame(const gt_uniforms* u, uint32_t width, uint32_t height, gt_dev_frame* f);
uint32_t gt_frame_tiles(const gt_tuning* t, gt_dev_frame* f, uint32_t tile_base);

namespace {

#define SH_C0 0.28209479177387814
#define SH_C1 0.4886025119029199
#define SH_C2_0 1.0925484305920792
#define SH_C2_1 -1.0925484305920792
#define SH_C2_2 0.31539156525252005
#define SH_C2_3 -1.0925484305920792
#define SH_C2_4 0.5462742152960396
#define SH_C3_0 -0.5900435899266435
#define SH_C3_1 2.890611442640554
#define SH_C3_2 -0.4570457994644658
#define SH_C3_3 0.3731763325901154
#define SH_C3_4 -0.4570457994644658
#define SH_C3_5 1.445305721320277
#define SH_C3_6 -0.5900435899266435

/* real SH basis of GS_shader.ts:142-175 and its gradient w.r.t. the direction */
__device__ __forceinline__ void sh_basis(double x, double y, double z, double* b, double (*db)[3]) {
  const double xx = x * x, yy = y * y, zz = z * z;
  b[0] = SH_C0; db[0][0] = 0; db[0][1] = 0; db[0][2] = 0;
  b[1] = -SH_C1 * y; db[1][0] = 0; db[1][1] = -SH_C1; db[1][2] = 0;
  b[2] = SH_C1 * z; db[2][0] = 0; db[2][1] = 0; db[2][2] = SH_C1;
  b[3] = -SH_C1 * x; db[3][0] = -SH_C1; db[3][1] = 0; db[3][2] = 0;
  b[4] = SH_C2_0 * x * y; db[4][0] = SH_C2_0 * y; db[4][1] = SH_C2_0 * x; db[4][2] = 0;
  b[5] = SH_C2_1 * y * z; db[5][0] = 0; db[5][1] = SH_C2_1 * z; db[5][2] = SH_C2_1 * y;
  b[6] = SH_C2_2 * (2 * zz - xx - yy); db[6][0] = -2 * SH_C2_2 * x; db[6][1] = -2 * SH_C2_2 * y; db[6][2] = 4 * SH_C2_2 * z;
  b[7] = SH_C2_3 * x * z; db[7][0] = SH_C2_3 * z; db[7][1] = 0; db[7][2] = SH_C2_3 * x;
  b[8] = SH_C2_4 * (xx - yy); db[8][0] = 2 * SH_C2_4 * x; db[8][1] = -2 * SH_C2_4 * y; db[8][2] = 0;
  b[9] = SH_C3_0 * y * (3 * xx - yy); db[9][0] = 6 * SH_C3_0 * x * y; db[9][1] = SH_C3_0 * (3 * xx - 3 * yy); db[9][2] = 0;
  b[10] = SH_C3_1 * x * y * z; db[10][0] = SH_C3_1 * y * z; db[10][1] = SH_C3_1 * x * z; db[10][2] = SH_C3_1 * x * y;
  b[11] = SH_C3_2 * y * (4 * zz - xx - yy);
  db[11][0] = -2 * SH_C3_2 * x * y; db[11][1] = SH_C3_2 * (4 * zz - xx - 3 * yy); db[11][2] = 8 * SH_C3_2 * y * z;
  b[12] = SH_C3_3 * z * (2 * zz - 3 * xx - 3 * yy);
  db[12][0] = -6 * SH_C3_3 * x * z; db[12][1] = -6 * SH_C3_3 * y * z; db[12][2] = SH_C3_3 * (6 * zz - 3 * xx - 3 * yy);
  b[13] = SH_C3_4 * x * (4 * zz - xx - yy);
  db[13][0] = SH_C3_4 * (4 * zz - 3 * xx - yy); db[13][1] = -2 * SH_C3_4 * x * y; db[13][2] = 8 * SH_C3_4 * x * z;
  b[14] = SH_C3_5 * z * (xx - yy); db[14][0] = 2 * SH_C3_5 * x * z; db[14][1] = -2 * SH_C3_5 * y * z; db[14][2] = SH_C3_5 * (xx - yy);
  b[15] = SH_C3_6 * x * (xx - 3 * yy); db[15][0] = SH_C3_6 * (3 * xx - 3 * yy); db[15][1] = -6 * SH_C3_6 * x * y; db[15][2] = 0;
}

/* Per-Gaussian accumulators of one frame (GT_ACC floats each).  All rays
 * of a frame share the camera origin o, so the colour c and its SH basis
 * are per-Gaussian constants; a hit contributes
 *   [0..2]   G_ch T alpha                 (dLoss/dc before the clamp)
 *   [3]      dLoss/dalpha * response      (opacity logit)
 *   [4..6]   dLoss/dmu through the response
 *   [7..9]   dLoss/dlog-scale
 *   [10..13] dLoss/dq of the unit quaternion (before the normalisation)
 * each evaluated per hit in FP64 (the per-hit terms cancel strongly: a
 * sum of products formed after FP32 accumulation would not), and
 * k_finalize adds the colour terms and the normalisation. */
#define GT_ACC 16

struct Gauss {
  double mu[3], is[3], q[4], qn, op; /* is: 1 / scale, one divide per load */
  double A[3][3]; /* A[r][c]: applied rotation (GLSL R[c][r]) */
};

__device__ __forceinline__ void load_gauss(const gt_dev_scene& S, uint32_t id, Gauss& g) {
  const float4* G = reinterpret_cast<const float4*>(S.geom) + 3 * (size_t)id;
  const float4 g0 = G[0], g1 = G[1], g2 = G[2];
  g.mu[0] = g0.x; g.mu[1] = g0.y; g.mu[2] = g0.z; g.op = g0.w;
  g.is[0] = 1.0 / (double)g1.x; g.is[1] = 1.0 / (double)g1.y; g.is[2] = 1.0 / (double)g1.z; g.qn = g1.w;
  g.q[0] = g2.x; g.q[1] = g2.y; g.q[2] = g2.z; g.q[3] = g2.w;
  const double r = g.q[0], x = g.q[1], y = g.q[2], z = g.q[3];
  g.A[0][0] = 1 - 2 * (y * y + z * z); g.A[0][1] = 2 * (x * y + r * z); g.A[0][2] = 2 * (x * z - r * y);
  g.A[1][0] = 2 * (x * y - r * z); g.A[1][1] = 1 - 2 * (x * x + z * z); g.A[1][2] = 2 * (y * z + r * x);
  g.A[2][0] = 2 * (x * z + r * y); g.A[2][1] = 2 * (y * z - r * x); g.A[2][2] = 1 - 2 * (x * x + y * y);
}

/* whitened-frame response of one hit: a, b, c = a x b, 1 / |b|^2, e, exp(-e) */
struct Resp {
  double a[3], b[3], cr[3], ibb, e, resp;
};

__device__ __forceinline__ void eval_resp(const Gauss& g, const double* o, const double* d, Resp& h) {
#pragma clang fp contract(fast)
  double u[3];
  for (int i = 0; i < 3; ++i) u[i] = o[i] - g.mu[i];
  for (int rr = 0; rr < 3; ++rr) {
    h.a[rr] = (g.A[rr][0] * u[0] + g.A[rr][1] * u[1] + g.A[rr][2] * u[2]) * g.is[rr];
    h.b[rr] = (g.A[rr][0] * d[0] + g.A[rr][1] * d[1] + g.A[rr][2] * d[2]) * g.is[rr];
  }
  h.cr[0] = h.a[1] * h.b[2] - h.a[2] * h.b[1];
  h.cr[1] = h.a[2] * h.b[0] - h.a[0] * h.b[2];
  h.cr[2] = h.a[0] * h.b[1] - h.a[1] * h.b[0];
  h.ibb = 1.0 / (h.b[0] * h.b[0] + h.b[1] * h.b[1] + h.b[2] * h.b[2]);
  h.e = (h.cr[0] * h.cr[0] + h.cr[1] * h.cr[1] + h.cr[2] * h.cr[2]) * h.ibb;
  h.resp = gt_exp64(-h.e);
}

#define BW_TABLE_LOG2 9
#define BW_TABLE (1u << BW_TABLE_LOG2) /* LDS accumulator entries per block */
#define BW_PROBES 8
#define BW_EMPTY 0xFFFFFFFFu
#define BW_VALS 14   /* accumulators in use of GT_ACC */

/* Block-local accumulation: a 16x16-pixel block hits a few hundred distinct
 * Gaussians many times each (neighbouring rays cross the same Gaussians),
 * so the accumulators are summed in an LDS hash table keyed by Gaussian id
 * (linear probing; keys are only ever set, never removed) and flushed to
 * HBM once per block.  A Gaussian that finds no slot adds straight to HBM. */
__device__ __forceinline__ int table_slot(uint32_t* key, uint32_t id) {
  const uint32_t h = (id * 2654435761u) >> (32 - BW_TABLE_LOG2);
  for (int pr = 0; pr < BW_PROBES; ++pr) {
    const uint32_t e = (h + (uint32_t)pr) & (BW_TABLE - 1u);
    const uint32_t k = __hip_atomic_load(key + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == id) return (int)e;
    if (k == BW_EMPTY) {
      const uint32_t old = atomicCAS(key + e, BW_EMPTY, id);
      if (old == BW_EMPTY || old == id) return (int)e;
    }
  }
  return -1;
}

/* The value of lane ^ X, for X in {1, 2, 8, 16, 32}.  Within a row of 16
 * lanes DPP moves: quad_perm [1,0,3,2] (^1), quad_perm [2,3,0,1] (^2) and
 * row_ror:8 (^8).  Across rows the gfx950 swaps: v_permlane16_swap_b32
 * with both operands x gives x[row 0] in odd rows' first result and
 * x[row 1] in even rows' second result (^16); v_permlane32_swap_b32 the
 * same for halves (^32). */
template <uint32_t X>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x, uint32_t lane) {
  if constexpr (X == 1u) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, true);
  else if constexpr (X == 2u) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, true);
  else if constexpr (X == 8u) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xF, 0xF, true);
  else if constexpr (X == 16u) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return (lane & 16u) ? r[0] : r[1];
  } else {
    static_assert(X == 32u, "lane_xor: unsupported distance");
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return (lane & 32u) ? r[0] : r[1];
  }
}
/* Merges lane L with lane L ^ X when both hold the same key: the lower lane
 * adds the partner's values, the upper one clears its key.  Called with
 * every lane of the wave active. */
template <uint32_t X>
__device__ __forceinline__ void merge_pair(uint32_t lane, uint32_t& key, float* v) {
  const uint32_t pk = lane_xor<X>(key, lane);
  const bool same = pk == key && key != BW_EMPTY;
  const bool upper = (lane & X) != 0u;
#pragma unroll
  for (int k = 0; k < BW_VALS; ++k) {
    const float pv = __uint_as_float(lane_xor<X>(__float_as_uint(v[k]), lane));
    if (same && !upper) v[k] += pv;
  }
  if (same && upper) key = BW_EMPTY;
}
/* Dispatch order of the backward's blocks, most expensive first.  A
 * block's waves loop over their rays' composited ids, so its duration
 * follows the largest composited count of its 16x16 pixels (correlation
 * 0.92 over the blocks of the 1M-Gaussian 1080p frame).  Dispatched in
 * index order, the last expensive blocks started late and ran on alone:
 * the blocks kept 80 % of the 1024 block slots busy on average, and fell
 * below 90 % of them at 78 % of the launch.  In cost order they keep 96 %
 * busy to the end.  k_bw_cost writes each block's largest count;
 * k_bw_order, one workgroup, counting-sorts the blocks by it (descending,
 * clamped to BW_COST_LEVELS - 1) in LDS.  The order within one count is the
 * order of the LDS atomics: the gradient sums are float atomics of
 * unspecified order anyway.  (Keeping index order within coarser levels,
 * for neighbouring blocks to run together, measured no faster.) */
constexpr uint32_t BW_COST_LEVELS = 1024u;
__global__ __launch_bounds__(256) void k_bw_cost(const gt_dev_frame F, uint32_t y0, uint32_t y1,
                                                 uint32_t* __restrict__ cost) {
  const uint32_t bx_n = (F.width + 15u) / 16u;
  const uint32_t px = (blockIdx.x % bx_n) * 16u + (threadIdx.x & 15u);
  const uint32_t py = y0 + (blockIdx.x / bx_n) * 16u + (threadIdx.x >> 4);
  uint32_t n = 0;
  if (px < F.width && py < y1) {
    n = F.n_comp[py * F.width + px - F.comp_base];
    if (n > F.comp_cap) n = 0;  // k_backward skips such a ray
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(n, off);
    n = o > n ? o : n;
  }
  __shared__ uint32_t s_max[4];
  if ((threadIdx.x & 63u) == 0u) s_max[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t a = s_max[0] > s_max[1] ? s_max[0] : s_max[1];
    const uint32_t b = s_max[2] > s_max[3] ? s_max[2] : s_max[3];
    cost[blockIdx.x] = a > b ? a : b;
  }
}
__global__ __launch_bounds__(1024) void k_bw_order(const uint32_t* __restrict__ cost, uint32_t nb,
                                                   uint32_t* __restrict__ order) {
  __shared__ uint32_t s_hist[BW_COST_LEVELS];
  __shared__ uint32_t s_wave[BW_COST_LEVELS / 64u];
  const uint32_t t = threadIdx.x, lane = t & 63u;
  s_hist[t] = 0u;
  __syncthreads();
  for (uint32_t i = t; i < nb; i += 1024u) {
    const uint32_t c = cost[i];
    atomicAdd(&s_hist[c < BW_COST_LEVELS - 1u ? c : BW_COST_LEVELS - 1u], 1u);
  }
  __syncthreads();
  /* exclusive scan over the levels in descending order: thread t holds
   * level BW_COST_LEVELS - 1 - t */
  const uint32_t v = s_hist[BW_COST_LEVELS - 1u - t];
  uint32_t x = v;
  for (uint32_t off = 1; off < 64u; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63u) s_wave[t >> 6] = x;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t w = 0; w < (t >> 6); ++w) base += s_wave[w];
  s_hist[BW_COST_LEVELS - 1u - t] = base + x - v;
  __syncthreads();
  for (uint32_t i = t; i < nb; i += 1024u) {
    const uint32_t c = cost[i];
    order[atomicAdd(&s_hist[c < BW_COST_LEVELS - 1u ? c : BW_COST_LEVELS - 1u], 1u)] = i;
  }
}

/* One thread per ray; a block covers 16x16 pixels of rows [y0, y1) as four
 * 8x8 wave tiles, block order[blockIdx.x] (blockIdx.x without an order).
 * Two passes over the ray's composited ids, in FP64: the first accumulates
 * the pixel L, the second walks the hits front to back with
 * R_i = L - L_{<=i}. */
__global__ __launch_bounds__(256, 4) void k_backward(const gt_dev_scene S, const gt_dev_frame F,
                                                  const float4* __restrict__ dl_drgba,
                                                  const float4* __restrict__ colors, uint32_t y0, uint32_t y1,
                                                  const uint32_t* __restrict__ order, float* __restrict__ acc,
                                                  unsigned int* overflow) {
#pragma clang fp contract(fast)
  __shared__ uint32_t s_key[BW_TABLE];
  __shared__ float s_val[BW_VALS * BW_TABLE];
  for (uint32_t i = threadIdx.x; i < BW_TABLE; i += 256u) {
    s_key[i] = BW_EMPTY;
#pragma unroll
    for (int k = 0; k < BW_VALS; ++k) s_val[k * BW_TABLE + i] = 0.0f;
  }
  __syncthreads();
  const uint32_t bx_n = (F.width + 15u) / 16u;
  const uint32_t blk = order ? order[blockIdx.x] : blockIdx.x;
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t px = (blk % bx_n) * 16u + (w & 1u) * 8u + (lane & 7u);
  const uint32_t py = y0 + (blk / bx_n) * 16u + (w >> 1) * 8u + (lane >> 3);
  const uint32_t pix = py * F.width + px;
  const uint32_t slot = pix - F.comp_base;
  uint32_t n = 0;
  float4 gr = make_float4(0.f, 0.f, 0.f, 0.f);
  if (px < F.width && py < y1) {
    n = F.n_comp[slot];
    if (n > F.comp_cap) {
      if (overflow) atomicAdd(overflow, 1u);
      n = 0;
    }
    if (n) gr = dl_drgba[pix];
  }
  const double G[3] = {gr.x, gr.y, gr.z};
  if (G[0] == 0.0 && G[1] == 0.0 && G[2] == 0.0) n = 0;
  const gt_f3 df = gt_ray_dir(F.to_world, F.ax, F.ay, px, py, F.width, F.height);
  const double o[3] = {F.to_world[12], F.to_world[13], F.to_world[14]};
  const double d[3] = {df.x, df.y, df.z};
  const uint32_t* ids = F.comp_ids + (size_t)slot * F.comp_cap;
  /* pass 1: L */
  double L[3] = {0, 0, 0}, T = 1.0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t id = ids[i];
    Gauss g;
    load_gauss(S, id, g);
    Resp h;
    eval_resp(g, o, d, h);
    const double alpha = h.resp * g.op;
    const float4 c = colors[id];
    L[0] += T * alpha * c.x;
    L[1] += T * alpha * c.y;
    L[2] += T * alpha * c.z;
    T *= 1.0 - alpha;
  }
  /* pass 2: adjoints front to back.  The loop runs in step across the
   * wave (lanes past their last hit idle), so that the contributions of
   * lanes holding the same Gaussian can be merged before the atomics. */
  double Lp[3] = {0, 0, 0};
  T = 1.0;
  for (uint32_t i = 0; __ballot(i < n) != 0ull; ++i) {
    float v[BW_VALS];
#pragma unroll
    for (int k = 0; k < BW_VALS; ++k) v[k] = 0.0f;
    uint32_t key = BW_EMPTY;
    if (i < n) {
      const uint32_t id = ids[i];
      key = id;
      Gauss g;
      load_gauss(S, id, g);
      Resp h;
      eval_resp(g, o, d, h);
      const double alpha = h.resp * g.op;
      const float4 c4 = colors[id];
      const double c[3] = {c4.x, c4.y, c4.z};
      const double one_m = 1.0 - alpha;
      const double i_one_m = one_m > 0.0 ? 1.0 / one_m : 0.0;
      double dL_dalpha = 0.0;
      for (int ch = 0; ch < 3; ++ch) {
        Lp[ch] += T * alpha * c[ch];
        const double rest = L[ch] - Lp[ch];
        dL_dalpha += G[ch] * (T * c[ch] - rest * i_one_m);
        v[ch] = (float)(G[ch] * T * alpha);
      }
      v[3] = (float)(dL_dalpha * h.resp);
      /* response: resp = exp(-e), e = |a x b|^2 / |b|^2 */
      const double dL_de = -h.resp * g.op * dL_dalpha;
      const double k2 = 2.0 * h.ibb;
      const double bxc[3] = {h.b[1] * h.cr[2] - h.b[2] * h.cr[1], h.b[2] * h.cr[0] - h.b[0] * h.cr[2],
                             h.b[0] * h.cr[1] - h.b[1] * h.cr[0]};
      const double cxa[3] = {h.cr[1] * h.a[2] - h.cr[2] * h.a[1], h.cr[2] * h.a[0] - h.cr[0] * h.a[2],
                             h.cr[0] * h.a[1] - h.cr[1] * h.a[0]};
      double dL_da[3], dL_db[3];
      for (int r = 0; r < 3; ++r) {
        dL_da[r] = dL_de * k2 * bxc[r];
        dL_db[r] = dL_de * k2 * (cxa[r] - h.e * h.b[r]);
      }
      /* a = S^-1 A (o - mu), b = S^-1 A d */
      double M[3][3];
      for (int r = 0; r < 3; ++r) {
        const double is = g.is[r];
        double sum = 0.0;
        for (int cc = 0; cc < 3; ++cc) {
          const double t = dL_da[r] * (o[cc] - g.mu[cc]) + dL_db[r] * d[cc];
          M[r][cc] = is * t;
          sum += g.A[r][cc] * t;
        }
        v[7 + r] = (float)(-is * sum); /* d(1/s)/dlog s = -1/s */
      }
      for (int cc = 0; cc < 3; ++cc)
        v[4 + cc] = (float)(-(dL_da[0] * g.A[0][cc] * g.is[0] + dL_da[1] * g.A[1][cc] * g.is[1] +
                              dL_da[2] * g.A[2][cc] * g.is[2]));
      /* quaternion (r, x, y, z) of A */
      const double qr = g.q[0], qx = g.q[1], qy = g.q[2], qz = g.q[3];
      v[10] = (float)(2 * (qz * M[0][1] - qy * M[0][2] - qz * M[1][0] + qx * M[1][2] + qy * M[2][0] - qx * M[2][1]));
      v[11] = (float)(2 * (qy * M[0][1] + qz * M[0][2] + qy * M[1][0] - 2 * qx * M[1][1] + qr * M[1][2] +
                           qz * M[2][0] - qr * M[2][1] - 2 * qx * M[2][2]));
      v[12] = (float)(2 * (-2 * qy * M[0][0] + qx * M[0][1] - qr * M[0][2] + qx * M[1][0] + qz * M[1][2] +
                           qr * M[2][0] + qz * M[2][1] - 2 * qy * M[2][2]));
      v[13] = (float)(2 * (-2 * qz * M[0][0] + qr * M[0][1] + qx * M[0][2] - qr * M[1][0] - 2 * qz * M[1][1] +
                           qy * M[1][2] + qx * M[2][0] + qy * M[2][1]));
      T *= one_m;
    }
    /* Neighbouring rays composite the same Gaussian at the same step, and
     * same-address LDS atomics serialise.  Lanes (x, y) = (lane & 7,
     * lane >> 3) of the 8x8 ray tile merge with their partners at x ^ 1,
     * x ^ 2, y ^ 1, y ^ 2 and y ^ 4 when the keys match: the lower lane
     * takes the sum and the upper one drops out (key BW_EMPTY). */
    merge_pair<1u>(lane, key, v);
    merge_pair<2u>(lane, key, v);
    merge_pair<8u>(lane, key, v);
    merge_pair<16u>(lane, key, v);
    merge_pair<32u>(lane, key, v);
    if (key != BW_EMPTY) {
      const int e = table_slot(s_key, key);
#pragma unroll
      for (int k = 0; k < BW_VALS; ++k) {
        if (v[k] == 0.0f) continue;
        if (e >= 0)
          atomicAdd(&s_val[k * BW_TABLE + (uint32_t)e], v[k]);
        else
          unsafeAtomicAdd(acc + (size_t)key * GT_ACC + k, v[k]);
      }
    }
  }
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < BW_TABLE; e += 256u) {
    const uint32_t id = s_key[e];
    if (id == BW_EMPTY) continue;
    float* A = acc + (size_t)id * GT_ACC;
#pragma unroll
    for (int k = 0; k < BW_VALS; ++k) {
      const float v = s_val[k * BW_TABLE + e];
      if (v != 0.0f) unsafeAtomicAdd(A + k, v);
    }
  }
}

__global__ __launch_bounds__(256) void k_zero_f32(float* __restrict__ p, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256ull) p[i] = 0.0f;
}

/* The chain rule of one Gaussian from its accumulators ac to the raw PLY
 * parameters (mean, f_dc/f_rest, opacity logit, log-scale, raw quaternion),
 * added into its gradient row out (59 floats). */
__device__ __forceinline__ void finalize_one(const gt_dev_scene& S, const gt_dev_frame& F, uint32_t id,
                                             const float* ac, float* out) {
  Gauss g;
  load_gauss(S, id, g);
  const double o[3] = {F.to_world[12], F.to_world[13], F.to_world[14]};
  const double u[3] = {o[0] - g.mu[0], o[1] - g.mu[1], o[2] - g.mu[2]};
  /* colour: SH coefficients and the direction normalize(mu - o) */
  const double dist = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const double dir[3] = {-u[0] / dist, -u[1] / dist, -u[2] / dist};
  const int nb = F.sh_degree == 0 ? 1 : F.sh_degree == 1 ? 4 : F.sh_degree == 2 ? 9 : 16;
  double bs[16], dbs[16][3];
  sh_basis(dir[0], dir[1], dir[2], bs, dbs);
  const float* sh = S.sh + 48 * (size_t)id;
  double dL_ddir[3] = {0, 0, 0};
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    double v = 0.5;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (j < nb) v += bs[j] * (double)sh[3 * j + ch];
    const double dc = v > 0.0 ? (double)ac[ch] : 0.0;
    if (dc == 0.0) continue;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      if (j >= nb) continue;
      const int idx = j == 0 ? 3 + ch : 6 + ch * 15 + (j - 1);
      out[idx] += (float)(dc * bs[j]);
      const double shv = sh[3 * j + ch];
      for (int k = 0; k < 3; ++k) dL_ddir[k] += dc * shv * dbs[j][k];
    }
  }
  double dL_dmu[3];
  {
    const double dd = dL_ddir[0] * dir[0] + dL_ddir[1] * dir[1] + dL_ddir[2] * dir[2];
    for (int k = 0; k < 3; ++k) dL_dmu[k] = (dL_ddir[k] - dir[k] * dd) / dist;
  }
  /* opacity logit: alpha = response * sigmoid(logit) */
  out[51] += (float)((double)ac[3] * g.op * (1.0 - g.op));
  for (int k = 0; k < 3; ++k) {
    out[k] += (float)(dL_dmu[k] + (double)ac[4 + k]);
    out[52 + k] += ac[7 + k];
  }
  /* the normalisation q = raw / |raw| */
  const double dq[4] = {ac[10], ac[11], ac[12], ac[13]};
  const double qd = dq[0] * g.q[0] + dq[1] * g.q[1] + dq[2] * g.q[2] + dq[3] * g.q[3];
  for (int k = 0; k < 4; ++k) out[55 + k] += (float)((dq[k] - g.q[k] * qd) / g.qn);
}

/* One thread per Gaussian, FIN_BLOCK Gaussians per block.  The gradient
 * rows of the block's Gaussians that have accumulators (a quarter of the
 * 1M-Gaussian scene's at 1080p) are copied to LDS, updated there and
 * written back; the other rows are neither read nor written.  Each wave
 * copies its own rows as one flat run of 59-float rows, 64 consecutive
 * floats per access instruction (two or three rows), so the accesses stay
 * coalesced; a thread updating its own row in HBM made each access
 * instruction touch 64 rows 236 B apart.  (Copying every row of a block
 * with any accumulator moved 0.47 GB instead of 0.13 GB.  Staging the SH
 * records in LDS as well halved the blocks per CU and doubled the time.) */
#define FIN_BLOCK 128u
__device__ __forceinline__ void fin_copy_rows(float* dst, const float* src, const uint8_t* rows, uint32_t nrows,
                                              uint32_t lane) {
  const uint32_t nf = nrows * GT_RAW_FLOATS;
#pragma unroll 4
  for (uint32_t j = lane; j < nf; j += 64u) {
    const uint32_t r = j / GT_RAW_FLOATS;
    const uint32_t e = (uint32_t)rows[r] * GT_RAW_FLOATS + (j - r * GT_RAW_FLOATS);
    dst[e] = src[e];
  }
}
__global__ __launch_bounds__(FIN_BLOCK) void k_finalize(const gt_dev_scene S, const gt_dev_frame F,
                                                        const float* __restrict__ acc, float* __restrict__ grads) {
  __shared__ float s_g[FIN_BLOCK * GT_RAW_FLOATS];
  __shared__ uint8_t s_rows[FIN_BLOCK];
  const uint32_t base = blockIdx.x * FIN_BLOCK;
  const uint32_t cnt = S.n - base < FIN_BLOCK ? S.n - base : FIN_BLOCK;
  const uint32_t id = base + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63u, w0 = threadIdx.x & ~63u;
  float ac[GT_ACC] = {};
  bool any = false;
  if (threadIdx.x < cnt) {
    const float4* A4 = reinterpret_cast<const float4*>(acc + (size_t)id * GT_ACC);
    const float4 a0 = A4[0], a1 = A4[1], a2 = A4[2], a3 = A4[3];
    const float v[GT_ACC] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w,
                             a2.x, a2.y, a2.z, a2.w, a3.x, a3.y, a3.z, a3.w};
    for (int i = 0; i < GT_ACC; ++i) {
      ac[i] = v[i];
      any |= v[i] != 0.0f;
    }
  }
  /* this wave's rows with accumulators, compacted in lane order */
  const uint64_t m = __ballot(any);
  if (!m) return;
  const uint32_t nrows = (uint32_t)__popcll(m);
  if (any) s_rows[w0 + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint8_t)threadIdx.x;
  __builtin_amdgcn_wave_barrier();
  float* const gblk = grads + (size_t)base * GT_RAW_FLOATS;
  fin_copy_rows(s_g, gblk, s_rows + w0, nrows, lane);
  __builtin_amdgcn_wave_barrier();
  if (any) finalize_one(S, F, id, ac, s_g + threadIdx.x * GT_RAW_FLOATS);
  __builtin_amdgcn_wave_barrier();
  fin_copy_rows(gblk, s_g, s_rows + w0, nrows, lane);
}

__global__ void k_zero_u32(uint32_t* p) {
  if (threadIdx.x == 0) p[0] = 0u;
}

__global__ __launch_bounds__(256) void k_max_u32(const uint32_t* __restrict__ v, uint32_t n, uint32_t* out) {
  uint32_t m = 0;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) m = v[i] > m ? v[i] : m;
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(m, off);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63u) == 0 && m) atomicMax(out, m);
}

}  // namespace

/* Backward of one frame.  All pointers are device pointers.  Per-Gaussian
 * accumulators live in the scene's backward scratch, so calls on one scene
 * must not overlap.  Checkpoint
 * mode: comp_ids / n_comp / comp_cap from a gt_render with those outputs.
 * Rays that composited more than comp_cap hits are skipped and counted in
 * *overflow.  Recompute mode (comp_ids == NULL): the forward is re-run over
 * bands of rows into a bounded scratch checkpoint (about `budget_bytes`),
 * each band followed by its backward; the scratch grows as needed, so no ray
 * overflows.  Recompute mode returns after the stream has drained. */
extern "C" int gt_backward(gt_scene* s, const gt_uniforms* u, uint32_t width, uint32_t height, const float* dl_drgba,
                           const uint32_t* comp_ids, const uint32_t* n_comp, uint32_t comp_cap, float* grads,
                           unsigned int* overflow, uint64_t budget_bytes, void* stream) {
  if (!s || !dl_drgba || !grads) { gt_set_error("gt_backward: scene, dl_drgba and grads are required"); return GT_E_INVALID; }
  int rc = gt_validate_uniforms(u, width, height);
  if (rc != GT_OK) return rc;
  if (u->debug_mode != 0) { gt_set_error("gt_backward: debugMode must be 0"); return GT_E_INVALID; }
  if (comp_ids && (!n_comp || comp_cap == 0)) {
    gt_set_error("gt_backward: checkpoint needs comp_ids, n_comp and comp_cap");
    return GT_E_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  const gt_dev_scene* S = gt_scene_dev(s);
  if (S->n == 0) return GT_OK;
  gt_dev_frame f;
  gt_fill_dev_frame(u, width, height, &f);
  const uint32_t R = width * height;
  /* per-Gaussian accumulators (zeroed by a kernel, so the sequence can be
   * captured in a graph) and this camera's colours */
  /* and the blocks' costs and dispatch order (the frame's block count each) */
  const uint32_t frame_blocks = ((width + 15u) / 16u) * ((height + 15u) / 16u);
  float* acc = (float*)gt_scene_bwd_scratch(
      s, 0, (size_t)S->n * (GT_ACC + 4) * sizeof(float) + (size_t)frame_blocks * 2u * sizeof(uint32_t));
  if (!acc) return GT_E_NOMEM;
  uint32_t* bw_cost = reinterpret_cast<uint32_t*>(acc + (size_t)S->n * (GT_ACC + 4));
  uint32_t* bw_order = bw_cost + frame_blocks;
  const float4* colors = reinterpret_cast<const float4*>(acc + (size_t)S->n * GT_ACC);
  {
    const size_t nf = (size_t)S->n * GT_ACC;
    const size_t blocks = (nf + 255) / 256 < 2048 ? (nf + 255) / 256 : 2048;
    hipLaunchKernelGGL(k_zero_f32, dim3((uint32_t)blocks), dim3(256), 0, st, acc, nf);
  }
  if ((rc = gt_launch_colors(S, &f, acc + (size_t)S->n * GT_ACC, stream)) != GT_OK) return rc;
  auto band_backward = [&](uint32_t y0, uint32_t y1) -> hipError_t {
    const uint32_t blocks = ((width + 15u) / 16u) * ((y1 - y0 + 15u) / 16u);
    hipLaunchKernelGGL(k_bw_cost, dim3(blocks), dim3(256), 0, st, f, y0, y1, bw_cost);
    hipLaunchKernelGGL(k_bw_order, dim3(1), dim3(1024), 0, st, (const uint32_t*)bw_cost, blocks, bw_order);
    hipLaunchKernelGGL(k_backward, dim3(blocks), dim3(256), 0, st, *S, f, (const float4*)dl_drgba, colors,
                       y0, y1, (const uint32_t*)bw_order, acc, overflow);
    return hipGetLastError();
  };
  auto finalize = [&]() -> hipError_t {
    hipLaunchKernelGGL(k_finalize, dim3((S->n + FIN_BLOCK - 1u) / FIN_BLOCK), dim3(FIN_BLOCK), 0, st, *S, f,
                       (const float*)acc, grads);
    return hipGetLastError();
  };
  if (comp_ids) {
    f.comp_ids = const_cast<uint32_t*>(comp_ids);
    f.n_comp = const_cast<uint32_t*>(n_comp);
    f.comp_cap = comp_cap;
    f.comp_base = 0;
    hipError_t e = band_backward(0, height);
    if (e == hipSuccess) e = finalize();
    if (e != hipSuccess) { gt_set_error("backward launch failed: %s", hipGetErrorString(e)); return GT_E_HIP; }
    return GT_OK;
  }
  /* recompute: bands of whole tile rows (8x8 tiles).  Each band's forward
   * writes its composited ids into the scratch checkpoint; if a ray of the
   * band composited more than the capacity, the capacity grows (the band
   * shrinks to stay within the budget, down to one tile row) and the band's
   * forward runs again before its backward.  The scratch and the capacity
   * the frame needed stay with the scene for the next call. */
  uint32_t* last_cap = gt_scene_rc_cap(s);
  uint32_t cap = comp_cap ? comp_cap : 64u;
  if (*last_cap > cap) cap = *last_cap;
  gt_tuning tile8{};
  tile8.path = 3;
  tile8.tile_w = 8;
  tile8.stack_lds = gt_scene_tuning(s)->stack_lds;
  tile8.rects = gt_scene_tuning(s)->rects;
  tile8.colors = gt_scene_tuning(s)->colors;
  /* the band forwards composite the colours computed above (the same
   * values their SH evaluation would give) */
  if (tile8.colors != 2u) f.colors = reinterpret_cast<const float*>(colors);
  (void)gt_frame_tiles(&tile8, &f, 0);
  const uint32_t tiles_y = (height + 7u) / 8u;
  const uint64_t budget = budget_bytes ? budget_bytes : (1ull << 30);
  const size_t fixed = 256 + (size_t)R * 4u + (size_t)R * 16u; /* max, counts, image */
  char* base = nullptr;
  hipError_t e = hipSuccess;
  uint32_t band = 0;
  for (uint32_t ty0 = 0; ty0 < tiles_y && rc == GT_OK && e == hipSuccess;) {
    if (band == 0) {
      /* (re)size the band and its checkpoint for the current capacity; the
       * stream is idle here (first band, or after the max read-back) */
      const uint64_t per_row = (uint64_t)width * 8u * cap * 4u;
      band = (uint32_t)(budget / per_row);
      if (band < 1) band = 1;
      if (band > tiles_y) band = tiles_y;
      base = (char*)gt_scene_bwd_scratch(s, 1, fixed + (size_t)band * 8u * width * cap * 4u);
      if (!base) { rc = GT_E_NOMEM; break; }
      f.rgba = (float*)(base + 256 + (size_t)R * 4u);
      f.n_comp = (uint32_t*)(base + 256);
      f.comp_ids = (uint32_t*)(base + fixed);
      f.comp_cap = cap;
    }
    uint32_t* dmax = (uint32_t*)base;
    const uint32_t ty1 = ty0 + band < tiles_y ? ty0 + band : tiles_y;
    const uint32_t y0 = ty0 * 8u, y1 = ty1 * 8u < height ? ty1 * 8u : height;
    const uint32_t npix = (y1 - y0) * width;
    f.comp_base = y0 * width;
    size_t rwords = 0;
    uint32_t* rbuf = gt_rects_wanted(&tile8, f.width, f.height, S->n_prox) ? gt_scene_rects(s, stream, 0, &rwords)
                                                                            : nullptr;
    rc = gt_launch_forward(S, &f, nullptr, 1, ty1 * f.tiles_x, &tile8, stream, ty0 * f.tiles_x, nullptr, rbuf,
                           rwords);
    if (rc != GT_OK) break;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, dmax);
    hipLaunchKernelGGL(k_max_u32, dim3((npix + 255) / 256 < 1024 ? (npix + 255) / 256 : 1024), dim3(256), 0, st,
                       (const uint32_t*)f.n_comp, npix, dmax);
    uint32_t mx = 0;
    e = hipMemcpyAsync(&mx, dmax, 4u, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) break;
    if (mx > cap) {
      uint32_t c = 1;
      while (c < mx) c <<= 1;
      cap = c;
      *last_cap = cap;
      band = 0; /* resize, then render this band again */
      continue;
    }
    e = band_backward(y0, y1);
    if (e != hipSuccess) break;
    ty0 = ty1;
  }
  if (rc == GT_OK && e == hipSuccess) e = finalize();
  const hipError_t es = hipStreamSynchronize(st);
  if (rc == GT_OK && e != hipSuccess) {
    gt_set_error("gt_backward (recompute): %s", hipGetErrorString(e));
    rc = GT_E_HIP;
  }
  if (rc == GT_OK && es != hipSuccess) {
    gt_set_error("gt_backward: %s", hipGetErrorString(es));
    rc = GT_E_HIP;
  }
  return rc;
}
