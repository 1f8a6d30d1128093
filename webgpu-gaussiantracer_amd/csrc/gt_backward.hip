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

/* Everything the two passes need about one hit. */
struct Hit {
  double mu[3], s[3], q[4], qn, op;
  double A[3][3]; /* A[r][c]: applied rotation (GLSL R[c][r]) */
  double u[3], a[3], b[3], cr[3], bb, e, resp, alpha;
  double dir[3], dist;
  double v[3], c[3];
};

__device__ __forceinline__ void eval_hit(const gt_dev_scene& S, uint32_t id, const double* o, const double* d,
                                         int nb, Hit& h) {
  const float4* G = reinterpret_cast<const float4*>(S.geom) + 3 * (size_t)id;
  const float4 g0 = G[0], g1 = G[1], g2 = G[2];
  h.mu[0] = g0.x; h.mu[1] = g0.y; h.mu[2] = g0.z; h.op = g0.w;
  h.s[0] = g1.x; h.s[1] = g1.y; h.s[2] = g1.z; h.qn = g1.w;
  h.q[0] = g2.x; h.q[1] = g2.y; h.q[2] = g2.z; h.q[3] = g2.w;
  const double r = h.q[0], x = h.q[1], y = h.q[2], z = h.q[3];
  h.A[0][0] = 1 - 2 * (y * y + z * z); h.A[0][1] = 2 * (x * y + r * z); h.A[0][2] = 2 * (x * z - r * y);
  h.A[1][0] = 2 * (x * y - r * z); h.A[1][1] = 1 - 2 * (x * x + z * z); h.A[1][2] = 2 * (y * z + r * x);
  h.A[2][0] = 2 * (x * z + r * y); h.A[2][1] = 2 * (y * z - r * x); h.A[2][2] = 1 - 2 * (x * x + y * y);
  for (int i = 0; i < 3; ++i) h.u[i] = o[i] - h.mu[i];
  for (int rr = 0; rr < 3; ++rr) {
    h.a[rr] = (h.A[rr][0] * h.u[0] + h.A[rr][1] * h.u[1] + h.A[rr][2] * h.u[2]) / h.s[rr];
    h.b[rr] = (h.A[rr][0] * d[0] + h.A[rr][1] * d[1] + h.A[rr][2] * d[2]) / h.s[rr];
  }
  h.cr[0] = h.a[1] * h.b[2] - h.a[2] * h.b[1];
  h.cr[1] = h.a[2] * h.b[0] - h.a[0] * h.b[2];
  h.cr[2] = h.a[0] * h.b[1] - h.a[1] * h.b[0];
  h.bb = h.b[0] * h.b[0] + h.b[1] * h.b[1] + h.b[2] * h.b[2];
  h.e = (h.cr[0] * h.cr[0] + h.cr[1] * h.cr[1] + h.cr[2] * h.cr[2]) / h.bb;
  h.resp = gt_exp64(-h.e);
  h.alpha = h.resp * h.op;
  double w[3] = {h.mu[0] - o[0], h.mu[1] - o[1], h.mu[2] - o[2]};
  h.dist = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (int i = 0; i < 3; ++i) h.dir[i] = w[i] / h.dist;
  double bs[16], dbs[16][3];
  sh_basis(h.dir[0], h.dir[1], h.dir[2], bs, dbs);
  const float* sh = S.sh + 48 * (size_t)id;
  for (int ch = 0; ch < 3; ++ch) {
    double acc = 0.5;
    for (int j = 0; j < nb; ++j) acc += bs[j] * (double)sh[3 * j + ch];
    h.v[ch] = acc;
    h.c[ch] = acc > 0 ? acc : 0.0;
  }
}

__device__ __forceinline__ void add(float* p, double v) {
  if (v != 0.0) unsafeAtomicAdd(p, (float)v);
}

/* dLoss/dparams of one hit given dLoss/dalpha and dLoss/dc (per channel,
 * already including the max(0, .) clamp); accumulates into g (59 floats). */
__device__ __forceinline__ void scatter_hit(const gt_dev_scene& S, uint32_t id, const double* o, const double* d,
                                            int nb, const Hit& h, double dL_dalpha, const double* dL_dc, float* g) {
  /* SH coefficients and the SH direction */
  double bs[16], dbs[16][3];
  sh_basis(h.dir[0], h.dir[1], h.dir[2], bs, dbs);
  const float* sh = S.sh + 48 * (size_t)id;
  double dL_ddir[3] = {0, 0, 0};
  for (int ch = 0; ch < 3; ++ch) {
    if (dL_dc[ch] == 0.0) continue;
    for (int j = 0; j < nb; ++j) {
      const int idx = j == 0 ? 3 + ch : 6 + ch * 15 + (j - 1);
      add(g + idx, dL_dc[ch] * bs[j]);
      const double shv = sh[3 * j + ch];
      for (int k = 0; k < 3; ++k) dL_ddir[k] += dL_dc[ch] * shv * dbs[j][k];
    }
  }
  double dL_dmu[3];
  {
    const double dd = dL_ddir[0] * h.dir[0] + dL_ddir[1] * h.dir[1] + dL_ddir[2] * h.dir[2];
    for (int k = 0; k < 3; ++k) dL_dmu[k] = (dL_ddir[k] - h.dir[k] * dd) / h.dist;
  }
  /* opacity logit: alpha = resp * op, op = sigmoid(logit) */
  add(g + 51, dL_dalpha * h.resp * h.op * (1.0 - h.op));
  /* response: resp = exp(-e), e = |a x b|^2 / |b|^2 */
  const double dL_de = -h.resp * h.op * dL_dalpha;
  double dL_da[3], dL_db[3];
  {
    const double k2 = 2.0 / h.bb;
    const double bxc[3] = {h.b[1] * h.cr[2] - h.b[2] * h.cr[1], h.b[2] * h.cr[0] - h.b[0] * h.cr[2],
                           h.b[0] * h.cr[1] - h.b[1] * h.cr[0]};
    const double cxa[3] = {h.cr[1] * h.a[2] - h.cr[2] * h.a[1], h.cr[2] * h.a[0] - h.cr[0] * h.a[2],
                           h.cr[0] * h.a[1] - h.cr[1] * h.a[0]};
    for (int k = 0; k < 3; ++k) {
      dL_da[k] = dL_de * k2 * bxc[k];
      dL_db[k] = dL_de * (k2 * cxa[k] - k2 * h.e * h.b[k]);
    }
  }
  /* a = diag(1/s) A u, b = diag(1/s) A d, u = o - mu */
  double dL_dA[3][3], dL_dlogs[3];
  for (int rr = 0; rr < 3; ++rr) {
    const double is = 1.0 / h.s[rr];
    double sum = 0.0;
    for (int cc = 0; cc < 3; ++cc) {
      const double t = dL_da[rr] * h.u[cc] + dL_db[rr] * d[cc];
      dL_dA[rr][cc] = is * t;
      sum += h.A[rr][cc] * t;
    }
    dL_dlogs[rr] = -is * sum; /* d(1/s)/dlog s = -1/s */
  }
  for (int cc = 0; cc < 3; ++cc) {
    double du = 0.0;
    for (int rr = 0; rr < 3; ++rr) du += dL_da[rr] * h.A[rr][cc] / h.s[rr];
    dL_dmu[cc] -= du;
  }
  add(g + 0, dL_dmu[0]);
  add(g + 1, dL_dmu[1]);
  add(g + 2, dL_dmu[2]);
  add(g + 52, dL_dlogs[0]);
  add(g + 53, dL_dlogs[1]);
  add(g + 54, dL_dlogs[2]);
  /* quaternion (r, x, y, z) of A, then the normalisation q = raw / |raw| */
  const double r = h.q[0], x = h.q[1], y = h.q[2], z = h.q[3];
  const double (*M)[3] = dL_dA;
  double dq[4];
  dq[0] = 2 * (z * M[0][1] - y * M[0][2] - z * M[1][0] + x * M[1][2] + y * M[2][0] - x * M[2][1]);
  dq[1] = 2 * (y * M[0][1] + z * M[0][2] + y * M[1][0] - 2 * x * M[1][1] + r * M[1][2] + z * M[2][0] - r * M[2][1] -
               2 * x * M[2][2]);
  dq[2] = 2 * (-2 * y * M[0][0] + x * M[0][1] - r * M[0][2] + x * M[1][0] + z * M[1][2] + r * M[2][0] + z * M[2][1] -
               2 * y * M[2][2]);
  dq[3] = 2 * (-2 * z * M[0][0] + r * M[0][1] + x * M[0][2] - r * M[1][0] - 2 * z * M[1][1] + y * M[1][2] +
               x * M[2][0] + y * M[2][1]);
  const double qd = dq[0] * r + dq[1] * x + dq[2] * y + dq[3] * z;
  for (int k = 0; k < 4; ++k) add(g + 55 + k, (dq[k] - h.q[k] * qd) / h.qn);
}

__global__ __launch_bounds__(256) void k_backward(const gt_dev_scene S, const gt_dev_frame F,
                                                  const float4* __restrict__ dl_drgba, uint32_t pix0, uint32_t npix,
                                                  float* __restrict__ grads, unsigned int* overflow) {
  const uint32_t li = blockIdx.x * 256u + threadIdx.x;
  if (li >= npix) return;
  const uint32_t pix = pix0 + li;
  const uint32_t slot = pix - F.comp_base;
  const uint32_t n = F.n_comp[slot];
  if (n == 0) return;
  if (n > F.comp_cap) {
    if (overflow) atomicAdd(overflow, 1u);
    return;
  }
  const float4 gr = dl_drgba[pix];
  const double G[3] = {gr.x, gr.y, gr.z};
  if (G[0] == 0.0 && G[1] == 0.0 && G[2] == 0.0) return;
  const uint32_t px = pix % F.width, py = pix / F.width;
  const gt_f3 df = gt_ray_dir(F.to_world, F.ax, F.ay, px, py, F.width, F.height);
  const double o[3] = {F.to_world[12], F.to_world[13], F.to_world[14]};
  const double d[3] = {df.x, df.y, df.z};
  const int nb = F.sh_degree == 0 ? 1 : F.sh_degree == 1 ? 4 : F.sh_degree == 2 ? 9 : 16;
  const uint32_t* ids = F.comp_ids + (size_t)slot * F.comp_cap;
  /* pass 1: L */
  double L[3] = {0, 0, 0}, T = 1.0;
  for (uint32_t i = 0; i < n; ++i) {
    Hit h;
    eval_hit(S, ids[i], o, d, nb, h);
    for (int ch = 0; ch < 3; ++ch) L[ch] += T * h.alpha * h.c[ch];
    T *= 1.0 - h.alpha;
  }
  /* pass 2: adjoints front to back */
  double Lp[3] = {0, 0, 0};
  T = 1.0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t id = ids[i];
    Hit h;
    eval_hit(S, id, o, d, nb, h);
    double dL_dalpha = 0.0, dL_dc[3];
    for (int ch = 0; ch < 3; ++ch) {
      Lp[ch] += T * h.alpha * h.c[ch];
      const double rest = L[ch] - Lp[ch];
      const double one_m = 1.0 - h.alpha;
      dL_dalpha += G[ch] * (T * h.c[ch] - (one_m > 0.0 ? rest / one_m : 0.0));
      dL_dc[ch] = h.v[ch] > 0.0 ? G[ch] * T * h.alpha : 0.0;
    }
    scatter_hit(S, id, o, d, nb, h, dL_dalpha, dL_dc, grads + (size_t)id * GT_RAW_FLOATS);
    T *= 1.0 - h.alpha;
  }
}

}  // namespace

/* Backward of one frame.  All pointers are device pointers.  Checkpoint
 * mode: comp_ids / n_comp / comp_cap from a gt_render with those outputs.
 * Recompute mode (comp_ids == NULL): the forward is re-run over bands of rows
 * into a bounded scratch checkpoint (at most `budget_bytes`), each band
 * followed by its backward. */
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
  gt_dev_frame f;
  gt_fill_dev_frame(u, width, height, &f);
  const uint32_t R = width * height;
  if (comp_ids) {
    f.comp_ids = const_cast<uint32_t*>(comp_ids);
    f.n_comp = const_cast<uint32_t*>(n_comp);
    f.comp_cap = comp_cap;
    f.comp_base = 0;
    hipLaunchKernelGGL(k_backward, dim3((R + 255) / 256), dim3(256), 0, st, *S, f, (const float4*)dl_drgba, 0u, R,
                       grads, overflow);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { gt_set_error("backward launch failed: %s", hipGetErrorString(e)); return GT_E_HIP; }
    return GT_OK;
  }
  /* recompute: bands of whole tile rows (8x8 tiles) */
  const uint32_t cap = comp_cap ? comp_cap : 256u;
  gt_tuning tile8{};
  tile8.path = 1;
  tile8.tile_w = 8;
  tile8.stack_lds = gt_scene_tuning(s)->stack_lds;
  const uint32_t tiles = gt_frame_tiles(&tile8, &f, 0);
  const uint32_t tiles_y = (height + 7u) / 8u;
  const uint64_t per_row = (uint64_t)width * 8u * (cap + 1u) * 4u;
  const uint64_t budget = budget_bytes ? budget_bytes : (256ull << 20);
  uint32_t band = (uint32_t)(budget / per_row);
  if (band < 1) band = 1;
  if (band > tiles_y) band = tiles_y;
  const uint64_t band_px = (uint64_t)band * 8u * width;
  void *ids_buf = nullptr, *cnt_buf = nullptr, *img = nullptr;
  hipError_t e = hipMalloc(&ids_buf, band_px * cap * 4u);
  if (e == hipSuccess) e = hipMalloc(&cnt_buf, band_px * 4u);
  if (e == hipSuccess) e = hipMalloc(&img, (size_t)R * 16u);
  if (e != hipSuccess) {
    if (ids_buf) (void)hipFree(ids_buf);
    if (cnt_buf) (void)hipFree(cnt_buf);
    if (img) (void)hipFree(img);
    gt_set_error("gt_backward: scratch allocation failed: %s", hipGetErrorString(e));
    return GT_E_NOMEM;
  }
  f.rgba = (float*)img;
  f.comp_ids = (uint32_t*)ids_buf;
  f.n_comp = (uint32_t*)cnt_buf;
  f.comp_cap = cap;
  for (uint32_t ty0 = 0; ty0 < tiles_y && rc == GT_OK; ty0 += band) {
    const uint32_t ty1 = ty0 + band < tiles_y ? ty0 + band : tiles_y;
    const uint32_t y0 = ty0 * 8u, y1 = ty1 * 8u < height ? ty1 * 8u : height;
    f.comp_base = y0 * width;
    rc = gt_launch_forward(S, &f, nullptr, 1, ty1 * f.tiles_x, &tile8, stream, ty0 * f.tiles_x);
    if (rc != GT_OK) break;
    const uint32_t npix = (y1 - y0) * width;
    hipLaunchKernelGGL(k_backward, dim3((npix + 255) / 256), dim3(256), 0, st, *S, f, (const float4*)dl_drgba,
                       y0 * width, npix, grads, overflow);
    e = hipGetLastError();
    if (e != hipSuccess) { gt_set_error("backward launch failed: %s", hipGetErrorString(e)); rc = GT_E_HIP; }
  }
  (void)tiles;
  /* scratch is freed after the stream drains */
  const hipError_t es = hipStreamSynchronize(st);
  (void)hipFree(ids_buf);
  (void)hipFree(cnt_buf);
  (void)hipFree(img);
  if (rc == GT_OK && es != hipSuccess) {
    gt_set_error("gt_backward: %s", hipGetErrorString(es));
    rc = GT_E_HIP;
  }
  return rc;
}
