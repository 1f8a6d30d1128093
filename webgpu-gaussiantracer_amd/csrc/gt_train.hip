/*
 * gt_train.hip - the optimizer step of a training loop over the PLY-raw
 * parameters (the reference plans training, README.md:68-69, and has no
 * optimizer of its own).
 *
 * gt_adam_step: Adam (Kingma & Ba 2015) in place on an n x 59 parameter
 * array, one lane per parameter, in the operation order of PyTorch's
 * torch.optim.Adam (single-tensor path):
 *   m = m + (1 - beta1) (g - m)                 (lerp)
 *   v = beta2 v + (1 - beta2) g g
 *   p = p - (lr_c / bc1) m / (sqrt(v) / sqrt(bc2) + eps)
 * with bc1 = 1 - beta1^step, bc2 = 1 - beta2^step and lr_c the learning
 * rate of the parameter's column (field).  Streaming and HBM-bound: 16 B
 * read and 12 B written per parameter.
 */
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/gt.h"
#include "gt_internal.h"

namespace {

#define ADAM_BLOCK 256

struct AdamArgs {
  float step_size[GT_RAW_FLOATS]; /* lr_c / bc1 */
  float beta1, beta2, eps, sqrt_bc2;
};

__global__ __launch_bounds__(ADAM_BLOCK) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                                     float* __restrict__ m, float* __restrict__ v, uint64_t total,
                                                     AdamArgs a) {
  __shared__ float s_step[GT_RAW_FLOATS];
  for (uint32_t i = threadIdx.x; i < GT_RAW_FLOATS; i += ADAM_BLOCK) s_step[i] = a.step_size[i];
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * ADAM_BLOCK;
  const uint32_t col_step = (uint32_t)(stride % GT_RAW_FLOATS);
  uint64_t i = (uint64_t)blockIdx.x * ADAM_BLOCK + threadIdx.x;
  uint32_t col = (uint32_t)(i % GT_RAW_FLOATS); /* field of parameter i, advanced with i */
  for (; i < total; i += stride) {
    const float gi = g[i];
    float mi = m[i];
    mi = mi + (1.0f - a.beta1) * (gi - mi);
    const float vi = a.beta2 * v[i] + (1.0f - a.beta2) * (gi * gi);
    const float denom = sqrtf(vi) / a.sqrt_bc2 + a.eps;
    p[i] = p[i] - s_step[col] * (mi / denom);
    m[i] = mi;
    v[i] = vi;
    col += col_step;
    if (col >= GT_RAW_FLOATS) col -= GT_RAW_FLOATS;
  }
}

}  // namespace

extern "C" int gt_adam_step(float* params, const float* grads, float* m, float* v, uint32_t n, const gt_adam* adam,
                            void* hip_stream) {
  if (!adam) { gt_set_error("gt_adam_step: adam is NULL"); return GT_E_INVALID; }
  if (n && (!params || !grads || !m || !v)) { gt_set_error("gt_adam_step: NULL buffer"); return GT_E_INVALID; }
  if (!(adam->beta1 >= 0.0f && adam->beta1 < 1.0f) || !(adam->beta2 >= 0.0f && adam->beta2 < 1.0f) ||
      !(adam->eps > 0.0f) || !std::isfinite(adam->eps) || adam->step < 1) {
    gt_set_error("gt_adam_step: need 0 <= beta1, beta2 < 1, eps > 0 and step >= 1 (beta1 %g, beta2 %g, eps %g, "
                 "step %u)", adam->beta1, adam->beta2, adam->eps, adam->step);
    return GT_E_INVALID;
  }
  AdamArgs a;
  const double bc1 = 1.0 - std::pow((double)adam->beta1, (double)adam->step);
  const double bc2 = 1.0 - std::pow((double)adam->beta2, (double)adam->step);
  for (int c = 0; c < GT_RAW_FLOATS; ++c) {
    if (!std::isfinite(adam->lr[c]) || adam->lr[c] < 0.0f) {
      gt_set_error("gt_adam_step: learning rate of field %d is %g", c, adam->lr[c]);
      return GT_E_INVALID;
    }
    a.step_size[c] = (float)((double)adam->lr[c] / bc1);
  }
  a.beta1 = adam->beta1;
  a.beta2 = adam->beta2;
  a.eps = adam->eps;
  a.sqrt_bc2 = (float)std::sqrt(bc2);
  if (n == 0) return GT_OK;
  const uint64_t total = (uint64_t)n * GT_RAW_FLOATS;
  const uint64_t want = (total + ADAM_BLOCK - 1) / ADAM_BLOCK;
  const uint32_t blocks = (uint32_t)(want < 8192 ? want : 8192);
  k_adam<<<blocks, ADAM_BLOCK, 0, (hipStream_t)hip_stream>>>(params, grads, m, v, total, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    gt_set_error("gt_adam_step launch failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  return GT_OK;
}
