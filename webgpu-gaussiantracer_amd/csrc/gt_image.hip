/*
 * gt_image.hip - presentation of a rendered frame: the reference's screen
 * pass (src/toneMapper.ts:11-26, fragment shader main) and its Save button
 * (src/main.ts:606-611, canvas.toDataURL('image/png')).
 *
 * The fragment shader reads pixelBuffer[x + y * W] for canvas pixel (x, y)
 * and returns hdrColor with alpha 1 (toneMapper.ts:16-22); the 8-bit canvas
 * stores each channel clamped to [0, 1] and scaled to 255 steps.  The
 * shader also computes an unused exposure curve ldr = 1 - exp(-1.66 hdr)
 * (toneMapper.ts:19) and has a commented-out gamma pow(ldr, 1/2.2)
 * (toneMapper.ts:21); both are offered as modes.
 *
 * gt_tonemap: one lane per pixel, 16 B read and 4 B written (HBM-bound), or
 * the same function on the host.  The curves run in FP64 with this
 * library's exp/log (gt_exp64, gt_log64), so host and device agree byte for
 * byte.
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <thread>
#include <vector>

#include "../../include/gt.h"
#include "../../include/gt_fp32.h"
#include "gt_internal.h"

namespace {

#define TONE_BLOCK 256

/* 255-step unorm conversion: clamp to [0, 1], round half up; NaN -> 0 */
__host__ __device__ inline uint8_t unorm8(double v) {
  if (!(v > 0.0)) return 0u;
  if (v >= 1.0) return 255u;
  return (uint8_t)(v * 255.0 + 0.5);
}

__host__ __device__ inline double tone_curve(float h, uint32_t mode) {
  const double x = (double)h;
  if (mode == GT_TONEMAP_IDENTITY) return x;
  const double ldr = 1.0 - gt_exp64(-(double)1.66f * x);
  if (mode == GT_TONEMAP_EXPOSURE) return ldr;
  if (!(ldr > 0.0)) return 0.0;
  return gt_exp64(gt_log64(ldr) / 2.2);
}

__host__ __device__ inline uchar4 tone_pixel(float4 v, uint32_t mode) {
  uchar4 o;
  o.x = unorm8(tone_curve(v.x, mode));
  o.y = unorm8(tone_curve(v.y, mode));
  o.z = unorm8(tone_curve(v.z, mode));
  o.w = 255u;
  return o;
}

__global__ __launch_bounds__(TONE_BLOCK) void k_tonemap(const float4* __restrict__ in, uint64_t n, uint32_t mode,
                                                        uchar4* __restrict__ out) {
  const uint64_t stride = (uint64_t)gridDim.x * TONE_BLOCK;
  for (uint64_t i = (uint64_t)blockIdx.x * TONE_BLOCK + threadIdx.x; i < n; i += stride) out[i] = tone_pixel(in[i], mode);
}

}  // namespace

extern "C" int gt_tonemap(const float* rgba, uint32_t width, uint32_t height, uint32_t mode, uint8_t* out,
                          int buffers, void* hip_stream) {
  if (mode > GT_TONEMAP_EXPOSURE_GAMMA) { gt_set_error("gt_tonemap: unknown mode %u", mode); return GT_E_INVALID; }
  if (buffers != GT_BUF_HOST && buffers != GT_BUF_DEVICE) {
    gt_set_error("gt_tonemap: buffers must be GT_BUF_HOST or GT_BUF_DEVICE");
    return GT_E_INVALID;
  }
  const uint64_t n = (uint64_t)width * height;
  if (n == 0) return GT_OK;
  if (!rgba || !out) { gt_set_error("gt_tonemap: NULL buffer"); return GT_E_INVALID; }
  const float4* in = reinterpret_cast<const float4*>(rgba);
  uchar4* o = reinterpret_cast<uchar4*>(out);
  if (buffers == GT_BUF_DEVICE) {
    const uint64_t want = (n + TONE_BLOCK - 1) / TONE_BLOCK;
    const uint32_t blocks = (uint32_t)(want < 16384 ? want : 16384);
    k_tonemap<<<blocks, TONE_BLOCK, 0, (hipStream_t)hip_stream>>>(in, n, mode, o);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      gt_set_error("gt_tonemap launch failed: %s", hipGetErrorString(e));
      return GT_E_HIP;
    }
    return GT_OK;
  }
  const unsigned hw = std::thread::hardware_concurrency();
  const uint64_t nt = std::min<uint64_t>(std::max(1u, std::min(hw, 16u)), (n + 65535) / 65536);
  std::vector<std::thread> th;
  for (uint64_t t = 0; t < nt; ++t)
    th.emplace_back([=]() {
      for (uint64_t i = n * t / nt; i < n * (t + 1) / nt; ++i) o[i] = tone_pixel(in[i], mode);
    });
  for (auto& x : th) x.join();
  return GT_OK;
}
