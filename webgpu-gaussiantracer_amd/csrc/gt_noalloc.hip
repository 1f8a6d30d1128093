/*
 * gt_noalloc.hip - the batched render of gt_batch_create + gt_batch_render
 * as one call that allocates no device memory.
 *
 * For callers that cannot call hipMalloc: code under HIP graph capture,
 * tight serving loops, harnesses that forbid it.  The batch path's scratch
 * is the camera table the render kernels read (one gt_dev_frame per camera,
 * with its tile numbering) and the persistent renderers' work counters.
 * Here both live in `scratch`, a file-scope device array of fixed size,
 * which every GPU the library runs on holds until the process ends.  Host
 * code never passes `scratch` itself to a kernel or a copy (in host code it
 * is only a stand-in for the device symbol): each device's address comes
 * from hipGetSymbolAddress, taken once per device with that device current.
 *
 * The camera table is written by a kernel that takes up to 16 cameras by
 * value as its argument, so no host buffer must outlive the call and
 * nothing is recorded as a memset or a host-to-device copy.
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/gt.h"
#include "gt_internal.h"

/* gt_scene.cpp */
const gt_dev_scene* gt_scene_dev(const gt_scene* s);
const gt_tuning* gt_scene_tuning(const gt_scene* s);
int gt_validate_uniforms(const gt_uniforms* u, uint32_t width, uint32_t height);
void gt_fill_dev_frame(const gt_uniforms* u, uint32_t width, uint32_t height, gt_dev_frame* f);
uint32_t gt_frame_tiles(const gt_tuning* t, gt_dev_frame* f, uint32_t tile_base);

namespace {

constexpr size_t kCounterBytes = 512; /* the work counters (GT_WORK_WORDS), padded to keep the table 256-B aligned */
constexpr size_t kScratchBytes = kCounterBytes + (size_t)GT_NOALLOC_MAX_CAMERAS * sizeof(gt_dev_frame);
constexpr uint32_t kChunk = 16; /* cameras per table-writing launch (16 x 208 B of kernel arguments) */

__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct FrameChunk {
  gt_dev_frame f[kChunk];
};
static_assert(sizeof(FrameChunk) <= 4000, "kernel argument space");

/* table[first + i] = c.f[i] for i < count */
__global__ void k_write_table(gt_dev_frame* __restrict__ table, uint32_t first, uint32_t count, const FrameChunk c) {
  const uint32_t words = (uint32_t)(sizeof(gt_dev_frame) / 4u);
  const uint32_t total = count * words;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&c.f[0]);
  uint32_t* dst = reinterpret_cast<uint32_t*>(table + first);
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) dst[i] = src[i];
}

std::mutex g_mu;
void* g_ptr[64];  /* scratch's device address on device d (set once) */

/* scratch's address on the current device, looked up once per device */
int scratch_ptr(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    gt_set_error("hipGetDevice failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  if (dev < 0 || dev >= 64) {
    gt_set_error("device %d out of the supported range 0..63", dev);
    return GT_E_LIMIT;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      gt_set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return GT_E_HIP;
    }
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return GT_OK;
}

}  // namespace

extern "C" int gt_batch_render_noalloc(gt_scene* s, uint32_t n, const gt_uniforms* u, const uint32_t* widths,
                                       const uint32_t* heights, const gt_frame* outs, void* stream) {
  if (!s || !u || !widths || !heights || !outs || n == 0) {
    gt_set_error("gt_batch_render_noalloc: scene, uniforms, sizes and outputs are required (n >= 1)");
    return GT_E_INVALID;
  }
  if (n > GT_NOALLOC_MAX_CAMERAS) {
    gt_set_error("gt_batch_render_noalloc: %u cameras exceed the limit of %u (GT_NOALLOC_MAX_CAMERAS)", n,
                 (unsigned)GT_NOALLOC_MAX_CAMERAS);
    return GT_E_LIMIT;
  }
  /* the same checks as gt_batch_create */
  const bool dump = outs[0].hit_ids != nullptr;
  for (uint32_t i = 0; i < n; ++i) {
    int rc = gt_validate_uniforms(&u[i], widths[i], heights[i]);
    if (rc != GT_OK) return rc;
    if (u[i].k != u[0].k) {
      gt_set_error("gt_batch_render_noalloc: camera %u has k=%u, camera 0 has k=%u (a batch shares k)", i, u[i].k,
                   u[0].k);
      return GT_E_INVALID;
    }
    if (u[i].debug_mode != 0) {
      gt_set_error("gt_batch_render_noalloc: camera %u: batches render debugMode 0 only", i);
      return GT_E_INVALID;
    }
    if (outs[i].antialiasing) {
      gt_set_error("gt_batch_render_noalloc: camera %u: antialiasing is a single-camera gt_render option", i);
      return GT_E_INVALID;
    }
    if (!outs[i].rgba || (outs[i].hit_ids != nullptr) != dump || (outs[i].accept != nullptr) != dump ||
        (dump && outs[i].max_rounds == 0)) {
      gt_set_error("gt_batch_render_noalloc: camera %u: rgba required; hit_ids/accept/max_rounds given for all or "
                   "none", i);
      return GT_E_INVALID;
    }
  }
  void* base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc != GT_OK) return rc;
  unsigned int* work = reinterpret_cast<unsigned int*>(base);
  gt_dev_frame* table = reinterpret_cast<gt_dev_frame*>(static_cast<unsigned char*>(base) + kCounterBytes);
  hipStream_t st = (hipStream_t)stream;
  const gt_tuning* t = gt_scene_tuning(s);
  /* number the tiles and write the table, kChunk cameras per launch */
  uint64_t tiles = 0;
  gt_dev_frame first{};
  FrameChunk c;
  for (uint32_t i0 = 0; i0 < n; i0 += kChunk) {
    const uint32_t cnt = n - i0 < kChunk ? n - i0 : kChunk;
    std::memset(&c, 0, sizeof(c));
    for (uint32_t j = 0; j < cnt; ++j) {
      const uint32_t i = i0 + j;
      gt_dev_frame& f = c.f[j];
      gt_fill_dev_frame(&u[i], widths[i], heights[i], &f);
      f.rgba = outs[i].rgba;
      f.hit_ids = outs[i].hit_ids;
      f.accept = outs[i].accept;
      f.n_rounds = outs[i].n_rounds;
      f.T_out = outs[i].T;
      f.stats = (unsigned long long*)outs[0].stats;
      f.max_rounds = outs[i].max_rounds;
      tiles += gt_frame_tiles(t, &f, (uint32_t)tiles);
      if (tiles * 64ull > 0xffffff00ull) {
        gt_set_error("gt_batch_render_noalloc: the batch has more than 2^32 rays");
        return GT_E_LIMIT;
      }
      if (i == 0) first = f;
    }
    hipLaunchKernelGGL(k_write_table, dim3(1), dim3(256), 0, st, table, i0, cnt, c);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      gt_set_error("camera table launch failed: %s", hipGetErrorString(e));
      return GT_E_HIP;
    }
  }
  const uint32_t total = (uint32_t)tiles;
  if (t->path == 2)
    return gt_launch_persistent(gt_scene_dev(s), &first, table, n, total, work,
                                (unsigned long long*)outs[0].stats, t, stream);
  return gt_launch_forward(gt_scene_dev(s), &first, table, n, total, t, stream, 0, work, nullptr, 0);
}
