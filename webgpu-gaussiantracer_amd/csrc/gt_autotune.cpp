/*
 * gt_autotune.cpp - per-scene choice of the launch tuning by measurement.
 *
 * Candidates differ only in how work is scheduled (wave-wide packet
 * traversal in the persistent kernel or the tile kernel, per-lane traversal,
 * the lane-refill persistent kernel, ray-tile shape of a wave, LDS stack
 * depth,
 * persistent refill / composite thresholds); none changes a result.  Each
 * candidate renders the caller's workload (one frame or one batch) once to
 * warm up and then `iters` times between HIP events on the caller's stream;
 * the fastest is stored in the scene.
 */
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/gt.h"
#include "gt_internal.h"

namespace {

std::vector<gt_tuning> candidates() {
  std::vector<gt_tuning> c;
  for (uint32_t tw : {8u, 4u, 16u}) c.push_back(gt_tuning{4, 0, 0, 0, tw, 0}); /* first: the default */
  for (uint32_t tw : {8u, 4u}) c.push_back(gt_tuning{3, 0, 0, 0, tw, 0});
  for (uint32_t tw : {8u, 16u, 32u})
    for (uint32_t lds : {16u, 32u}) c.push_back(gt_tuning{1, 0, 0, 0, tw, lds});
  for (uint32_t tw : {8u, 16u})
    for (uint32_t th : {32u, 48u}) c.push_back(gt_tuning{2, th, th, 4, tw, 32});
  return c;
}

template <typename Render>
int tune(gt_scene* s, uint32_t iters, hipStream_t st, gt_autotune_result* out, Render render) {
  if (iters == 0) iters = 3;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    gt_set_error("hipEventCreate failed");
    return GT_E_HIP;
  }
  gt_tuning saved;
  gt_scene_get_tuning(s, &saved);
  double best = 1e300;
  gt_tuning best_t = saved;
  uint32_t tried = 0;
  int rc = GT_OK;
  for (gt_tuning t : candidates()) {
    t.rects = saved.rects; /* the caller's choices of rectangle culling and colours are kept */
    t.colors = saved.colors;
    if ((rc = gt_scene_set_tuning(s, &t)) != GT_OK) break;
    if ((rc = render()) != GT_OK) break; /* warm-up, graph capture */
    hipEventRecord(e0, st);
    for (uint32_t i = 0; i < iters && rc == GT_OK; ++i) rc = render();
    hipEventRecord(e1, st);
    if (rc != GT_OK) break;
    if (hipEventSynchronize(e1) != hipSuccess) { gt_set_error("hipEventSynchronize failed"); rc = GT_E_HIP; break; }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    const double per = ms / iters;
    if (tried == 0) out->first_ms = per;
    ++tried;
    if (per < best) {
      best = per;
      best_t = t;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != GT_OK) {
    gt_scene_set_tuning(s, &saved);
    return rc;
  }
  gt_scene_set_tuning(s, &best_t);
  out->best = best_t;
  out->best_ms = best;
  out->n_tried = tried;
  return GT_OK;
}

}  // namespace

extern "C" int gt_autotune_frame(gt_scene* s, const gt_uniforms* u, uint32_t width, uint32_t height, float* rgba_dev,
                                 uint32_t iters, void* stream, gt_autotune_result* out) {
  if (!s || !u || !rgba_dev || !out) { gt_set_error("gt_autotune_frame: NULL argument"); return GT_E_INVALID; }
  gt_frame f{};
  f.rgba = rgba_dev;
  hipStream_t st = (hipStream_t)stream;
  return tune(s, iters, st, out, [&]() { return gt_render(s, u, width, height, &f, GT_BUF_DEVICE, stream); });
}

gt_scene* gt_batch_scene(const gt_batch* b); /* gt_batch.cpp */

extern "C" int gt_autotune_batch(gt_scene* s, gt_batch* b, int use_graph, uint32_t iters, void* stream,
                                 gt_autotune_result* out) {
  if (!s || !b || !out) { gt_set_error("gt_autotune_batch: NULL argument"); return GT_E_INVALID; }
  if (gt_batch_scene(b) != s) {
    /* the candidates are set on s but the batch renders its own scene */
    gt_set_error("gt_autotune_batch: the batch renders another scene than the one being tuned");
    return GT_E_INVALID;
  }
  hipStream_t st = (hipStream_t)stream;
  return tune(s, iters, st, out, [&]() { return gt_batch_render(b, use_graph, stream); });
}
