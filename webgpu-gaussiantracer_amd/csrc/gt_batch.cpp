/*
 * gt_batch.cpp - many cameras of uneven resolution rendered by one launch.
 *
 * The reference renders one camera per traceRays dispatch
 * (webrtx/src/pass_encoder.ts:35-65, src/main.ts:563-593).  A batch keeps
 * the frames of n cameras (uniform block, size, caller-owned device outputs)
 * in one device table; gt_batch_render numbers the 64-ray tiles of all
 * cameras consecutively and renders them with one tile-kernel launch or one
 * persistent-kernel launch over a global ray counter (the scene's tuning
 * picks).  With use_graph the launch sequence is captured once into a
 * hipGraph and replayed on later calls.  Every camera's pixels, hit lists
 * and round counts are those of gt_render on that camera alone.
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <vector>

#include "../../include/gt.h"
#include "gt_internal.h"

/* gt_scene.cpp */
const gt_dev_scene* gt_scene_dev(const gt_scene* s);
const gt_tuning* gt_scene_tuning(const gt_scene* s);
int gt_validate_uniforms(const gt_uniforms* u, uint32_t width, uint32_t height);
void gt_fill_dev_frame(const gt_uniforms* u, uint32_t width, uint32_t height, gt_dev_frame* f);
uint32_t gt_frame_tiles(const gt_tuning* t, gt_dev_frame* f, uint32_t tile_base);

struct gt_batch {
  gt_scene* scene = nullptr;
  std::vector<gt_dev_frame> frames;
  void* table = nullptr;
  size_t table_bytes = 0;
  void* work = nullptr; /* the persistent renderers' work counters (GT_WORK_WORDS), this batch's own */
  uint32_t total_tiles = 0;
  gt_tuning tune_used{};
  bool table_valid = false;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  /* what the captured launch took by value: the scene's device arrays and
   * counts, and the persistent renderer's work counter.  A rebuild may move
   * or resize the arrays (gt_scene_set_alpha_min, a GPU rebuild or a refit
   * that falls back to one), after which the graph would read freed memory;
   * the graph is recaptured whenever these differ from the scene's. */
  gt_dev_scene dev_used{};

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~gt_batch() {
    drop_graph();
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    if (table) (void)hipFree(table);
    if (work) (void)hipFree(work);
  }
};

/* (re)numbers the tiles for the scene's current tile shape and uploads the
 * table; drops a captured graph whose launch arguments are out of date */
static int prepare_table(gt_batch* b) {
  const gt_tuning* t = gt_scene_tuning(b->scene);
  if (b->exec && std::memcmp(gt_scene_dev(b->scene), &b->dev_used, sizeof(gt_dev_scene)) != 0) b->drop_graph();
  if (b->table_valid && std::memcmp(t, &b->tune_used, sizeof(gt_tuning)) == 0) return GT_OK;
  uint64_t base = 0;
  for (auto& f : b->frames) {
    base += gt_frame_tiles(t, &f, (uint32_t)base);
    if (base * 64ull > 0xffffff00ull) {
      gt_set_error("batch has more than 2^32 rays");
      return GT_E_LIMIT;
    }
  }
  b->total_tiles = (uint32_t)base;
  const size_t bytes = b->frames.size() * sizeof(gt_dev_frame);
  if (!b->table || b->table_bytes < bytes) {
    if (b->table) (void)hipFree(b->table);
    b->table = nullptr;
    hipError_t e = hipMalloc(&b->table, bytes);
    if (e != hipSuccess) {
      gt_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
      return GT_E_NOMEM;
    }
    b->table_bytes = bytes;
  }
  hipError_t e = hipMemcpy(b->table, b->frames.data(), bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    gt_set_error("camera table upload failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  b->tune_used = *t;
  b->table_valid = true;
  b->drop_graph();
  return GT_OK;
}

static int launch(gt_batch* b, hipStream_t st) {
  const gt_tuning* t = gt_scene_tuning(b->scene);
  const gt_dev_frame* cams = (const gt_dev_frame*)b->table;
  const uint32_t n = (uint32_t)b->frames.size();
  unsigned int* work = (unsigned int*)b->work;
  if (t->path == 2)
    return gt_launch_persistent(gt_scene_dev(b->scene), &b->frames[0], cams, n, b->total_tiles, work,
                                (unsigned long long*)b->frames[0].stats, t, st);
  return gt_launch_forward(gt_scene_dev(b->scene), &b->frames[0], cams, n, b->total_tiles, t, st, 0, work, nullptr,
                           0);
}

extern "C" int gt_batch_create(gt_scene* s, uint32_t n, const gt_uniforms* u, const uint32_t* widths,
                               const uint32_t* heights, const gt_frame* outs, gt_batch** out) {
  if (!out) { gt_set_error("gt_batch_create: out is NULL"); return GT_E_INVALID; }
  *out = nullptr;
  if (!s || !u || !widths || !heights || !outs || n == 0) {
    gt_set_error("gt_batch_create: scene, uniforms, sizes and outputs are required (n >= 1)");
    return GT_E_INVALID;
  }
  /* the constant face table is uploaded here, never inside a graph capture */
  int prc = gt_forward_prepare();
  if (prc != GT_OK) return prc;
  gt_batch* b = new (std::nothrow) gt_batch();
  if (!b) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  b->scene = s;
  if (hipMalloc(&b->work, GT_WORK_WORDS * sizeof(unsigned int)) != hipSuccess) {
    b->work = nullptr;
    delete b;
    gt_set_error("hipMalloc of the batch's work counters failed");
    return GT_E_NOMEM;
  }
  b->frames.resize(n);
  const bool dump = outs[0].hit_ids != nullptr;
  for (uint32_t i = 0; i < n; ++i) {
    int rc = gt_validate_uniforms(&u[i], widths[i], heights[i]);
    if (rc == GT_OK && u[i].k != u[0].k) {
      gt_set_error("gt_batch_create: camera %u has k=%u, camera 0 has k=%u (a batch shares k)", i, u[i].k, u[0].k);
      rc = GT_E_INVALID;
    }
    if (rc == GT_OK && u[i].debug_mode != 0) {
      gt_set_error("gt_batch_create: camera %u: batches render debugMode 0 only", i);
      rc = GT_E_INVALID;
    }
    if (rc == GT_OK && outs[i].antialiasing) {
      gt_set_error("gt_batch_create: camera %u: antialiasing is a single-camera gt_render option", i);
      rc = GT_E_INVALID;
    }
    if (rc == GT_OK && (!outs[i].rgba || (outs[i].hit_ids != nullptr) != dump ||
                        (outs[i].accept != nullptr) != dump || (dump && outs[i].max_rounds == 0))) {
      gt_set_error("gt_batch_create: camera %u: rgba required; hit_ids/accept/max_rounds given for all or none", i);
      rc = GT_E_INVALID;
    }
    if (rc != GT_OK) {
      delete b;
      return rc;
    }
    gt_dev_frame& f = b->frames[i];
    gt_fill_dev_frame(&u[i], widths[i], heights[i], &f);
    f.rgba = outs[i].rgba;
    f.hit_ids = outs[i].hit_ids;
    f.accept = outs[i].accept;
    f.n_rounds = outs[i].n_rounds;
    f.T_out = outs[i].T;
    f.stats = (unsigned long long*)outs[0].stats;
    f.max_rounds = outs[i].max_rounds;
  }
  const int rc = prepare_table(b);
  if (rc != GT_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return GT_OK;
}

extern "C" int gt_batch_render(gt_batch* b, int use_graph, void* stream) {
  if (!b) { gt_set_error("gt_batch_render: batch is NULL"); return GT_E_INVALID; }
  int rc = prepare_table(b);
  if (rc != GT_OK) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return launch(b, st);
  if (!b->exec) {
    if (!b->cap_stream && hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess) {
      gt_set_error("capture stream creation failed");
      return GT_E_HIP;
    }
    hipError_t e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      gt_set_error("hipStreamBeginCapture failed: %s", hipGetErrorString(e));
      return GT_E_HIP;
    }
    rc = launch(b, b->cap_stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc != GT_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      gt_set_error("hipStreamEndCapture failed: %s", hipGetErrorString(e));
      return GT_E_HIP;
    }
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      b->drop_graph();
      gt_set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
      return GT_E_HIP;
    }
    b->dev_used = *gt_scene_dev(b->scene);
  }
  hipError_t e = hipGraphLaunch(b->exec, st);
  if (e != hipSuccess) {
    gt_set_error("hipGraphLaunch failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  return GT_OK;
}

gt_scene* gt_batch_scene(const gt_batch* b) { return b->scene; }

extern "C" int gt_batch_rays(const gt_batch* b, uint64_t* rays) {
  if (!b || !rays) { gt_set_error("gt_batch_rays: NULL argument"); return GT_E_INVALID; }
  uint64_t r = 0;
  for (const auto& f : b->frames) r += (uint64_t)f.width * f.height;
  *rays = r;
  return GT_OK;
}

extern "C" void gt_batch_destroy(gt_batch* b) { delete b; }
