/*
 * gt_scene.cpp - scene (acceleration container) upload and the gt_render
 * host wrapper.
 *
 * Replaces, for the Gaussian path:
 *   TLAS/BLAS upload                        webrtx/src/wasm_bvh_builder.ts:260-377
 * The reference builds a binary SAH TLAS over instance boxes and walks a
 * 39-node BLAS per instance; here a binned-SAH binary tree over padded
 * world boxes of the proxies (gt_host_build.cpp) is collapsed into 4-wide
 * 128-B nodes whose leaves hold runs of 64-B proxy records (world-to-object
 * matrix + id), and the 20 icosahedron faces are tested directly (no BLAS).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include "../../include/gt.h"
#include "../../include/gt_fp32.h"
#include "gt_host_build.h"
#include "gt_internal.h"
#include "gt_prep.h"
#include "gt_scene_impl.h"

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      gt_set_error("%s failed: %s", #expr, hipGetErrorString(e_));                 \
      return GT_E_HIP;                                                             \
    }                                                                              \
  } while (0)

static void face_table(float* f) {
  for (int i = 0; i < GT_NUM_FACES; ++i) {
    const gt_face fc = gt_ico_face(i);
    const float v[12] = {fc.p0.x, fc.p0.y, fc.p0.z, fc.e0.x, fc.e0.y, fc.e0.z,
                         fc.e1.x, fc.e1.y, fc.e1.z, fc.n.x,  fc.n.y,  fc.n.z};
    std::memcpy(f + 12 * i, v, sizeof(v));
  }
}

template <typename T>
static int upload(DevBuf& b, const T* src, size_t count) {
  if (b.p) { (void)hipFree(b.p); b.p = nullptr; }
  b.bytes = std::max<size_t>(count * sizeof(T), 256);
  hipError_t e = hipMalloc(&b.p, b.bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    gt_set_error("hipMalloc(%zu) failed: %s", b.bytes, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? GT_E_NOMEM : GT_E_HIP;
  }
  if (count) HIP_TRY(hipMemcpy(b.p, src, count * sizeof(T), hipMemcpyHostToDevice));
  return GT_OK;
}


int gt_devbuf_reserve(DevBuf& b, size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  if (b.p && b.bytes >= bytes) return GT_OK;
  if (b.p) { (void)hipFree(b.p); b.p = nullptr; b.bytes = 0; }
  const hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    gt_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? GT_E_NOMEM : GT_E_HIP;
  }
  b.bytes = bytes;
  return GT_OK;
}

int gt_scene_static_upload(gt_scene* s) {
  int rc;
  /* the render kernels' constant face table, before any render (which may
   * be captured into a graph) needs it */
  if ((rc = gt_forward_prepare()) != GT_OK) return rc;
  if (!s->faces.p) {
    float faces[12 * GT_NUM_FACES];
    face_table(faces);
    if ((rc = upload(s->faces, faces, 12 * GT_NUM_FACES)) != GT_OK) return rc;
  }
  if (!s->ctl.p) { /* the persistent renderers' work counters, one set per stream */
    const std::vector<uint32_t> zero((size_t)GT_WORK_SLOTS * GT_WORK_WORDS, 0u);
    if ((rc = upload(s->ctl, zero.data(), zero.size())) != GT_OK) return rc;
  }
  return GT_OK;
}

static int build_and_upload(gt_scene* s) {
  const auto t0 = std::chrono::steady_clock::now();
  HostScene& h = s->h;
  int rc = gt_host_build_proxies(h, s->alpha_min);
  if (rc != GT_OK) return rc;
  HostTree tree;
  if ((rc = gt_host_build_tree(h, tree)) != GT_OK) return rc;
  const uint32_t np = (uint32_t)h.proxy_ids.size();
  std::vector<gt_node4>& nodes = tree.nodes;
  std::vector<float>& prox = tree.prox;
  const uint32_t root = tree.root, need = tree.need, depth = tree.depth;
  std::vector<float> geom, sh;
  gt_host_shading_records(h, geom, sh);
  if ((rc = upload(s->geom, geom.data(), geom.size())) != GT_OK) return rc;
  if ((rc = upload(s->sh, sh.data(), sh.size())) != GT_OK) return rc;
  if ((rc = upload(s->prox, prox.data(), prox.size())) != GT_OK) return rc;
  if ((rc = upload(s->nodes, nodes.data(), nodes.size())) != GT_OK) return rc;
  if ((rc = gt_scene_static_upload(s)) != GT_OK) return rc;
  s->dev.geom = (const float*)s->geom.p;
  s->dev.sh = (const float*)s->sh.p;
  s->dev.prox = (const float*)s->prox.p;
  s->dev.nodes = (const gt_node4*)s->nodes.p;
  s->dev.faces = (const float*)s->faces.p;
  s->dev.root = root;
  s->n = h.n;
  s->dev.n = h.n;
  s->dev.n_prox = np;
  s->dev.n_nodes = (uint32_t)nodes.size();
  s->n_nodes = (uint32_t)nodes.size();
  s->max_stack = need;
  s->depth = depth;
  s->rsched = false; /* depth-first numbering: the refit schedule is made on first use */
  s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return GT_OK;
}

static int finish_create(std::unique_ptr<gt_scene>& s, gt_scene** out) {
  if (hipGetDevice(&s->device) != hipSuccess) {
    gt_set_error("no HIP device available");
    return GT_E_HIP;
  }
  const int rc = build_and_upload(s.get());
  if (rc != GT_OK) return rc;
  *out = s.release();
  return GT_OK;
}

extern "C" int gt_scene_create_raw(const float* raw, uint32_t n, double alpha_min, gt_scene** out) {
  if (!out) { gt_set_error("gt_scene_create_raw: out is NULL"); return GT_E_INVALID; }
  *out = nullptr;
  if (n && !raw) { gt_set_error("gt_scene_create_raw: raw is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_create_raw: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  std::unique_ptr<gt_scene> s(new (std::nothrow) gt_scene());
  if (!s) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  try {
    s->h.n = n;
    s->n = n;
    s->h.raw.assign(raw, raw + (size_t)n * GT_RAW_FLOATS);
    s->alpha_min = alpha_min;
    const int rc = gt_host_activate_raw(s->h);
    if (rc != GT_OK) return rc;
    return finish_create(s, out);
  } catch (const std::bad_alloc&) {
    gt_set_error("out of host memory");
    return GT_E_NOMEM;
  }
}

extern "C" int gt_scene_create_packed(const void* records, uint32_t n, double alpha_min, gt_scene** out) {
  if (!out) { gt_set_error("gt_scene_create_packed: out is NULL"); return GT_E_INVALID; }
  *out = nullptr;
  if (n && !records) { gt_set_error("gt_scene_create_packed: records is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_create_packed: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  std::unique_ptr<gt_scene> s(new (std::nothrow) gt_scene());
  if (!s) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  try {
    s->h.n = n;
    s->n = n;
    const float* f = (const float*)records;
    s->h.packed.assign(f, f + (size_t)n * 80);
    s->alpha_min = alpha_min;
    const int rc = gt_host_activate_packed(s->h);
    if (rc != GT_OK) return rc;
    return finish_create(s, out);
  } catch (const std::bad_alloc&) {
    gt_set_error("out of host memory");
    return GT_E_NOMEM;
  }
}

extern "C" int gt_scene_set_alpha_min(gt_scene* s, double alpha_min) {
  if (!s) { gt_set_error("gt_scene_set_alpha_min: scene is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_set_alpha_min: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  s->alpha_min = alpha_min;
  if (s->device_built) return gt_device_build(s, nullptr);
  try {
    return build_and_upload(s);
  } catch (const std::bad_alloc&) {
    gt_set_error("out of host memory");
    return GT_E_NOMEM;
  }
}

extern "C" void gt_scene_destroy(gt_scene* s) { delete s; }

/* ---------------- device build (GPU LBVH) ---------------- */
static int device_rebuild(gt_scene* s, const float* raw_dev, void* stream) {
  const size_t bytes = (size_t)s->n * GT_RAW_FLOATS * sizeof(float);
  int rc = gt_devbuf_reserve(s->raw_dev, bytes);
  if (rc != GT_OK) return rc;
  if (bytes) HIP_TRY(hipMemcpyAsync(s->raw_dev.p, raw_dev, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return gt_device_build(s, (hipStream_t)stream);
}

extern "C" int gt_scene_create_device(const float* raw_dev, uint32_t n, double alpha_min, void* stream,
                                      gt_scene** out) {
  if (!out) { gt_set_error("gt_scene_create_device: out is NULL"); return GT_E_INVALID; }
  *out = nullptr;
  if (n && !raw_dev) { gt_set_error("gt_scene_create_device: raw_dev is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_create_device: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  if (n > GT_MAX_PROXIES) {
    gt_set_error("gt_scene_create_device: %u Gaussians exceed the limit %u", n, GT_MAX_PROXIES);
    return GT_E_LIMIT;
  }
  std::unique_ptr<gt_scene> s(new (std::nothrow) gt_scene());
  if (!s) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  if (hipGetDevice(&s->device) != hipSuccess) {
    gt_set_error("no HIP device available");
    return GT_E_HIP;
  }
  s->n = n;
  s->device_built = true;
  s->alpha_min = alpha_min;
  const int rc = device_rebuild(s.get(), raw_dev, stream);
  if (rc != GT_OK) return rc;
  *out = s.release();
  return GT_OK;
}

extern "C" int gt_scene_create_raw_gpu(const float* raw, uint32_t n, double alpha_min, gt_scene** out) {
  if (!out) { gt_set_error("gt_scene_create_raw_gpu: out is NULL"); return GT_E_INVALID; }
  *out = nullptr;
  if (n && !raw) { gt_set_error("gt_scene_create_raw_gpu: raw is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_create_raw_gpu: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  if (n > GT_MAX_PROXIES) {
    gt_set_error("gt_scene_create_raw_gpu: %u Gaussians exceed the limit %u", n, GT_MAX_PROXIES);
    return GT_E_LIMIT;
  }
  std::unique_ptr<gt_scene> s(new (std::nothrow) gt_scene());
  if (!s) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  if (hipGetDevice(&s->device) != hipSuccess) {
    gt_set_error("no HIP device available");
    return GT_E_HIP;
  }
  s->n = n;
  s->device_built = true;
  s->alpha_min = alpha_min;
  const size_t bytes = (size_t)n * GT_RAW_FLOATS * sizeof(float);
  int rc = gt_devbuf_reserve(s->raw_dev, bytes);
  if (rc != GT_OK) return rc;
  if (bytes) HIP_TRY(hipMemcpy(s->raw_dev.p, raw, bytes, hipMemcpyHostToDevice));
  rc = gt_device_build(s.get(), nullptr);
  if (rc != GT_OK) return rc;
  *out = s.release();
  return GT_OK;
}

extern "C" int gt_scene_update_device(gt_scene* s, const float* raw_dev, void* stream) {
  if (!s) { gt_set_error("gt_scene_update_device: scene is NULL"); return GT_E_INVALID; }
  if (!s->device_built) {
    gt_set_error("gt_scene_update_device: the scene was not created by gt_scene_create_device");
    return GT_E_INVALID;
  }
  if (s->n && !raw_dev) { gt_set_error("gt_scene_update_device: raw_dev is NULL"); return GT_E_INVALID; }
  return device_rebuild(s, raw_dev, stream);
}

extern "C" int gt_scene_refit_device(gt_scene* s, const float* raw_dev, void* stream, uint32_t* rebuilt) {
  if (!s) { gt_set_error("gt_scene_refit_device: scene is NULL"); return GT_E_INVALID; }
  if (s->n && !raw_dev) { gt_set_error("gt_scene_refit_device: raw_dev is NULL"); return GT_E_INVALID; }
  uint32_t r = 0;
  const int rc = gt_device_refit(s, raw_dev, (hipStream_t)stream, &r);
  if (rebuilt) *rebuilt = r;
  return rc;
}

extern "C" int gt_scene_get_info(const gt_scene* s, gt_scene_info* info) {
  if (!s || !info) { gt_set_error("gt_scene_get_info: NULL argument"); return GT_E_INVALID; }
  std::memset(info, 0, sizeof(*info));
  info->n_gaussians = s->n;
  info->n_proxies = s->dev.n_prox;
  info->n_nodes = s->n_nodes;
  info->max_stack = s->max_stack;
  info->bvh_depth = s->depth;
  info->device_bytes = s->geom.bytes + s->sh.bytes + s->prox.bytes + s->nodes.bytes + s->faces.bytes +
                      s->raw_dev.bytes;
  info->build_ms = s->build_ms;
  return GT_OK;
}

extern "C" int gt_scene_get_host_data(const gt_scene* s, float* packed, uint32_t* proxy_ids, float* o2w,
                                      float* w2o) {
  if (!s) { gt_set_error("gt_scene_get_host_data: scene is NULL"); return GT_E_INVALID; }
  if (s->device_built) return gt_device_host_data(s, packed, proxy_ids, o2w, w2o);
  const HostScene& h = s->h;
  if (packed) std::memcpy(packed, h.packed.data(), h.packed.size() * sizeof(float));
  if (proxy_ids) std::memcpy(proxy_ids, h.proxy_ids.data(), h.proxy_ids.size() * sizeof(uint32_t));
  if (o2w) std::memcpy(o2w, h.o2w.data(), h.o2w.size() * sizeof(float));
  if (w2o) std::memcpy(w2o, h.w2o.data(), h.w2o.size() * sizeof(float));
  return GT_OK;
}

/* ---------------- render ---------------- */
int gt_validate_uniforms(const gt_uniforms* u, uint32_t width, uint32_t height) {
  if (!u) { gt_set_error("uniforms are NULL"); return GT_E_INVALID; }
  if (width < 1 || height < 1 || width > 16384 || height > 16384) {
    gt_set_error("width(%u), height(%u) must be in 1..16384", width, height);
    return GT_E_INVALID;
  }
  if (u->k < 1 || u->k > GT_MAX_K) {
    gt_set_error("hit_array_size k=%u must be in 1..%d", u->k, GT_MAX_K);
    return GT_E_LIMIT;
  }
  if (!(u->t_min >= 0.0f) || !std::isfinite(u->t_min) || !(u->t_max > u->t_min)) {
    gt_set_error("need 0 <= t_min < t_max (t_min=%g, t_max=%g)", u->t_min, u->t_max);
    return GT_E_INVALID;
  }
  if (u->sh_degree_max > 3) { gt_set_error("sh_degree_max %u > 3", u->sh_degree_max); return GT_E_INVALID; }
  if (u->debug_mode > 2) { gt_set_error("debugMode %u not in {0,1,2}", u->debug_mode); return GT_E_INVALID; }
  for (int i = 0; i < 16; ++i)
    if (!std::isfinite(u->to_world[i])) { gt_set_error("to_world has a non-finite entry"); return GT_E_INVALID; }
  if (!std::isfinite(u->camera_aspect[0]) || !std::isfinite(u->camera_aspect[1]) || !std::isfinite(u->T_min) ||
      !std::isfinite(u->alpha_min)) {
    gt_set_error("non-finite uniform value");
    return GT_E_INVALID;
  }
  return GT_OK;
}

void gt_fill_dev_frame(const gt_uniforms* u, uint32_t width, uint32_t height, gt_dev_frame* f) {
  std::memset(f, 0, sizeof(*f));
  std::memcpy(f->to_world, u->to_world, sizeof(f->to_world));
  f->ax = u->camera_aspect[0];
  f->ay = u->camera_aspect[1];
  f->t_min = u->t_min;
  f->t_max = u->t_max;
  f->T_min = u->T_min;
  f->alpha_min = u->alpha_min;
  f->k = u->k;
  f->sh_degree = u->sh_degree_max;
  f->early_stop = u->early_stop;
  f->debug_mode = u->debug_mode;
  f->width = width;
  f->height = height;
  const float om = std::max(std::fabs(u->to_world[12]), std::max(std::fabs(u->to_world[13]), std::fabs(u->to_world[14])));
  f->slack_abs = 2e-5f * (om + 1.0f);
  f->accumulated_frames = u->accumulated_frames;
  f->tiles_x = (width + 7u) / 8u;
  f->tile_base = 0;
}

/* The per-tile kernel by default (faster on a single camera: the rays of a
 * wave stay coherent), the persistent kernel with tuning path 2. */
uint32_t gt_tune_tile_w(const gt_tuning* t) { return t->tile_w ? t->tile_w : 8u; }

/* tiles_x / tile_base of a frame for the tuning's tile shape; returns its tile count */
uint32_t gt_frame_tiles(const gt_tuning* t, gt_dev_frame* f, uint32_t tile_base) {
  const uint32_t tw = gt_tune_tile_w(t), th = 64u / tw;
  f->tiles_x = (f->width + tw - 1u) / tw;
  f->tile_base = tile_base;
  return f->tiles_x * ((f->height + th - 1u) / th);
}

/* The slot of `stream` in this scene (assigned on the stream's first
 * render), or -1 once GT_WORK_SLOTS streams hold one.  Caller holds ctl_mu. */
static int stream_slot(gt_scene* s, void* stream) {
  for (uint32_t i = 0; i < s->ctl_used; ++i)
    if (s->ctl_streams[i] == stream) return (int)i;
  if (s->ctl_used == GT_WORK_SLOTS) return -1;
  s->ctl_streams[s->ctl_used] = stream;
  return (int)s->ctl_used++;
}

/* The work counters of `stream` in this scene, or NULL once GT_WORK_SLOTS
 * streams hold a slot. */
unsigned int* gt_scene_work(gt_scene* s, void* stream) {
  std::lock_guard<std::mutex> lk(s->ctl_mu);
  unsigned int* base = (unsigned int*)s->ctl.p;
  if (!base) return nullptr;
  const int i = stream_slot(s, stream);
  return i < 0 ? nullptr : base + (size_t)i * GT_WORK_WORDS;
}

/* The proxy-rectangle table of `stream`: GT_RECT_ENTRIES(n_prox) entries
 * and extra_words more (the tile-order workspace), its capacity in 32-bit
 * words in *words; or NULL: no slot, an allocation that fails, or one that
 * would be needed while the stream captures a graph.  The rectangles only
 * cull work and order it, so a frame without them renders the same
 * result. */
uint32_t* gt_scene_rects(gt_scene* s, void* stream, size_t extra_words, size_t* words) {
  *words = 0;
  std::lock_guard<std::mutex> lk(s->ctl_mu);
  const int i = stream_slot(s, stream);
  if (i < 0) return nullptr;
  DevBuf& b = s->rects[i];
  const size_t bytes = (GT_RECT_ENTRIES(s->dev.n_prox) * 2 + extra_words) * sizeof(uint32_t);
  if (!b.p || b.bytes < bytes) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing((hipStream_t)stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
      return nullptr;
    if (gt_devbuf_reserve(b, bytes) != GT_OK) {
      (void)hipGetLastError(); /* the failed allocation is not the launch's error */
      return nullptr;
    }
  }
  *words = b.bytes / sizeof(uint32_t);
  return (uint32_t*)b.p;
}

/* This stream's buffer for the frame's per-Gaussian colours (n float4), or
 * NULL: on a stream rendered from no slot, while the stream is being
 * captured and the buffer would have to grow, or when it cannot be
 * allocated.  The frame then evaluates SH per hit (the same values). */
static float* scene_colors(gt_scene* s, void* stream) {
  std::lock_guard<std::mutex> lk(s->ctl_mu);
  const int i = stream_slot(s, stream);
  if (i < 0) return nullptr;
  DevBuf& b = s->colors[i];
  const size_t bytes = (size_t)s->dev.n * 4u * sizeof(float);
  if (!b.p || b.bytes < bytes) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing((hipStream_t)stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
      return nullptr;
    if (gt_devbuf_reserve(b, bytes) != GT_OK) {
      (void)hipGetLastError(); /* the failed allocation is not the launch's error */
      return nullptr;
    }
  }
  return (float*)b.p;
}

/* The persistent packet kernel by default (tuning path 0 or 4); the tile
 * kernels with paths 1 and 3; the lane-refill persistent kernel with path 2
 * (which leaves debug modes, the backward checkpoint and antialiasing to the
 * tile kernel).  A persistent kernel needs this stream's work counters; a
 * scene rendered from more than GT_WORK_SLOTS streams renders on the
 * further streams with the tile kernel (results never depend on the path). */
static int launch_frame(gt_scene* s, gt_dev_frame* f, void* stream) {
  const uint32_t tiles = gt_frame_tiles(&s->tune, f, 0);
  const uint32_t path = s->tune.path == 0 ? 4u : s->tune.path;
  unsigned int* work = (path == 2 || path == 4) ? gt_scene_work(s, stream) : nullptr;
  gt_tuning t = s->tune;
  t.path = path;
  if ((path == 2 || path == 4) && !work) t.path = 3;
  if (t.path != 2 || f->debug_mode != 0 || f->comp_ids || f->n_comp || f->antialias) {
    if (t.path == 2) t.path = 3;
    size_t words = 0;
    uint32_t* rects = gt_rects_wanted(&t, f->width, f->height, s->dev.n_prox)
                          ? gt_scene_rects(s, stream, gt_order_words(f->tiles_x, tiles / f->tiles_x), &words)
                          : nullptr;
    if (f->debug_mode == 0 && gt_colors_wanted(&t, f->width, f->height, s->dev.n)) {
      float* colors = scene_colors(s, stream);
      if (colors) {
        const int rc = gt_launch_colors(&s->dev, f, colors, stream);
        if (rc != GT_OK) return rc;
        f->colors = colors;
      }
    }
    return gt_launch_forward(&s->dev, f, nullptr, 1, tiles, &t, stream, 0, work, rects, words);
  }
  return gt_launch_persistent(&s->dev, f, nullptr, 1, tiles, work, (unsigned long long*)f->stats, &t, stream);
}

extern "C" int gt_scene_set_tuning(gt_scene* s, const gt_tuning* t) {
  if (!s || !t) { gt_set_error("gt_scene_set_tuning: NULL argument"); return GT_E_INVALID; }
  const bool tile_ok = t->tile_w == 0 || t->tile_w == 4 || t->tile_w == 8 || t->tile_w == 16 || t->tile_w == 32;
  const bool lds_ok = t->stack_lds == 0 || t->stack_lds == 8 || t->stack_lds == 16 || t->stack_lds == 32 ||
                      t->stack_lds == 64;
  if (t->path > 4 || t->refill_min > 64 || t->comp_min > 64 || t->blocks_per_cu > 8 || !tile_ok || !lds_ok ||
      t->rects > 2 || t->colors > 2) {
    gt_set_error("gt_scene_set_tuning: value out of range (path %u, refill %u, comp %u, blocks %u, tile_w %u, "
                 "stack_lds %u, rects %u, colors %u)", t->path, t->refill_min, t->comp_min, t->blocks_per_cu,
                 t->tile_w, t->stack_lds, t->rects, t->colors);
    return GT_E_INVALID;
  }
  s->tune = *t;
  return GT_OK;
}

extern "C" int gt_scene_get_tuning(const gt_scene* s, gt_tuning* t) {
  if (!s || !t) { gt_set_error("gt_scene_get_tuning: NULL argument"); return GT_E_INVALID; }
  *t = s->tune;
  return GT_OK;
}

const gt_dev_scene* gt_scene_dev(const gt_scene* s) { return &s->dev; }
const gt_tuning* gt_scene_tuning(const gt_scene* s) { return &s->tune; }

static void* grow(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return b.p;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  const hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    gt_set_error("backward scratch hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return nullptr;
  }
  b.bytes = bytes;
  return b.p;
}

/* Backward scratch of at least `bytes` (contents not kept when it grows),
 * allocated on first use and kept with the scene; NULL (with the error set)
 * when the allocation fails.  `which` 0: accumulators and colours, 1: the
 * recompute mode's band buffers. */
void* gt_scene_bwd_scratch(gt_scene* s, int which, size_t bytes) { return grow(which ? s->rc : s->bwd, bytes); }
uint32_t* gt_scene_rc_cap(gt_scene* s) { return &s->rc_cap; }

extern "C" int gt_render(gt_scene* s, const gt_uniforms* u, uint32_t width, uint32_t height, const gt_frame* out,
                         int buf_flags, void* stream) {
  if (!s || !out || !out->rgba) { gt_set_error("gt_render: scene, out and out->rgba are required"); return GT_E_INVALID; }
  int rc = gt_validate_uniforms(u, width, height);
  if (rc != GT_OK) return rc;
  if ((out->hit_ids || out->accept) && out->max_rounds == 0) {
    gt_set_error("gt_render: max_rounds must be >= 1 when hit_ids/accept are requested");
    return GT_E_INVALID;
  }
  if (out->comp_ids && (!out->n_comp || out->comp_cap == 0)) {
    gt_set_error("gt_render: the checkpoint needs comp_ids, n_comp and comp_cap >= 1");
    return GT_E_INVALID;
  }
  if ((out->comp_ids || out->n_comp) && u->debug_mode != 0) {
    gt_set_error("gt_render: the checkpoint is written in debugMode 0 only");
    return GT_E_INVALID;
  }
  if (out->antialiasing && (out->comp_ids || out->n_comp)) {
    gt_set_error("gt_render: antialiasing renders do not write the backward checkpoint");
    return GT_E_INVALID;
  }
  gt_dev_frame f;
  gt_fill_dev_frame(u, width, height, &f);
  f.antialias = out->antialiasing ? 1u : 0u;
  const size_t R = (size_t)width * height;
  const size_t MR = out->max_rounds;
  if (buf_flags == GT_BUF_DEVICE) {
    f.rgba = out->rgba;
    f.hit_ids = out->hit_ids;
    f.accept = out->accept;
    f.n_rounds = out->n_rounds;
    f.T_out = out->T;
    f.stats = (unsigned long long*)out->stats;
    f.max_rounds = (uint32_t)MR;
    f.comp_ids = out->comp_ids;
    f.n_comp = out->n_comp;
    f.comp_cap = out->comp_cap;
    if (f.antialias && f.accumulated_frames) { /* render into scratch, then average into rgba */
      if ((rc = gt_devbuf_reserve(s->aa, R * 16)) != GT_OK) return rc;
      f.rgba = (float*)s->aa.p;
      if ((rc = launch_frame(s, &f, stream)) != GT_OK) return rc;
      return gt_launch_aa_mix(f.rgba, out->rgba, (uint32_t)R, f.accumulated_frames, stream);
    }
    return launch_frame(s, &f, stream);
  }
  // host buffers: device temporaries, synchronous
  struct Tmp {
    void* p = nullptr;
    ~Tmp() { if (p) (void)hipFree(p); }
  } t_hits, t_acc, t_rounds, t_T, t_stats, t_cids, t_ncomp;
  auto alloc = [&](Tmp& t, size_t bytes) -> int {
    hipError_t e = hipMalloc(&t.p, std::max<size_t>(bytes, 4));
    if (e != hipSuccess) {
      t.p = nullptr;
      gt_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
      return GT_E_NOMEM;
    }
    return GT_OK;
  };
  /* the image goes through a buffer the scene keeps: no allocation per
   * frame once the largest frame size has been seen */
  if ((rc = gt_devbuf_reserve(s->himg, R * 16)) != GT_OK) return rc;
  f.rgba = (float*)s->himg.p;
  if (out->hit_ids) {
    if ((rc = alloc(t_hits, R * MR * u->k * 4)) != GT_OK) return rc;
    f.hit_ids = (uint32_t*)t_hits.p;
  }
  if (out->accept) {
    if ((rc = alloc(t_acc, R * MR * 4)) != GT_OK) return rc;
    f.accept = (uint32_t*)t_acc.p;
  }
  if (out->n_rounds) {
    if ((rc = alloc(t_rounds, R * 4)) != GT_OK) return rc;
    f.n_rounds = (uint32_t*)t_rounds.p;
  }
  if (out->T) {
    if ((rc = alloc(t_T, R * 4)) != GT_OK) return rc;
    f.T_out = (float*)t_T.p;
  }
  if (out->stats) {
    if ((rc = alloc(t_stats, GT_NSTATS * 8)) != GT_OK) return rc;
    HIP_TRY(hipMemset(t_stats.p, 0, GT_NSTATS * 8));
    f.stats = (unsigned long long*)t_stats.p;
  }
  if (out->comp_ids) {
    if ((rc = alloc(t_cids, R * out->comp_cap * 4)) != GT_OK) return rc;
    f.comp_ids = (uint32_t*)t_cids.p;
    f.comp_cap = out->comp_cap;
  }
  if (out->n_comp) {
    if ((rc = alloc(t_ncomp, R * 4)) != GT_OK) return rc;
    f.n_comp = (uint32_t*)t_ncomp.p;
  }
  f.max_rounds = (uint32_t)MR;
  if ((rc = launch_frame(s, &f, nullptr)) != GT_OK) return rc;
  if (f.antialias && f.accumulated_frames) { /* average the new frame into the caller's image */
    Tmp t_avg;
    if ((rc = alloc(t_avg, R * 16)) != GT_OK) return rc;
    HIP_TRY(hipMemcpy(t_avg.p, out->rgba, R * 16, hipMemcpyHostToDevice));
    if ((rc = gt_launch_aa_mix(f.rgba, (float*)t_avg.p, (uint32_t)R, f.accumulated_frames, nullptr)) != GT_OK)
      return rc;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out->rgba, t_avg.p, R * 16, hipMemcpyDeviceToHost));
  } else {
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out->rgba, f.rgba, R * 16, hipMemcpyDeviceToHost));
  }
  if (out->hit_ids) HIP_TRY(hipMemcpy(out->hit_ids, f.hit_ids, R * MR * u->k * 4, hipMemcpyDeviceToHost));
  if (out->accept) HIP_TRY(hipMemcpy(out->accept, f.accept, R * MR * 4, hipMemcpyDeviceToHost));
  if (out->n_rounds) HIP_TRY(hipMemcpy(out->n_rounds, f.n_rounds, R * 4, hipMemcpyDeviceToHost));
  if (out->T) HIP_TRY(hipMemcpy(out->T, f.T_out, R * 4, hipMemcpyDeviceToHost));
  if (out->stats) HIP_TRY(hipMemcpy(out->stats, f.stats, GT_NSTATS * 8, hipMemcpyDeviceToHost));
  if (out->comp_ids) HIP_TRY(hipMemcpy(out->comp_ids, f.comp_ids, R * out->comp_cap * 4, hipMemcpyDeviceToHost));
  if (out->n_comp) HIP_TRY(hipMemcpy(out->n_comp, f.n_comp, R * 4, hipMemcpyDeviceToHost));
  return GT_OK;
}

/* Self-check of the face pre-test on the host: for n (proxy, ray) pairs,
 * cand[i] = gt_face_candidates mask, exact[i] = faces accepted by the exact
 * triangle test + facing rule over (tmin, tmax).  The kernel evaluates the
 * same shared functions; exact & ~cand must be 0. */
extern "C" int gt_debug_face_test(const float* w2o, const float* o, const float* d, uint32_t n, float tmin, float tmax,
                                  float margin_scale, uint32_t* cand, uint32_t* exact) {
  if (!w2o || !o || !d || !cand || !exact) { gt_set_error("gt_debug_face_test: NULL argument"); return GT_E_INVALID; }
  if (!gt_ico_literals_match()) { gt_set_error("icosahedron literal tables out of date"); return GT_E_LIMIT; }
  for (uint32_t i = 0; i < n; ++i) {
    const float* m = w2o + 12 * (size_t)i;
    const gt_f3 ro = gt_f3_make(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
    const gt_f3 rd = gt_f3_make(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
    const gt_f3 oo = gt_xform_point(m, ro), od = gt_xform_dir(m, rd);
    float mg, mn;
    gt_face_margins(oo, od, m, &mg, &mn);
    cand[i] = gt_face_candidates(oo, od, mg * margin_scale, mn * margin_scale);
    uint32_t ex = 0;
    for (int f = 0; f < GT_NUM_FACES; ++f) {
      const gt_face fc = gt_ico_face(f);
      float t;
      if (gt_tri_hit(oo, od, tmin, tmax, fc, &t) && gt_front_facing(m, fc.n, rd)) ex |= 1u << f;
    }
    exact[i] = ex;
  }
  return GT_OK;
}
