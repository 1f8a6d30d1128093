/*
 * gt_scene_impl.h - the scene handle behind gt_scene* (C++ only), shared by
 * the host builder and render wrappers (gt_scene.cpp) and the device
 * builder (gt_build.hip).
 */
#ifndef GT_SCENE_IMPL_H
#define GT_SCENE_IMPL_H

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "gt_host_build.h"
#include "gt_internal.h"

#define GT_WORK_SLOTS 32u

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
};

/* Device-build workspace (gt_build.hip): temporaries kept between rebuilds. */
struct gt_devbuild;
void gt_devbuild_destroy(gt_devbuild* w);

struct gt_scene {
  HostScene h;                /* host-built scenes only */
  uint32_t n = 0;             /* Gaussians */
  bool device_built = false;  /* built on the GPU from raw_dev (gt_scene_create_device) */
  double alpha_min = 0.0;
  // device
  int device = 0;
  DevBuf geom, sh, prox, nodes, faces;
  DevBuf raw_dev; /* device-built scenes: copy of the n x 59 PLY-raw parameters */
  gt_devbuild* db = nullptr;
  /* work counters of the persistent renderers: one set of GT_WORK_WORDS per
   * stream that rendered the scene (up to GT_WORK_SLOTS streams), so renders
   * of one scene on different streams never share a queue */
  DevBuf ctl;
  void* ctl_streams[GT_WORK_SLOTS] = {};
  uint32_t ctl_used = 0;
  std::mutex ctl_mu;
  /* the proxy rectangles of one-camera frames, one table per stream slot
   * of ctl (sized on the stream's first render after a rebuild) */
  DevBuf rects[GT_WORK_SLOTS];
  /* the per-Gaussian colours of a stream's current frame, one per stream
   * slot of ctl (sized on the stream's first render after a rebuild) */
  DevBuf colors[GT_WORK_SLOTS];
  DevBuf bwd; /* backward scratch: per-Gaussian accumulators and colours */
  DevBuf rc;  /* recompute-mode scratch: band checkpoint, counts, image */
  DevBuf aa;  /* antialiasing: the new frame before it is averaged in */
  DevBuf himg; /* gt_render with host buffers: the frame on the device */
  uint32_t rc_cap = 0; /* checkpoint capacity the last recompute needed */
  gt_tuning tune{};
  gt_dev_scene dev{};
  uint32_t n_nodes = 0, max_stack = 0, depth = 0;
  double build_ms = 0.0;
  /* refit schedule (gt_scene_refit_device): the nodes of level L are entries
   * [rlevels[L], rlevels[L + 1]) of rorder, or those indices themselves when
   * ridentity (the device builder numbers nodes level by level) */
  std::vector<uint32_t> rlevels;
  DevBuf rorder;
  bool ridentity = false;
  bool rsched = false; /* schedule valid for the current tree */
  ~gt_scene() { gt_devbuild_destroy(db); }
};

/* gt_scene.cpp: (re)allocates b to exactly max(bytes, 256) bytes unless it
 * already holds at least bytes; GT_OK or an error code (message set). */
int gt_devbuf_reserve(DevBuf& b, size_t bytes);
/* uploads the icosahedron face table and the persistent renderer's counter once */
int gt_scene_static_upload(gt_scene* s);

/* gt_build.hip: builds every device array of s from s->raw_dev on the GPU */
int gt_device_build(gt_scene* s, hipStream_t stream);
/* gt_build.hip: new parameters, same tree topology (boxes refit); falls back
 * to gt_device_build when the set of proxies changed (*rebuilt = 1) */
int gt_device_refit(gt_scene* s, const float* raw_dev, hipStream_t stream, uint32_t* rebuilt);
/* gt_build.hip: host copies of a device-built scene (gt_scene_get_host_data) */
int gt_device_host_data(const gt_scene* s, float* packed, uint32_t* proxy_ids, float* o2w, float* w2o);

#endif
