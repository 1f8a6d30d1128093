/*
 * gt_internal.h - layouts shared by the host builder (gt_scene.cpp) and the
 * HIP kernels (gt_forward.hip).
 *
 * HBM layout of a scene (all arrays 256-B aligned):
 *   geom   N  x 48 B  {mu.xyz, opacity} {scale.xyz, 0} {quat r x y z}
 *                     (what the FP64 max-response evaluation reads)
 *   sh     N  x 192 B sh[16][3] without the vec3 padding of the 320-B record
 *   prox   N' x 64 B  {w2o c0.xyz c1.x} {c1.yz c2.xy} {c2.z c3.xyz} {gid 0 0 0}
 *                     stored in BVH leaf order, so a leaf is a contiguous run
 *   nodes  M  x 128 B 4-wide BVH node: child boxes as SoA float4 rows
 *   faces  20 x 48 B  icosahedron face constants {p0, e0, e1, n}
 */
#ifndef GT_INTERNAL_H
#define GT_INTERNAL_H

#include <stdint.h>

#include "../../include/gt.h"

#define GT_NODE_EMPTY 0xffffffffu
#define GT_LEAF_BIT 0x80000000u
#ifndef GT_LEAF_COUNT_BITS
#define GT_LEAF_COUNT_BITS 3          /* leaf ref: 1 | (count-1):BITS | first:(31-BITS) */
#endif
#ifndef GT_LEAF_MAX
#define GT_LEAF_MAX (1u << GT_LEAF_COUNT_BITS) /* proxies per leaf */
#endif
#define GT_LEAF_COUNT_SHIFT (31 - GT_LEAF_COUNT_BITS)
#define GT_LEAF_COUNT_MASK ((1u << GT_LEAF_COUNT_BITS) - 1u)
#define GT_LEAF_FIRST_MASK ((1u << GT_LEAF_COUNT_SHIFT) - 1u)
#define GT_MAX_PROXIES GT_LEAF_FIRST_MASK
#define GT_STACK_MAX 96               /* traversal stack entries per ray */
/* Work counters of the persistent renderers: one per XCD (gfx950 has 8),
 * GT_WORK_STRIDE words apart so each sits in its own 64-B line. */
#define GT_XCDS 8u
#define GT_WORK_STRIDE 16u
#define GT_WORK_WORDS (GT_XCDS * GT_WORK_STRIDE)
/* SAH cost of a proxy test relative to a 4-wide node visit (host and
 * device builders).  0.5 measured fastest at 1M Gaussians, 1080p, k 16:
 * 14.6 ms per frame against 15.8 ms at 2.0 (host tree) and 15.4 against
 * 16.7 ms (GPU tree).  Most proxy tests end at the bounding-sphere test,
 * so larger leaves and fewer node visits pay. */
#ifndef GT_SAH_ISECT
#define GT_SAH_ISECT 0.5
#endif
#ifndef GT_SAH_LEAF
#define GT_SAH_LEAF 0.0 /* SAH cost of a leaf besides its proxy tests (C_L A), both builders */
#endif

static inline uint32_t gt_leaf_ref(uint32_t first, uint32_t count) {
  return GT_LEAF_BIT | ((count - 1u) << GT_LEAF_COUNT_SHIFT) | first;
}

/* 128-B BVH4 node.  Row r of lo/hi holds axis r of the four child boxes. */
typedef struct gt_node4 {
  float lo_x[4], hi_x[4];
  float lo_y[4], hi_y[4];
  float lo_z[4], hi_z[4];
  uint32_t child[4]; /* inner node index, leaf ref, or GT_NODE_EMPTY */
  uint32_t _pad[4];
} gt_node4;

typedef struct gt_dev_scene {
  const float* geom;     /* N x 12 */
  const float* sh;       /* N x 48 */
  const float* prox;     /* N' x 16 */
  const gt_node4* nodes; /* M */
  const float* faces;    /* 20 x 12 */
  uint32_t root;         /* ref of the root (inner index, leaf ref or EMPTY) */
  uint32_t n, n_prox, n_nodes;
} gt_dev_scene;

typedef struct gt_dev_frame {
  float to_world[16];
  float ax, ay;
  float t_min, t_max, T_min, alpha_min;
  uint32_t k, sh_degree, early_stop, debug_mode;
  uint32_t width, height;
  float slack_abs; /* box-test slack, |camera origin| dependent */
  uint32_t antialias;          /* nonzero: jittered rays, running average into rgba */
  uint32_t accumulated_frames; /* frames already averaged into rgba (AA) */
  uint32_t max_rounds;
  uint32_t tiles_x;   /* 8x8 pixel tiles per row */
  uint32_t tile_base; /* first global tile index of this camera in a batch */
  /* optional checkpoint for the backward pass: ids of the composited hits of
   * pixel p (in order) at comp_ids[(p - comp_base) * comp_cap + i], their
   * number at n_comp[p - comp_base] (may exceed comp_cap: overflow) */
  uint32_t* comp_ids;
  uint32_t* n_comp;
  uint32_t comp_cap;
  uint32_t comp_base;
  float* rgba;
  uint32_t* hit_ids;
  uint32_t* accept;
  uint32_t* n_rounds;
  float* T_out;
  unsigned long long* stats;
  /* NULL, or the frame's per-Gaussian colours (float4, from the camera
   * origin, gt_launch_colors) that a one-camera render composites instead
   * of evaluating each hit's SH record */
  const float* colors;
} gt_dev_frame;

#ifdef __cplusplus
extern "C" {
#endif
/* implemented in gt_forward.hip */
/* Tile renderer over n_cams cameras (frame F used directly when n_cams == 1,
 * else the device array cams_dev with tile_base/tiles_x set for tune's
 * tile shape); total_tiles = all cameras' tiles.  Tuning path 4 renders the
 * same tiles with the persistent packet kernel, whose queue counters are
 * `work` (GT_WORK_WORDS device words, reset by the launch; required for
 * path 4, unused otherwise).  rect_buf: NULL, or device memory of
 * rect_words 32-bit words, at least 2 GT_RECT_ENTRIES(n_prox), that a
 * one-camera launch of the packet traversal may fill with the frame's proxy
 * rectangles (gt_rect.h) and cull with; with gt_order_words more, a
 * persistent launch also orders its tile queue by the rectangles' coverage.
 * Results never depend on either. */
#define GT_RECT_ENTRIES(n_prox) ((size_t)(n_prox) + GT_LEAF_MAX)
/* Cost-ordered tile queue of a one-camera persistent frame (gt_forward.hip).
 * Coverage is counted on a grid of cells of cell x cell tiles, the
 * smallest cell for which the grid fits GT_COVER_CELLS counters (LDS grids),
 * by GT_COVER_PARTS workgroups into partial grids that one workgroup then
 * sums and orders.  The workspace (32-bit words, after the rectangles in
 * rect_buf) holds the partial grids and the order (one word per tile). */
#define GT_ORDER_LEVELS 512u
#define GT_ORDER_BUCKETS (GT_XCDS * GT_ORDER_LEVELS)
#define GT_COVER_CELLS 8192u
#define GT_COVER_PARTS 64u
static inline uint32_t gt_cover_cell(uint32_t tiles_x, uint32_t tiles_y) {
  uint32_t c = 1;
  while ((size_t)((tiles_x + c - 1u) / c) * ((tiles_y + c - 1u) / c) > GT_COVER_CELLS) ++c;
  return c;
}
static inline size_t gt_order_words(uint32_t tiles_x, uint32_t tiles_y) {
  const uint32_t c = gt_cover_cell(tiles_x, tiles_y);
  const size_t cells = (size_t)((tiles_x + c - 1u) / c) * ((tiles_y + c - 1u) / c);
  return GT_COVER_PARTS * cells + (size_t)tiles_x * tiles_y;
}
/* does a one-camera frame of width x height use rectangles (tuning field
 * rects: 0 when the frame has at least a quarter as many rays as the scene
 * has proxies, 1 always, 2 never)?  Measured with the tile ordering, views
 * from outside the scene: 1280x720 gains 5-27 % at 1M and 3M proxies,
 * 640x480 is even at 1M, and smaller frames lose 8-28 % (few tiles, little
 * to cull, and the leaf-record prefetch that rectangles turn off matters
 * most to waves that run alone). */
static inline int gt_rects_wanted(const gt_tuning* t, uint32_t width, uint32_t height, uint32_t n_prox) {
  const uint32_t mode = t ? t->rects : 0u;
  if (mode == 2u || n_prox == 0) return 0;
  return mode == 1u || 4ull * width * height >= n_prox;
}
/* Precomputed colours pay when the frame composites many hits per
 * Gaussian: a colour costs about 85 ps per Gaussian to compute and saves
 * about 14 ps per composited hit (1M Gaussians, 1080p, k 16: 17 hits per
 * ray), so they are used from half as many rays as Gaussians. */
static inline int gt_colors_wanted(const gt_tuning* t, uint32_t width, uint32_t height, uint32_t n) {
  const uint32_t mode = t ? t->colors : 0u;
  if (mode == 2u || n == 0) return 0;
  return mode == 1u || 2ull * width * height >= n;
}
int gt_launch_forward(const gt_dev_scene* s, const gt_dev_frame* f, const gt_dev_frame* cams_dev, uint32_t n_cams,
                      uint32_t total_tiles, const gt_tuning* tune, void* stream, uint32_t tile_offset,
                      unsigned int* work, uint32_t* rect_buf, size_t rect_words);
/* Persistent renderer over n_cams cameras sharing k.  cams_dev: device
 * array of n_cams frames (tile_base set, ascending); cam0: host copy of
 * frame 0, used directly when n_cams == 1.  work: one device counter the
 * launch resets and consumes; stats: optional device counters. */
int gt_launch_persistent(const gt_dev_scene* s, const gt_dev_frame* cam0, const gt_dev_frame* cams_dev,
                         uint32_t n_cams, uint32_t total_tiles, unsigned int* work, unsigned long long* stats,
                         const gt_tuning* tune, void* stream);
/* out[i] (float4, N entries) = FP32 radiance of Gaussian i seen from f's
 * camera origin, bit-identical to the colour the render kernels composite. */
int gt_launch_colors(const gt_dev_scene* s, const gt_dev_frame* f, float* out, void* stream);
/* Uploads the icosahedron face table to the current device once (a
 * synchronous copy: call it outside any stream capture); GT_OK or GT_E_HIP. */
int gt_forward_prepare(void);
/* avg[i] = gt_aa_mix(avg[i], img[i], frames) per RGB channel, A = 1 (float4 pixels) */
int gt_launch_aa_mix(const float* img, float* avg, uint32_t n_pixels, uint32_t frames, void* stream);
void gt_set_error(const char* fmt, ...);
#ifdef __cplusplus
}
#endif

#endif
