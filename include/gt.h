/*
 * gt.h - C-ABI of the MI355X Gaussian ray tracer (libgt.so).
 *
 * Drop-in for the reference's ray-tracing path: everything the WebRTX
 * extension to WebGPU plus the app's frame loop do for one frame
 *   createRayTracingAccelerationContainer + hostBuildRayTracingAccelerationContainer
 *       (webrtx/src/patch.ts:211-219, src/main.ts:30-111)      -> gt_scene_create_*
 *   uniform block update (src/main.ts:520-534)                 -> gt_uniforms_init
 *   beginRayTracingPass().traceRays(device, sbt, W, H)
 *       (webrtx/src/pass_encoder.ts:35-65, src/main.ts:563-593) -> gt_render
 * behind plain C types: pointers, sizes, status codes.  No C++ exception
 * crosses this boundary; every entry point returns GT_OK (0) or a negative
 * GT_E* code, and gt_last_error() gives the calling thread's message
 * (the reference throws JS strings, pass_encoder.ts:27-57).
 *
 * Threading: calls on one scene handle are serialised by the caller (as the
 * reference's single JS thread does); different handles may be used from
 * different threads.  Buffers are caller-owned.  With GT_BUF_DEVICE the
 * pointers are HIP device pointers and the call is asynchronous on the given
 * stream; otherwise they are host pointers and the call is synchronous.
 * Renders of one scene may overlap on the GPU when issued on different
 * streams: the persistent kernel keeps its queue counters per stream (a
 * batch keeps its own), so only one batch replayed on two streams at once
 * would share them.
 */
#ifndef GT_H
#define GT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_OK 0
#define GT_E_INVALID (-1)   /* bad argument (null pointer, size, range)        */
#define GT_E_HIP (-2)       /* HIP runtime error                               */
#define GT_E_NOMEM (-3)     /* allocation failed                               */
#define GT_E_IO (-4)        /* file could not be read / parsed                 */
#define GT_E_LIMIT (-5)     /* exceeds a documented limit (k, rounds, stack)   */

#define GT_RAW_FLOATS 59    /* PLY-raw record: x y z f_dc_0..2 f_rest_0..44 opacity scale_0..2 rot_0..3 */
#define GT_PACKED_BYTES 320 /* GaussianParticle record, src/GS_shader.ts:81-87, src/GS_ply.ts:199-205 */
#define GT_MAX_K 32         /* hit_array_size limit (the reference compiles MAX_K = k, GS_shader.ts:377) */
#define GT_SENTINEL_ID 0xffffffffu /* empty k-buffer slot, GS_shader.ts:239 */

/* Buffer flags */
#define GT_BUF_HOST 0
#define GT_BUF_DEVICE 1

typedef struct gt_scene gt_scene;

/* 112-byte uniform block, byte for byte as src/GS_shader.ts:96-108 and the
 * packing of src/main.ts:516-534 (fview/uview indices in comments). */
typedef struct gt_uniforms {
  float to_world[16];          /* 0   camera matrix, column-major (camera.ts)  [0..15] */
  float camera_aspect[2];      /* 64  tan(hFov/2), tan(hFov/2)*H/W            [16,17] */
  float t_min;                 /* 72  [18] */
  float t_max;                 /* 76  [19] */
  float T_min;                 /* 80  [20] */
  float alpha_min;             /* 84  [21] */
  uint32_t k;                  /* 88  hit_array_size [22] */
  uint32_t sh_degree_max;      /* 92  [23] */
  uint32_t accumulated_frames; /* 96  [24] (AA is compiled out in the reference) */
  uint32_t early_stop;         /* 100 0xffffffff = never [25] */
  uint32_t debug_mode;         /* 104 0 render, 1 closest proxy, 2 hit count [26] */
  uint32_t _pad;               /* 108 */
} gt_uniforms;

/* Fills the block as upadateUniformData does (src/main.ts:520-534):
 * aspect = (f32(tan(0.5*hFov*pi/180)), f32(aspect.x * H / W)),
 * early_stop <= 0 -> 0xffffffff. */
void gt_uniforms_init(gt_uniforms* u, const float to_world[16], double hfov_deg, uint32_t width,
                      uint32_t height, float t_min, float t_max, float T_min, float alpha_min,
                      uint32_t k, uint32_t sh_degree_max, int32_t early_stop, uint32_t debug_mode);

/* WASDCamera matrix (src/camera.ts:126-140, 153-204) for a camera at pos
 * looking at target after its first update(): rotY(yaw)*rotX(pitch) with
 * yaw = atan2(b.x, b.z), pitch = -asin(b.y), b = normalize(pos - target). */
void gt_camera_look_at(const float pos[3], const float target[3], float to_world[16]);

/* ---- scene (acceleration container) ---- */

/* Every scene entry point checks its records: a non-finite value, a
 * quaternion of zero norm, or a scale exp(log-scale) that is 0 or inf in
 * f32 (for packed records: a non-finite field, a scale <= 0 or a zero
 * quaternion) is GT_E_INVALID, and gt_last_error() names the first such
 * Gaussian.  So is a Gaussian with a proxy (opacity > alpha_min) whose
 * proxy transform, its inverse, its world box or the box centre is not
 * finite in f32: a position near the f32 limit, a scale whose product with
 * sqrt(2 ln(opacity / alpha_min)) overflows, or one so small that the
 * inverse overflows.  The reference has no check (GS_ply.ts:225-235
 * divides by the quaternion norm and builds NaN proxies).  After such an
 * error from gt_scene_update_device / gt_scene_refit_device the scene's
 * contents are unspecified until an update with valid parameters succeeds. */

/* From PLY-raw parameters (n x 59 floats): applies the activations of
 * src/GS_ply.ts:225-235 in double, packs 320-B records, filters proxies by
 * opacity > alpha_min and builds their transforms as src/main.ts:73-94 does,
 * then builds and uploads the BVH.  alpha_min is the JS double of
 * main.ts:463 (the uniform's f32 alpha_min is separate). */
int gt_scene_create_raw(const float* raw, uint32_t n, double alpha_min, gt_scene** out);

/* From already packed 320-B records (n x 320 bytes, the reference's
 * gsData.gaussiansBuffer).  Proxy transforms are then built from the f32
 * activated values, since the JS doubles are not part of the buffer. */
int gt_scene_create_packed(const void* records, uint32_t n, double alpha_min, gt_scene** out);

/* Re-filter and rebuild for a new alpha_min (the reference rebuilds the TLAS). */
int gt_scene_set_alpha_min(gt_scene* s, double alpha_min);

/* The same scene built on the GPU from DEVICE-resident PLY-raw parameters
 * (raw_dev: n x 59 floats), for callers whose parameters live in HBM (a
 * training loop).  Replaces the host stages gaussianSceneToBvh
 * (src/main.ts:30-111) and the wasm TLAS build
 * (webrtx/src/wasm_bvh_builder.ts:260-377, webrtx/bvh/src/lib.rs:253-347):
 * activations, records, proxy transforms and boxes are computed by the same
 * code as gt_scene_create_raw (bit-identical); the BVH is a GPU radix tree
 * over Morton codes whose subtrees of up to 256 leaves are rebuilt by
 * locally-ordered clustering, with SAH-chosen leaves and the least-SAH-cost
 * collapse to the same 4-wide node format.  Hit lists and pixels equal
 * those of the host-built scene.
 * raw_dev is copied into the scene; the call returns when the build is done
 * (it synchronises hip_stream).  gt_scene_set_alpha_min rebuilds on the GPU. */
int gt_scene_create_device(const float* raw_dev, uint32_t n, double alpha_min, void* hip_stream, gt_scene** out);

/* gt_scene_create_raw's inputs (HOST PLY-raw floats) built on the GPU as
 * gt_scene_create_device does: one upload, then the GPU build (synchronous).
 * For hosts without device memory (the JavaScript host); results equal
 * gt_scene_create_raw's. */
int gt_scene_create_raw_gpu(const float* raw, uint32_t n, double alpha_min, gt_scene** out);

/* Rebuilds a GPU-built scene (gt_scene_create_device or _raw_gpu) from
 * updated DEVICE parameters (same n, same alpha_min), reusing its device
 * memory: the step after an optimizer update.  Must not overlap renders or
 * backward calls on the scene. */
int gt_scene_update_device(gt_scene* s, const float* raw_dev, void* hip_stream);

/* Updates any scene (host- or GPU-built) from DEVICE parameters (same n, same
 * alpha_min) and keeps its tree: records, proxy transforms and proxy boxes
 * are recomputed, and the node boxes are refit level by level.  Which hits a
 * ray finds does not depend on the tree, so renders equal those of a fresh
 * build of the same parameters, bit for bit.  A host-built tree keeps its
 * binned-SAH quality, and a refit costs less than a rebuild.  If the
 * set of Gaussians with a proxy changed, the scene is rebuilt on the GPU
 * instead and *rebuilt (if not NULL) is set to 1, else 0.  The scene then
 * holds device data only (gt_scene_get_host_data downloads it).  Returns
 * when done (synchronises hip_stream).  No counterpart in the reference,
 * which does not train (README.md:68-69). */
int gt_scene_refit_device(gt_scene* s, const float* raw_dev, void* hip_stream, uint32_t* rebuilt);

void gt_scene_destroy(gt_scene* s);

typedef struct gt_scene_info {
  uint32_t n_gaussians;
  uint32_t n_proxies;       /* opacity > alpha_min */
  uint32_t n_nodes;         /* wide BVH nodes */
  uint32_t max_stack;       /* traversal stack entries the BVH can need */
  uint32_t bvh_depth;
  uint32_t _pad;
  uint64_t device_bytes;    /* HBM held by the scene */
  double build_ms;          /* last build time (activations, proxies, BVH; host or GPU) */
} gt_scene_info;
int gt_scene_get_info(const gt_scene* s, gt_scene_info* info);

/* Host copies of what the scene built (for tests): packed records
 * (n x 80 floats), proxy ids / object-to-world / world-to-object
 * (n_proxies, n_proxies x 12 floats each, column-major 4x3).  Any may be NULL. */
int gt_scene_get_host_data(const gt_scene* s, float* packed, uint32_t* proxy_ids, float* o2w, float* w2o);

/* Host stage of gt_scene_create_raw, callable without a GPU: activations,
 * 320-B packing and proxy construction (src/GS_ply.ts:225-252,
 * src/main.ts:73-94).  packed: n x 80 floats; proxy_ids: n; o2w, w2o:
 * n x 12 floats; *n_proxies receives how many were written. */
int gt_host_prepare(const float* raw, uint32_t n, double alpha_min, float* packed, uint32_t* proxy_ids, float* o2w,
                    float* w2o, uint32_t* n_proxies);

/* Self-check of the kernel's face pre-test on the host (no counterpart in
 * the reference; it guards the bit-exactness of hit lists): for n pairs of
 * a proxy world-to-object matrix (12 floats) and a world ray (o, d), writes
 * the candidate-face mask and the mask of faces the exact triangle test and
 * facing rule accept in (tmin, tmax).  With margin_scale 1 (the kernel's
 * margins) exact & ~cand is 0 by design; smaller scales show how close the
 * inputs come to the FP32 error regime. */
int gt_debug_face_test(const float* w2o, const float* o, const float* d, uint32_t n, float tmin, float tmax,
                       float margin_scale, uint32_t* cand, uint32_t* exact);

/* Launch tuning of a scene (no counterpart in the reference; chosen per
 * scene by the autotuner).  Zero fields mean "library default".  None of
 * them changes a result: hit lists and pixels are identical for all values.
 *   path          0 default (= 4), 1 tile kernel (one ray tile per wave,
 *                 per-lane traversal), 2 persistent kernel whose lanes refill
 *                 from a global ray queue (per-lane traversal), 3 tile kernel
 *                 with wave-wide packet traversal (the wave walks the BVH
 *                 together from one stack of (node, lane mask)), 4 persistent
 *                 kernel with packet traversal: resident waves take 64-ray
 *                 tiles from a global queue (one counter per XCD)
 *   refill_min    path 2: idle lanes of a wave that trigger a refill (1..64)
 *   comp_min      path 2: lanes waiting to composite that trigger a
 *                 composite pass (1..64)
 *   blocks_per_cu paths 2 and 4: resident 256-thread blocks per CU (1..8; 0 = one per
 *                 wave slot: 4 on paths 2 and 4)
 *   tile_w        ray tile of a wave: tile_w x (64 / tile_w) pixels, tile_w in {4, 8, 16, 32}
 *   stack_lds     traversal stack entries per ray kept in LDS (8, 16, 32 or 64);
 *                 deeper entries spill to scratch memory
 *   rects         one-camera frames of paths 3 and 4: 0 cull leaf proxies by
 *                 per-frame screen rectangles (and, on path 4, claim tiles in
 *                 order of decreasing rectangle coverage) when the frame has
 *                 at least a quarter as many rays as the scene has proxies,
 *                 1 for every frame, 2 never
 *   colors        one-camera rendering frames of paths 1, 3 and 4: 0 composite
 *                 from per-Gaussian colours computed for the frame's origin
 *                 before the traversal when the frame has at least half as
 *                 many rays as the scene has Gaussians (else each composited
 *                 hit evaluates its SH record), 1 for every frame, 2 never;
 *                 the colours are the same FP32 values either way */
typedef struct gt_tuning {
  uint32_t path;
  uint32_t refill_min;
  uint32_t comp_min;
  uint32_t blocks_per_cu;
  uint32_t tile_w;
  uint32_t stack_lds;
  uint32_t rects;
  uint32_t colors;
} gt_tuning;
int gt_scene_set_tuning(gt_scene* s, const gt_tuning* t);
int gt_scene_get_tuning(const gt_scene* s, gt_tuning* t);

/* ---- render (traceRays) ---- */

/* Per-frame outputs.  rgba: W*H float4, row-major y*W+x, A = 1
 * (GS_shader.ts:300-301).  Optional parity outputs (NULL to skip):
 *   hit_ids  W*H*max_rounds*k Gaussian ids per ray and round, sentinel-filled
 *   accept   W*H*max_rounds   bit i set = entry i was composited (alpha > alpha_min)
 *   n_rounds W*H              rounds run
 *   T        W*H              final transmittance
 *   stats    GT_NSTATS u64    accumulated counters (device pointer with GT_BUF_DEVICE) */
#define GT_NSTATS 8
/* stats: 0 rays, 1 rounds, 2 H_eval (evaluated k-buffer entries), 3 H_comp
 * (composited), 4 proxies tested, 5 front-face hits, 6 BVH nodes visited, 7 rounds beyond max_rounds */
typedef struct gt_frame {
  float* rgba;
  uint32_t* hit_ids;
  uint32_t* accept;
  uint32_t* n_rounds;
  float* T;
  uint64_t* stats;
  uint32_t max_rounds;
  uint32_t comp_cap; /* checkpoint capacity per ray (with comp_ids) */
  /* Optional checkpoint for gt_backward: ids of the composited hits of ray
   * y*W+x in order at comp_ids[(y*W+x)*comp_cap + i], their number at
   * n_comp[y*W+x] (a count above comp_cap means the list overflowed).
   * n_comp without comp_ids only counts, to size a checkpoint. */
  uint32_t* comp_ids;
  uint32_t* n_comp;
  /* Nonzero: the reference's antialiasing path (compiled out there,
   * GS_shader.ts:54, 282-307): each pixel's ray goes through a jittered
   * offset drawn from (y*W+x, u.accumulated_frames), and rgba, which holds
   * the average of the previous frames on input, receives
   * frames == 0 ? new : mix(old, new, 1 / (frames + 1)).  Single-camera
   * renders only (not batches), and not with the backward checkpoint. */
  uint32_t antialiasing;
  uint32_t _pad;
} gt_frame;

/* traceRays(W, H) of the reference with the uniform block u.  Unlike the
 * reference (pass_encoder.ts:52-56) W and H need not be multiples of 8;
 * limits: 1..16384 each, 1 <= k <= 32, t_min >= 0, sh_degree_max <= 3. */
int gt_render(gt_scene* s, const gt_uniforms* u, uint32_t width, uint32_t height, const gt_frame* out,
              int buf_flags, void* hip_stream);

/* ---- backward (no counterpart in the reference, which has no
 * differentiation; README.md:68-69 plans training) ----
 * Accumulates dLoss/d(raw parameters) of one frame into grads (n x 59
 * floats, the PLY-raw layout of gt_scene_create_raw: mean, f_dc, f_rest,
 * opacity logit, log-scale, raw quaternion) given dl_drgba (W*H float4,
 * alpha channel ignored: A = 1 is constant).  Hit lists are held fixed.
 * Checkpoint mode: comp_ids / n_comp / comp_cap from gt_render's outputs.
 * Rays whose list exceeds comp_cap are skipped and counted in *overflow
 * (device counter, optional).  Recompute mode (comp_ids NULL): the forward
 * traversal is re-run in bands of rows whose scratch checkpoint stays near
 * budget_bytes (0: 1 GiB), starting with comp_cap (0: 64) ids per ray and
 * growing it (with smaller bands, down to 8 rows) when a ray composites
 * more, so no ray overflows; the scratch (about budget_bytes + 20 B per
 * pixel) and the capacity stay with the scene until gt_scene_destroy.
 * Both modes also keep 80 B per Gaussian of accumulators with the scene,
 * so backward calls on one scene must not overlap.  All pointers are
 * device pointers; checkpoint mode is
 * asynchronous on the stream, recompute mode returns after the stream
 * drains.  FP32 atomics: the summation order across rays is unspecified. */
#define GT_GRAD_FLOATS 59
int gt_backward(gt_scene* s, const gt_uniforms* u, uint32_t width, uint32_t height, const float* dl_drgba,
                const uint32_t* comp_ids, const uint32_t* n_comp, uint32_t comp_cap, float* grads,
                unsigned int* overflow, uint64_t budget_bytes, void* hip_stream);

/* ---- training (no counterpart in the reference, which has no optimizer;
 * README.md:68-69 plans training) ----
 * One Adam update (Kingma & Ba 2015, the operation order of PyTorch's
 * torch.optim.Adam) in place on n x 59 PLY-raw parameters given their
 * gradient (gt_backward's layout), with one learning rate per field
 * (column; 0 keeps a field fixed).  m and v are the moment buffers (same
 * shape, zero before step 1).  Device pointers; asynchronous on the stream.
 * A training step is gt_render (with checkpoint) + dL/dRGBA + gt_backward
 * per view, gt_adam_step, then gt_scene_update_device. */
typedef struct gt_adam {
  float lr[GT_RAW_FLOATS];
  float beta1, beta2, eps;
  uint32_t step; /* 1 for the first update (bias correction) */
} gt_adam;
int gt_adam_step(float* params, const float* grads, float* m, float* v, uint32_t n, const gt_adam* adam,
                 void* hip_stream);

/* ---- batched cameras (no counterpart in the reference, which renders one
 * camera per traceRays dispatch, pass_encoder.ts:35-65) ---- */
typedef struct gt_batch gt_batch;

/* n cameras of uneven resolution sharing k and debugMode 0.  outs[i] holds
 * DEVICE pointers (rgba required; hit_ids/accept with max_rounds for all
 * cameras or none; outs[0].stats, if set, accumulates the batch's
 * counters).  The camera table is copied to HBM here. */
int gt_batch_create(gt_scene* s, uint32_t n, const gt_uniforms* u, const uint32_t* widths, const uint32_t* heights,
                    const gt_frame* outs, gt_batch** out);
/* Renders every camera with one launch on stream (asynchronous).  With
 * use_graph != 0 the launches are captured into a hipGraph on first use and
 * the graph is replayed afterwards; it is recaptured when the scene's tuning
 * changes or a rebuild has moved its device arrays (gt_scene_set_alpha_min,
 * gt_scene_update_device, a refit that rebuilds).  Each camera's results
 * equal gt_render on that camera. */
int gt_batch_render(gt_batch* b, int use_graph, void* hip_stream);
int gt_batch_rays(const gt_batch* b, uint64_t* rays);
void gt_batch_destroy(gt_batch* b);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gt_batch_create + gt_batch_render (no graph) as one
 * call that allocates nothing.  Same arguments, checks and results: each
 * camera's outputs equal gt_render on that camera.  The batch's scratch
 * (the camera table the render kernels read and the persistent renderer's
 * work counter) lives in a file-scope device array sized for
 * GT_NOALLOC_MAX_CAMERAS cameras (about 54 KB), which every GPU the library
 * runs on holds until the process ends; more cameras is GT_E_LIMIT.  All
 * calls on one GPU share that array, so calls that overlap on one GPU
 * (other streams, other threads) race: order them on one stream or
 * synchronise between them.  Asynchronous on hip_stream. */
#define GT_NOALLOC_MAX_CAMERAS 256
int gt_batch_render_noalloc(gt_scene* s, uint32_t n, const gt_uniforms* u, const uint32_t* widths,
                            const uint32_t* heights, const gt_frame* outs, void* hip_stream);

/* ---- autotuner (no counterpart in the reference) ----
 * Times the result-neutral gt_tuning candidates on the caller's workload
 * (iters timed renders each, after one warm-up) and stores the fastest in
 * the scene.  Device buffers; synchronises the stream. */
typedef struct gt_autotune_result {
  gt_tuning best;
  double best_ms;   /* per render, best candidate */
  double first_ms;  /* per render, first candidate (the default: persistent packet kernel, 8x8 tiles) */
  uint32_t n_tried;
  uint32_t _pad;
} gt_autotune_result;
int gt_autotune_frame(gt_scene* s, const gt_uniforms* u, uint32_t width, uint32_t height, float* rgba_dev,
                      uint32_t iters, void* hip_stream, gt_autotune_result* out);
int gt_autotune_batch(gt_scene* s, gt_batch* b, int use_graph, uint32_t iters, void* hip_stream,
                      gt_autotune_result* out);

/* ---- synthetic scenes (seeded SplitMix64 + Box-Muller) ---- */
typedef struct gt_gen_params {
  uint32_t n;
  uint32_t n_clusters;      /* 0: means uniform in the box; else mixture */
  float box_half;           /* means / cluster centres uniform in [-box_half, box_half]^3 */
  float cluster_sigma;      /* member offsets N(0, sigma) per axis */
  float log_scale_mean, log_scale_std;
  float logit_mean, logit_std;
  float dc_std, rest_std;
  uint64_t seed;
} gt_gen_params;
int gt_generate_scene(const gt_gen_params* g, float* raw /* n x 59 */);

/* Reads a binary-little-endian 3DGS .ply (float/uchar properties, SH degree
 * 0..3) into PLY-raw records (src/GS_ply.ts:58-132, 171-196).  Missing SH
 * coefficients are zero.  *raw is malloc'ed; release with gt_free. */
int gt_ply_read(const char* path, float** raw, uint32_t* n, uint32_t* sh_degree);
void gt_free(void* p);

/* ---- presentation: the screen pass and the Save button ---- */
#define GT_TONEMAP_IDENTITY 0       /* what the reference's fragment shader returns: hdrColor (toneMapper.ts:22) */
#define GT_TONEMAP_EXPOSURE 1       /* its unused ldrColor = 1 - exp(-1.66 hdr) (toneMapper.ts:19) */
#define GT_TONEMAP_EXPOSURE_GAMMA 2 /* with the commented-out pow(ldr, 1/2.2) (toneMapper.ts:21) */
/* A W x H float RGBA frame (gt_render's layout, pixel x + y W) as the
 * reference's canvas shows it (src/toneMapper.ts:11-26): canvas pixel
 * (x, y) takes pixel x + y W; RGB through the curve of `mode`; alpha 1; each
 * channel clamped to [0, 1] and rounded to 255 steps (round half up, NaN to
 * 0) as 8-bit RGBA.  buffers: GT_BUF_HOST (synchronous) or GT_BUF_DEVICE
 * (asynchronous on hip_stream); rgba and out are both of that kind.  The
 * curves are evaluated in FP64 with this library's exp/log, so host and
 * device results are identical. */
int gt_tonemap(const float* rgba, uint32_t width, uint32_t height, uint32_t mode, uint8_t* out, int buffers,
               void* hip_stream);
/* Writes 8-bit RGBA (rows top to bottom) as a PNG file: colour type 6,
 * depth 8, filter 0, stored deflate blocks.  The reference's Save button
 * downloads the canvas this way (src/main.ts:606-611, toDataURL('image/png')).
 * Width and height 1..16384. */
int gt_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height);

const char* gt_last_error(void);
const char* gt_version(void);

#ifdef __cplusplus
}
#endif
#endif /* GT_H */
