/*
 * gt_oracle.h - CPU restatement of the reference's Gaussian ray-tracing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the
 * checker (or the timed CPU baseline), never as the product path.
 *
 * Restated reference files:
 *   activations          src/GS_ply.ts:225-235
 *   320-B record         src/GS_ply.ts:198-205, src/GS_packing.ts:146-231
 *   proxy filter/xform   src/main.ts:73-94 (wgpu-matrix mat3.fromQuat / mat4.mul),
 *                        from the JS doubles or from a packed buffer's f32 values
 *   world-to-object      webrtx/bvh/src/lib.rs:501-505 (glam Affine3A::inverse)
 *   ray generation       src/GS_shader.ts:275-301
 *   round loop/composite src/GS_shader.ts:192-273
 *   any-hit k-buffer     src/GS_shader.ts:323-368 + commit rule trace.glsl:206-215
 *   traversal semantics  webrtx/src/glsl/trace.glsl:42-245
 *   triangle test        webrtx/src/glsl/intersect.glsl:73-92
 *   SH radiance          src/GS_shader.ts:33-56, 138-190
 *   debug modes          src/GS_shader.ts:201-228, 289-296, 312-321
 */
#ifndef GT_ORACLE_H
#define GT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GTO_RAW_FLOATS 59     /* x y z f_dc[3] f_rest[45] opacity scale[3] rot[4] */
#define GTO_PACKED_FLOATS 80  /* 320-B GaussianParticle record */

typedef struct gto_scene {
  uint32_t n;                /* Gaussians */
  const float* packed;       /* n x 80 floats */
  uint32_t n_proxies;        /* Gaussians with opacity > alpha_min */
  const uint32_t* proxy_ids; /* original (PLY) index per proxy */
  const float* o2w;          /* n_proxies x 12, column-major 4x3 */
  const float* w2o;          /* n_proxies x 12, column-major 4x3 */
  const void* bvh;           /* optional prebuilt gto_bvh_build result */
} gto_scene;

/* Mirrors the 112-B uniform block (src/GS_shader.ts:96-108) + launch size. */
typedef struct gto_params {
  float to_world[16];
  float aspect[2];
  float t_min, t_max, T_min, alpha_min;
  uint32_t k, sh_degree_max, accumulated_frames, early_stop, debug_mode;
  uint32_t width, height;
  uint32_t antialias; /* GS_shader.ts's antialiasing path: jittered rays, running average (rgba in/out) */
} gto_params;

/* stats[]: 0 rays, 1 rounds, 2 H_eval, 3 H_comp, 4 proxy candidates tested,
 *          5 front-face hits found, 6 exact-t ties observed, 7 hit-list overflows */
#define GTO_NSTATS 8

/* GS_ply.ts:225-235: raw -> packed (f32) and the JS-double activated values
 * used by the proxy builder (opacity, scale[3], quat[4] per Gaussian). */
void gto_activate(const float* raw, uint32_t n, float* packed, double* act8);

/* main.ts:73-94 + lib.rs:501-505.  Returns the number of proxies. */
uint32_t gto_build_proxies(const float* raw, const double* act8, uint32_t n, double alpha_min,
                           uint32_t* proxy_ids, float* o2w, float* w2o);

/* The same from packed 320-B records (GS_ply.ts:198-205): act8 from the
 * f32 fields (gto_activate_packed), positions from the record. */
void gto_activate_packed(const float* packed, uint32_t n, double* act8);
uint32_t gto_build_proxies_packed(const float* packed, const double* act8, uint32_t n, double alpha_min,
                                  uint32_t* proxy_ids, float* o2w, float* w2o);

/* Render.  ray_list (optional) selects rays by linear index y*W+x; outputs
 * With p->antialias and accumulated_frames > 0, rgba holds the running
 * average on input and receives the new average (GS_shader.ts:303-307).
 * are indexed by position in the list.  hit_ids: per ray max_rounds*k ids
 * (sentinel-filled), accept: per ray max_rounds u32 masks (bit i = entry i
 * composited), n_rounds: rounds run.  Any output may be NULL.
 * use_bvh=0 is brute force over all proxies.  Returns 0 or a negative error. */
int gto_render(const gto_scene* s, const gto_params* p, int use_bvh, int n_threads,
               const uint32_t* ray_list, uint32_t n_list,
               float* rgba, float* T_out, uint32_t* n_rounds, uint32_t max_rounds,
               uint32_t* hit_ids, uint32_t* accept, uint64_t* stats);

/* Prebuilt oracle BVH over the scene's proxies (use_bvh renders reuse it
 * when gto_scene.bvh is set). */
void* gto_bvh_build(const gto_scene* s);
void gto_bvh_free(void* b);

/* FP64 re-evaluation of given hit lists and accept masks: pixels in double
 * (of the frame itself: the antialiasing average is not applied). */
void gto_reeval_f64(const gto_scene* s, const gto_params* p, const uint32_t* ray_list, uint32_t n_list,
                    const uint32_t* n_rounds, uint32_t max_rounds, const uint32_t* hit_ids,
                    const uint32_t* accept, double* rgb_out, double* T_out);

/* Deliberate-mistake switch used by tests to show the pinning checks fail
 * when the oracle is wrong.  0 = faithful. */
void gto_set_mutation(int m);

/* Pieces exposed for pinning tests. */
void gto_ray_dir(const gto_params* p, uint32_t px, uint32_t py, float* d3);
double gto_exp64(double x);
double gto_log64(double x);
double gto_response(const float* packed_rec, const float* o3, const float* d3);
void gto_radiance(const float* packed_rec, uint32_t sh_degree, const float* dir3, float* rgb);
int gto_tri_hit(const float* o3, const float* d3, float tmin, float tmax, int face, float* t);

#ifdef __cplusplus
}
#endif
#endif
