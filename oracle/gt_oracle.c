/*
 * gt_oracle.c - CPU restatement of the reference's Gaussian ray-tracing path.
 *
 * TEST INFRASTRUCTURE: the checker for the HIP path and the CPU baseline of
 * bench.py.  The product never links or calls it.
 *
 * Semantics restated (reference file:line in each function):
 *  - per ray, rounds of "trace + composite" (src/GS_shader.ts:192-273);
 *  - a round keeps the k smallest-t front-face proxy hits with
 *    t_curr < t < t_max (any-hit insertion GS_shader.ts:349-353 with the
 *    commit rule GS_shader.ts:358-366 / trace.glsl:206-215);
 *  - exact-t ties between different Gaussians depend on the reference's BVH
 *    traversal order, which is not reproducible; they are broken here by
 *    (t, Gaussian id), equal (t, id) pairs keep arrival order;
 *  - proxy hits use the exact FP32 triangle test of intersect.glsl:73-92
 *    in object space (trace.glsl:102-105) and the facing rule of
 *    trace.glsl:153-164 (shared primitives in include/gt_fp32.h).
 * The BVH variant culls with conservative double-precision boxes, so it
 * finds exactly the hits of the brute-force variant.
 */
#include "gt_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "../include/gt_fp32.h"

static int g_mutation = 0;
void gto_set_mutation(int m) { g_mutation = m; }

/* ---------------- activations: src/GS_ply.ts:187-196, 225-235 ---------------- */
/* exp and log are gt_exp64 / gt_log64 (the library's definitions of the JS
 * engine's Math.exp / Math.log, which are not correctly rounded either). */
void gto_activate(const float* raw, uint32_t n, float* packed, double* act8) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = raw + (size_t)i * GTO_RAW_FLOATS;
    float* o = packed + (size_t)i * GTO_PACKED_FLOATS;
    double* a = act8 + (size_t)i * 8;
    memset(o, 0, GTO_PACKED_FLOATS * sizeof(float));
    o[0] = r[0]; o[1] = r[1]; o[2] = r[2];
    double sc[3], q[4];
    for (int j = 0; j < 3; ++j) sc[j] = gt_exp64((double)r[52 + j]);
    double q0 = r[55], q1 = r[56], q2 = r[57], q3 = r[58];
    double norm = sqrt(((q0 * q0 + q1 * q1) + q2 * q2) + q3 * q3);
    q[0] = q0 / norm; q[1] = q1 / norm; q[2] = q2 / norm; q[3] = q3 / norm;
    double op = 1.0 / (1.0 + gt_exp64(-(double)r[51]));
    for (int j = 0; j < 3; ++j) o[4 + j] = (float)sc[j];
    for (int j = 0; j < 4; ++j) o[8 + j] = (float)q[j];
    o[12] = (float)op;
    /* sh[0] = f_dc_{0,1,2}; sh[1+j] = (f_rest_j, f_rest_{15+j}, f_rest_{30+j}) */
    for (int c = 0; c < 3; ++c) o[16 + c] = r[3 + c];
    for (int j = 0; j < 15; ++j)
      for (int c = 0; c < 3; ++c) o[16 + 4 * (1 + j) + c] = r[6 + c * 15 + j];
    a[0] = op;
    a[1] = sc[0]; a[2] = sc[1]; a[3] = sc[2];
    a[4] = q[0]; a[5] = q[1]; a[6] = q[2]; a[7] = q[3];
  }
}

/* ---------------- proxies: src/main.ts:73-94, webrtx/bvh/src/lib.rs:501-505 ---------------- */
/* wgpu-matrix mat3.fromQuat (gl-matrix formulation), computed in double and
 * stored to f32 (Float32Array), column-major with 3 used rows. */
static void from_quat_f32(double x, double y, double z, double w, float m[3][3]) {
  double x2 = x + x, y2 = y + y, z2 = z + z;
  double xx = x * x2, yx = y * x2, yy = y * y2, zx = z * x2, zy = z * y2, zz = z * z2;
  double wx = w * x2, wy = w * y2, wz = w * z2;
  m[0][0] = (float)(1 - yy - zz); m[0][1] = (float)(yx + wz); m[0][2] = (float)(zx - wy);
  m[1][0] = (float)(yx - wz); m[1][1] = (float)(1 - xx - zz); m[1][2] = (float)(zy + wx);
  m[2][0] = (float)(zx + wy); m[2][1] = (float)(zy - wx); m[2][2] = (float)(1 - xx - yy);
}

/* glam Affine3A::inverse in f32 (scalar path): rows of the 3x3 inverse are
 * cross products of the columns times 1/det; translation = -(inv * t). */
static void affine_inverse_f32(const float* m, float* out) {
  gt_f3 x = gt_f3_make(m[0], m[1], m[2]);
  gt_f3 y = gt_f3_make(m[3], m[4], m[5]);
  gt_f3 z = gt_f3_make(m[6], m[7], m[8]);
  gt_f3 t = gt_f3_make(m[9], m[10], m[11]);
  gt_f3 tmp0 = gt_cross(y, z), tmp1 = gt_cross(z, x), tmp2 = gt_cross(x, y);
  float det = gt_dot(z, tmp2);
  float inv_det = 1.0f / det;
  gt_f3 r0 = gt_scale(inv_det, tmp0), r1 = gt_scale(inv_det, tmp1), r2 = gt_scale(inv_det, tmp2);
  /* inverse column c = (r0[c], r1[c], r2[c]) */
  gt_f3 c0 = gt_f3_make(r0.x, r1.x, r2.x), c1 = gt_f3_make(r0.y, r1.y, r2.y), c2 = gt_f3_make(r0.z, r1.z, r2.z);
  gt_f3 it = gt_add(gt_add(gt_scale(t.x, c0), gt_scale(t.y, c1)), gt_scale(t.z, c2));
  out[0] = c0.x; out[1] = c0.y; out[2] = c0.z;
  out[3] = c1.x; out[4] = c1.y; out[5] = c1.z;
  out[6] = c2.x; out[7] = c2.y; out[8] = c2.z;
  out[9] = -it.x; out[10] = -it.y; out[11] = -it.z;
}

/* The proxies of main.ts:79-94 from activated values act8 and positions at
 * pos + i * pos_stride (the raw record or the packed record). */
static uint32_t build_proxies_at(const float* pos, size_t pos_stride, const double* act8, uint32_t n,
                                 double alpha_min, uint32_t* proxy_ids, float* o2w, float* w2o) {
  uint32_t np = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const double* a = act8 + (size_t)i * 8;
    const float* r = pos + (size_t)i * pos_stride;
    double op = a[0];
    if (!(op > alpha_min)) continue;
    double s = sqrt(2 * gt_log64(op / alpha_min));
    if (g_mutation == 7) s = 1.0;
    float S[3];
    for (int j = 0; j < 3; ++j) S[j] = (float)(s * a[1 + j]);
    float R[3][3];
    from_quat_f32(a[5], a[6], a[7], a[4], R);
    float* m = o2w + (size_t)np * 12;
    for (int c = 0; c < 3; ++c)
      for (int rr = 0; rr < 3; ++rr) m[c * 3 + rr] = (float)((double)R[c][rr] * (double)S[c]);
    m[9] = r[0]; m[10] = r[1]; m[11] = r[2];
    affine_inverse_f32(m, w2o + (size_t)np * 12);
    proxy_ids[np++] = i;
  }
  return np;
}

uint32_t gto_build_proxies(const float* raw, const double* act8, uint32_t n, double alpha_min,
                           uint32_t* proxy_ids, float* o2w, float* w2o) {
  return build_proxies_at(raw, GTO_RAW_FLOATS, act8, n, alpha_min, proxy_ids, o2w, w2o);
}

/* A packed 320-B buffer (GS_ply.ts:198-205) holds the activated values only
 * as f32: gaussianData's doubles (GS_ply.ts:225-235, read by main.ts:73-78)
 * are not part of it, so the proxy builder's inputs are those f32 values
 * widened to double. */
void gto_activate_packed(const float* packed, uint32_t n, double* act8) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* o = packed + (size_t)i * GTO_PACKED_FLOATS;
    double* a = act8 + (size_t)i * 8;
    a[0] = o[12];
    for (int j = 0; j < 3; ++j) a[1 + j] = o[4 + j];
    for (int j = 0; j < 4; ++j) a[4 + j] = o[8 + j];
  }
}

uint32_t gto_build_proxies_packed(const float* packed, const double* act8, uint32_t n, double alpha_min,
                                  uint32_t* proxy_ids, float* o2w, float* w2o) {
  return build_proxies_at(packed, GTO_PACKED_FLOATS, act8, n, alpha_min, proxy_ids, o2w, w2o);
}

/* ---------------- shading pieces ---------------- */
double gto_exp64(double x) { return gt_exp64(x); }
double gto_log64(double x) { return gt_log64(x); }

/* pixel offset (GS_shader.ts:281-283): 0.5, or the getNewSeed/rng jitter with antialiasing */
static void pixel_offset(const gto_params* p, uint32_t px, uint32_t py, float* ox, float* oy) {
  *ox = 0.5f;
  *oy = 0.5f;
  if (p->antialias) gt_aa_offset(px, py, p->width, p->accumulated_frames, ox, oy);
}

void gto_ray_dir(const gto_params* p, uint32_t px, uint32_t py, float* d3) {
  gt_f3 d;
  float ox, oy;
  pixel_offset(p, px, py, &ox, &oy);
  if (g_mutation == 3) {
    float tw[16];
    memcpy(tw, p->to_world, sizeof(tw));
    tw[0] = -tw[0]; tw[1] = -tw[1]; tw[2] = -tw[2];
    d = gt_ray_dir_off(tw, p->aspect[0], p->aspect[1], px, py, p->width, p->height, ox, oy);
  } else {
    d = gt_ray_dir_off(p->to_world, p->aspect[0], p->aspect[1], px, py, p->width, p->height, ox, oy);
  }
  d3[0] = d.x; d3[1] = d.y; d3[2] = d.z;
}

/* computeResponse (GS_shader.ts:111-136) through gt_response64 */
static double response_of(const float* rec, gt_f3 o, gt_f3 d) {
  double r = gt_response64(rec, rec + 4, rec + 8, o, d);
  if (g_mutation == 1) r = sqrt(r); /* exp(-e/2): a 1/2 factor in the exponent */
  return r;
}

double gto_response(const float* rec, const float* o3, const float* d3) {
  return response_of(rec, gt_f3_make(o3[0], o3[1], o3[2]), gt_f3_make(d3[0], d3[1], d3[2]));
}

/* SH constants, src/GS_shader.ts:39-52 (f32) */
#define SH_C0 0.28209479177387814f
#define SH_C1 0.4886025119029199f
#define SH_C2_0 1.0925484305920792f
#define SH_C2_1 -1.0925484305920792f
#define SH_C2_2 0.31539156525252005f
#define SH_C2_3 -1.0925484305920792f
#define SH_C2_4 0.5462742152960396f
#define SH_C3_0 -0.5900435899266435f
#define SH_C3_1 2.890611442640554f
#define SH_C3_2 -0.4570457994644658f
#define SH_C3_3 0.3731763325901154f
#define SH_C3_4 -0.4570457994644658f
#define SH_C3_5 1.445305721320277f
#define SH_C3_6 -0.5900435899266435f

/* shToRadiance + computeRadiance (GS_shader.ts:138-190), channel-wise. */
void gto_radiance(const float* rec, uint32_t deg, const float* dir, float* rgb) {
  const float* sh = rec + 16;
  float x = dir[0], y = dir[1], z = dir[2];
  float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  for (int c = 0; c < 3; ++c) {
#define SH(i) sh[4 * (i) + c]
    float L = 0.5f + SH_C0 * SH(0);
    if (deg >= 1) {
      float ym = (g_mutation == 2) ? y : -y;
      L += SH_C1 * ((ym * SH(1) + z * SH(2)) - x * SH(3));
    }
    if (deg >= 2) {
      L += (((SH_C2_0 * xy * SH(4) + SH_C2_1 * yz * SH(5)) + SH_C2_2 * ((2.f * zz - xx) - yy) * SH(6)) +
            SH_C2_3 * xz * SH(7)) +
           SH_C2_4 * (xx - yy) * SH(8);
    }
    if (deg >= 3) {
      L += (((((SH_C3_0 * y * (3.0f * xx - yy) * SH(9) + SH_C3_1 * xy * z * SH(10)) +
               SH_C3_2 * y * ((4.0f * zz - xx) - yy) * SH(11)) +
              SH_C3_3 * z * ((2.0f * zz - 3.0f * xx) - 3.0f * yy) * SH(12)) +
             SH_C3_4 * x * ((4.0f * zz - xx) - yy) * SH(13)) +
            SH_C3_5 * z * (xx - yy) * SH(14)) +
           SH_C3_6 * x * (xx - 3.0f * yy) * SH(15);
    }
#undef SH
    rgb[c] = L > 0.0f ? L : 0.0f; /* max(vec3(0), L) */
  }
}

int gto_tri_hit(const float* o3, const float* d3, float tmin, float tmax, int face, float* t) {
  return gt_tri_hit(gt_f3_make(o3[0], o3[1], o3[2]), gt_f3_make(d3[0], d3[1], d3[2]), tmin, tmax,
                    gt_ico_face(face), t);
}

/* ---------------- oracle BVH (conservative double boxes) ---------------- */
typedef struct obvh_node {
  double lo[3], hi[3];
  int32_t left, right; /* children; leaf when left < 0: proxies [first, first+count) */
  int32_t first, count;
} obvh_node;

typedef struct obvh {
  obvh_node* nodes;
  uint32_t* order; /* proxy indices */
  int32_t n_nodes;
} obvh;

typedef struct box_ctx {
  const double* boxes; /* n_proxies x 6 */
  const double* cent;
  int axis;
} box_ctx;

static box_ctx g_sort_ctx;
static int cmp_cent(const void* a, const void* b) {
  double ca = g_sort_ctx.cent[3 * (*(const uint32_t*)a) + g_sort_ctx.axis];
  double cb = g_sort_ctx.cent[3 * (*(const uint32_t*)b) + g_sort_ctx.axis];
  if (ca < cb) return -1;
  if (ca > cb) return 1;
  return (*(const uint32_t*)a < *(const uint32_t*)b) ? -1 : 1;
}

static int32_t obvh_build_rec(obvh* b, const double* boxes, const double* cent, uint32_t first, uint32_t count) {
  int32_t id = b->n_nodes++;
  obvh_node* nd = &b->nodes[id];
  for (int j = 0; j < 3; ++j) { nd->lo[j] = 1e300; nd->hi[j] = -1e300; }
  double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
  for (uint32_t i = first; i < first + count; ++i) {
    const double* bx = boxes + 6 * b->order[i];
    const double* c = cent + 3 * b->order[i];
    for (int j = 0; j < 3; ++j) {
      if (bx[j] < nd->lo[j]) nd->lo[j] = bx[j];
      if (bx[3 + j] > nd->hi[j]) nd->hi[j] = bx[3 + j];
      if (c[j] < clo[j]) clo[j] = c[j];
      if (c[j] > chi[j]) chi[j] = c[j];
    }
  }
  if (count <= 4) {
    nd->left = -1; nd->right = -1; nd->first = (int32_t)first; nd->count = (int32_t)count;
    return id;
  }
  int axis = 0;
  for (int j = 1; j < 3; ++j)
    if (chi[j] - clo[j] > chi[axis] - clo[axis]) axis = j;
  g_sort_ctx.boxes = boxes; g_sort_ctx.cent = cent; g_sort_ctx.axis = axis;
  qsort(b->order + first, count, sizeof(uint32_t), cmp_cent);
  uint32_t half = count / 2;
  int32_t l = obvh_build_rec(b, boxes, cent, first, half);
  int32_t r = obvh_build_rec(b, boxes, cent, first + half, count - half);
  nd = &b->nodes[id];
  nd->left = l; nd->right = r; nd->first = -1; nd->count = 0;
  return id;
}

/* World box of proxy i: O2W applied to the 12 icosahedron vertices in
 * double, padded outward far beyond the FP32 error of the object-space test. */
static void proxy_box(const float* m, double* box) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int v = 0; v < 12; ++v) {
    gt_f3 p = gt_ico_vertex(v);
    for (int r = 0; r < 3; ++r) {
      double w = (double)m[r] * p.x + (double)m[3 + r] * p.y + (double)m[6 + r] * p.z + (double)m[9 + r];
      if (w < lo[r]) lo[r] = w;
      if (w > hi[r]) hi[r] = w;
    }
  }
  for (int r = 0; r < 3; ++r) {
    double pad = 1e-3 * (hi[r] - lo[r]) + 1e-5 * (1.0 + fabs(lo[r]) + fabs(hi[r]));
    box[r] = lo[r] - pad;
    box[3 + r] = hi[r] + pad;
  }
}

static void obvh_build(obvh* b, const gto_scene* s) {
  uint32_t n = s->n_proxies;
  double* boxes = (double*)malloc(sizeof(double) * 6 * (n ? n : 1));
  double* cent = (double*)malloc(sizeof(double) * 3 * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) {
    proxy_box(s->o2w + (size_t)12 * i, boxes + 6 * i);
    for (int j = 0; j < 3; ++j) cent[3 * i + j] = 0.5 * (boxes[6 * i + j] + boxes[6 * i + 3 + j]);
  }
  b->order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t i = 0; i < n; ++i) b->order[i] = i;
  b->nodes = (obvh_node*)malloc(sizeof(obvh_node) * (2 * (size_t)n + 2));
  b->n_nodes = 0;
  if (n) obvh_build_rec(b, boxes, cent, 0, n);
  free(boxes);
  free(cent);
}

static void obvh_free(obvh* b) { free(b->nodes); free(b->order); }

void* gto_bvh_build(const gto_scene* s) {
  obvh* b = (obvh*)malloc(sizeof(obvh));
  obvh_build(b, s);
  return b;
}

void gto_bvh_free(void* b) {
  if (!b) return;
  obvh_free((obvh*)b);
  free(b);
}

/* slab interval of a double box; returns 1 if [t0, t1] overlaps [lo, hi] */
static int box_overlap(const obvh_node* nd, const double* o, const double* inv, double lo, double hi) {
  double t0 = lo, t1 = hi;
  for (int j = 0; j < 3; ++j) {
    double a = (nd->lo[j] - o[j]) * inv[j];
    double c = (nd->hi[j] - o[j]) * inv[j];
    if (a != a) a = -1e300; /* 0 * inf: ray in the slab plane, no constraint */
    if (c != c) c = 1e300;
    double mn = a < c ? a : c, mx = a < c ? c : a;
    if (mn > t0) t0 = mn;
    if (mx < t1) t1 = mx;
  }
  return t0 <= t1;
}

/* ---------------- k-buffer (canonical key: t bits, then id) ---------------- */
typedef struct kentry { float t; uint32_t id; } kentry;

static int key_less(float ta, uint32_t ia, float tb, uint32_t ib) {
  uint32_t ba = gt_f32_to_bits(ta), bb = gt_f32_to_bits(tb);
  if (ba != bb) return ba < bb;
  return ia < ib;
}

typedef struct ray_ctx {
  gt_f3 o, d;
  float tmin, tmax;
  uint32_t k;
  kentry kc[33];
  int full;
  uint64_t cand, hits, ties;
  int backfaces;    /* debug mode 1: no culling */
  int count_only;   /* debug mode 2 */
  uint64_t count;
} ray_ctx;

static void kbuf_insert(ray_ctx* r, float t, uint32_t id) {
  uint32_t k = r->k;
  if (!key_less(t, id, r->kc[k - 1].t, r->kc[k - 1].id)) {
    if (gt_f32_to_bits(t) == gt_f32_to_bits(r->kc[k - 1].t) && id != r->kc[k - 1].id) r->ties++;
    return;
  }
  /* stable insertion: after every entry with key <= new key */
  uint32_t pos = k - 1;
  while (pos > 0 && key_less(t, id, r->kc[pos - 1].t, r->kc[pos - 1].id)) {
    r->kc[pos] = r->kc[pos - 1];
    --pos;
  }
  if (pos > 0 && gt_f32_to_bits(r->kc[pos - 1].t) == gt_f32_to_bits(t) && r->kc[pos - 1].id != id) r->ties++;
  r->kc[pos].t = t;
  r->kc[pos].id = id;
}

static void test_proxy(ray_ctx* r, const gto_scene* s, uint32_t pi) {
  const float* w2o = s->w2o + (size_t)12 * pi;
  gt_f3 oo = gt_xform_point(w2o, r->o);
  gt_f3 od = gt_xform_dir(w2o, r->d);
  r->cand++;
  for (int f = 0; f < GT_NUM_FACES; ++f) {
    gt_face fc = gt_ico_face(f);
    float t;
    if (!gt_tri_hit(oo, od, r->tmin, r->tmax, fc, &t)) continue;
    int front = gt_front_facing(w2o, fc.n, r->d);
    if (g_mutation == 4) front = !front;
    if (g_mutation == 11 && r->count_only) front = 1; /* debugMode 2 counting back faces too */
    if (!r->backfaces && !front) continue;
    r->hits++;
    if (r->count_only) { r->count++; continue; }
    kbuf_insert(r, t, s->proxy_ids[pi]);
  }
}

static void gather(ray_ctx* r, const gto_scene* s, const obvh* b) {
  if (!b) {
    for (uint32_t i = 0; i < s->n_proxies; ++i) test_proxy(r, s, i);
    return;
  }
  if (b->n_nodes == 0) return;
  double o[3] = {r->o.x, r->o.y, r->o.z};
  double inv[3] = {1.0 / (double)r->d.x, 1.0 / (double)r->d.y, 1.0 / (double)r->d.z};
  int32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const obvh_node* nd = &b->nodes[stack[--sp]];
    double hi = (double)r->kc[r->k - 1].t;
    if (r->count_only) hi = r->tmax;
    hi = hi * (1.0 + 1e-5) + 1e-5;
    double lo = (double)r->tmin * (1.0 - 1e-5) - 1e-5;
    if (!box_overlap(nd, o, inv, lo, hi)) continue;
    if (nd->left < 0) {
      for (int32_t i = nd->first; i < nd->first + nd->count; ++i) test_proxy(r, s, b->order[i]);
    } else {
      stack[sp++] = nd->right;
      stack[sp++] = nd->left;
    }
  }
}

/* ---------------- render ---------------- */
typedef struct render_job {
  const gto_scene* s;
  const gto_params* p;
  const obvh* b;
  const uint32_t* ray_list;
  uint32_t n_list;
  float* rgba;
  float* T_out;
  uint32_t* n_rounds;
  uint32_t max_rounds;
  uint32_t* hit_ids;
  uint32_t* accept;
  volatile uint32_t next;
  uint64_t stats[GTO_NSTATS];
  pthread_mutex_t mu;
} render_job;

static void render_ray(render_job* J, uint32_t li, uint64_t* st) {
  const gto_params* p = J->p;
  const gto_scene* s = J->s;
  uint32_t ray = J->ray_list ? J->ray_list[li] : li;
  uint32_t px = ray % p->width, py = ray / p->width;
  float d3[3];
  gto_ray_dir(p, px, py, d3);
  ray_ctx r;
  memset(&r, 0, sizeof(r));
  r.o = gt_f3_make(p->to_world[12], p->to_world[13], p->to_world[14]);
  r.d = gt_f3_make(d3[0], d3[1], d3[2]);
  r.k = p->k;
  r.tmax = p->t_max;
  uint32_t MR = J->max_rounds;
  uint32_t* ids_out = J->hit_ids ? J->hit_ids + (size_t)li * MR * p->k : NULL;
  uint32_t* acc_out = J->accept ? J->accept + (size_t)li * MR : NULL;
  if (ids_out)
    for (size_t i = 0; i < (size_t)MR * p->k; ++i) ids_out[i] = GT_SENTINEL;
  if (acc_out)
    for (uint32_t i = 0; i < MR; ++i) acc_out[i] = 0;
  float L[3] = {0, 0, 0};
  float T = 1.0f;
  uint32_t rounds = 0;

  if (p->debug_mode == 1) {
    /* GS_shader.ts:289-296 + 312-321: one trace, no culling, every hit
     * commits (zero payload), closest hit colours by SH at the ray dir. */
    r.k = 1;
    r.kc[0].t = p->t_max; r.kc[0].id = GT_SENTINEL;
    r.tmin = p->t_min;
    r.backfaces = g_mutation == 8 ? 0 : 1; /* mutation 8: culling back faces as mode 0 does */
    gather(&r, s, J->b);
    if (r.kc[0].id != GT_SENTINEL) {
      const float* rec = s->packed + (size_t)GTO_PACKED_FLOATS * r.kc[0].id;
      if (g_mutation == 9) { /* the commented-out alternative, GS_shader.ts:320 */
        gt_f3 dir = gt_normalize(gt_sub(gt_f3_make(rec[0], rec[1], rec[2]), r.o));
        float dd[3] = {dir.x, dir.y, dir.z};
        gto_radiance(rec, p->sh_degree_max, dd, L);
      } else {
        gto_radiance(rec, p->sh_degree_max, d3, L);
      }
    }
    rounds = 1;
    if (ids_out) ids_out[0] = r.kc[0].id;
  } else if (p->debug_mode == 2) {
    /* GS_shader.ts:201-228: any-hit call count binned into colours; counted
     * here as all front-face hits in (t_min, t_max), independent of order. */
    r.k = 1;
    r.kc[0].t = p->t_max; r.kc[0].id = GT_SENTINEL;
    r.tmin = p->t_min;
    r.count_only = 1;
    gather(&r, s, J->b);
    uint64_t c = r.count;
    float col[3];
    if (c == 0) { col[0] = 0; col[1] = 0; col[2] = 0; }
    else if (c < 20 || (g_mutation == 10 && c == 20)) { col[0] = 0.5f; col[1] = 0.5f; col[2] = 0; }
    else if (c < 40) { col[0] = 0.5f; col[1] = 0; col[2] = 0.5f; }
    else if (c < 60) { col[0] = 0; col[1] = 0.5f; col[2] = 0.5f; }
    else if (c < 80) { col[0] = 0.5f; col[1] = 0; col[2] = 0; }
    else if (c < 100) { col[0] = 0; col[1] = 0.5f; col[2] = 0; }
    else if (c < 150) { col[0] = 0; col[1] = 0; col[2] = 0.5f; }
    else { col[0] = 0.5f; col[1] = 0.5f; col[2] = 0.5f; }
    L[0] = col[0]; L[1] = col[1]; L[2] = col[2];
    rounds = 1;
    if (ids_out) ids_out[0] = (uint32_t)(c > 0xfffffffeu ? 0xfffffffeu : c);
  } else {
    float t_curr = p->t_min;
    uint32_t step = 0;
    while (p->T_min < T && t_curr < p->t_max) {
      if (p->early_stop == step++) break;
      for (uint32_t i = 0; i < p->k; ++i) { r.kc[i].t = p->t_max; r.kc[i].id = GT_SENTINEL; }
      r.tmin = t_curr;
      gather(&r, s, J->b);
      t_curr = r.kc[p->k - 1].t + (g_mutation == 5 ? 0.0f : GT_RELAUNCH_EPS);
      uint32_t mask = 0;
      for (uint32_t i = 0; i < p->k; ++i) {
        uint32_t id = r.kc[i].id;
        if (id == GT_SENTINEL) { t_curr = p->t_max; break; }
        st[2]++;
        const float* rec = s->packed + (size_t)GTO_PACKED_FLOATS * id;
        float alpha = (float)(response_of(rec, r.o, r.d) * (double)rec[12]);
        if (p->alpha_min < alpha) {
          gt_f3 mu = gt_f3_make(rec[0], rec[1], rec[2]);
          gt_f3 dir = gt_normalize(gt_sub(mu, r.o));
          float dd[3] = {dir.x, dir.y, dir.z}, Lh[3];
          gto_radiance(rec, p->sh_degree_max, dd, Lh);
          float Ta = (g_mutation == 6) ? alpha : T * alpha;
          for (int c = 0; c < 3; ++c) L[c] += Ta * Lh[c];
          T *= 1.0f - alpha;
          mask |= 1u << i;
          st[3]++;
        }
      }
      if (rounds < MR) {
        if (ids_out)
          for (uint32_t i = 0; i < p->k; ++i) ids_out[(size_t)rounds * p->k + i] = r.kc[i].id;
        if (acc_out) acc_out[rounds] = mask;
      } else {
        st[7]++;
      }
      rounds++;
    }
  }
  if (J->rgba) {
    float* o = J->rgba + 4 * (size_t)li;
    if (p->antialias) /* GS_shader.ts:303-307: running average over accumulated frames */
      for (int c = 0; c < 3; ++c) L[c] = gt_aa_mix(o[c], L[c], p->accumulated_frames);
    o[0] = L[0]; o[1] = L[1]; o[2] = L[2]; o[3] = 1.0f;
  }
  if (J->T_out) J->T_out[li] = T;
  if (J->n_rounds) J->n_rounds[li] = rounds;
  st[0]++;
  st[1] += rounds;
  st[4] += r.cand;
  st[5] += r.hits;
  st[6] += r.ties;
}

static void* render_worker(void* arg) {
  render_job* J = (render_job*)arg;
  uint64_t st[GTO_NSTATS] = {0};
  for (;;) {
    uint32_t b = __atomic_fetch_add(&J->next, 16u, __ATOMIC_RELAXED);
    if (b >= J->n_list) break;
    uint32_t e = b + 16 < J->n_list ? b + 16 : J->n_list;
    for (uint32_t i = b; i < e; ++i) render_ray(J, i, st);
  }
  pthread_mutex_lock(&J->mu);
  for (int i = 0; i < GTO_NSTATS; ++i) J->stats[i] += st[i];
  pthread_mutex_unlock(&J->mu);
  return NULL;
}

int gto_render(const gto_scene* s, const gto_params* p, int use_bvh, int n_threads,
               const uint32_t* ray_list, uint32_t n_list,
               float* rgba, float* T_out, uint32_t* n_rounds, uint32_t max_rounds,
               uint32_t* hit_ids, uint32_t* accept, uint64_t* stats) {
  if (!s || !p) return -1;
  if (p->k < 1 || p->k > 32) return -2;
  if (p->width == 0 || p->height == 0) return -3;
  if (!(p->t_min >= 0.0f)) return -4;
  obvh bv;
  const obvh* bp = NULL;
  int own = 0;
  if (use_bvh && s->bvh) {
    bp = (const obvh*)s->bvh;
  } else if (use_bvh) {
    obvh_build(&bv, s);
    bp = &bv;
    own = 1;
  }
  render_job J;
  memset(&J, 0, sizeof(J));
  J.s = s; J.p = p; J.b = bp;
  J.ray_list = ray_list;
  J.n_list = ray_list ? n_list : p->width * p->height;
  J.rgba = rgba; J.T_out = T_out; J.n_rounds = n_rounds;
  J.max_rounds = max_rounds ? max_rounds : 1;
  J.hit_ids = max_rounds ? hit_ids : NULL;
  J.accept = max_rounds ? accept : NULL;
  J.next = 0;
  pthread_mutex_init(&J.mu, NULL);
  if (n_threads < 1) n_threads = 1;
  if (n_threads > 256) n_threads = 256;
  pthread_t th[256];
  for (int i = 1; i < n_threads; ++i) pthread_create(&th[i], NULL, render_worker, &J);
  render_worker(&J);
  for (int i = 1; i < n_threads; ++i) pthread_join(th[i], NULL);
  pthread_mutex_destroy(&J.mu);
  if (stats)
    for (int i = 0; i < GTO_NSTATS; ++i) stats[i] = J.stats[i];
  if (own) obvh_free(&bv);
  return 0;
}

/* ---------------- FP64 re-evaluation of given hit lists ---------------- */
static void radiance64(const float* rec, uint32_t deg, const double* dir, double* rgb) {
  static const double C0 = 0.28209479177387814, C1 = 0.4886025119029199;
  static const double C2[5] = {1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
                               -1.0925484305920792, 0.5462742152960396};
  static const double C3[7] = {-0.5900435899266435, 2.890611442640554, -0.4570457994644658,
                               0.3731763325901154, -0.4570457994644658, 1.445305721320277,
                               -0.5900435899266435};
  const float* sh = rec + 16;
  double x = dir[0], y = dir[1], z = dir[2];
  double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  double b[16];
  b[0] = C0;
  b[1] = -C1 * y; b[2] = C1 * z; b[3] = -C1 * x;
  b[4] = C2[0] * xy; b[5] = C2[1] * yz; b[6] = C2[2] * (2 * zz - xx - yy); b[7] = C2[3] * xz;
  b[8] = C2[4] * (xx - yy);
  b[9] = C3[0] * y * (3 * xx - yy); b[10] = C3[1] * xy * z; b[11] = C3[2] * y * (4 * zz - xx - yy);
  b[12] = C3[3] * z * (2 * zz - 3 * xx - 3 * yy); b[13] = C3[4] * x * (4 * zz - xx - yy);
  b[14] = C3[5] * z * (xx - yy); b[15] = C3[6] * x * (xx - 3 * yy);
  int nb = deg == 0 ? 1 : deg == 1 ? 4 : deg == 2 ? 9 : 16;
  for (int c = 0; c < 3; ++c) {
    double L = 0.5;
    for (int i = 0; i < nb; ++i) L += b[i] * (double)sh[4 * i + c];
    rgb[c] = L > 0 ? L : 0;
  }
}

static double response64(const float* rec, const double* o, const double* d) {
  double sc[3] = {rec[4], rec[5], rec[6]};
  double r = rec[8], x = rec[9], y = rec[10], z = rec[11];
  double R[3][3]; /* GLSL column-major R[c][r] as in the shader */
  R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - r * z); R[0][2] = 2 * (x * z + r * y);
  R[1][0] = 2 * (x * y + r * z); R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - r * x);
  R[2][0] = 2 * (x * z - r * y); R[2][1] = 2 * (y * z + r * x); R[2][2] = 1 - 2 * (x * x + y * y);
  double M[3][3], A[3][3];
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) M[c][rr] = R[c][rr] / sc[rr];
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) A[c][rr] = M[rr][0] * M[c][0] + M[rr][1] * M[c][1] + M[rr][2] * M[c][2];
  double mu[3] = {rec[0], rec[1], rec[2]};
  double tmp[3], mo[3];
  for (int rr = 0; rr < 3; ++rr) tmp[rr] = A[0][rr] * d[0] + A[1][rr] * d[1] + A[2][rr] * d[2];
  for (int j = 0; j < 3; ++j) mo[j] = mu[j] - o[j];
  double s = (mo[0] * tmp[0] + mo[1] * tmp[1] + mo[2] * tmp[2]) / (d[0] * tmp[0] + d[1] * tmp[1] + d[2] * tmp[2]);
  double pm[3], q[3];
  for (int j = 0; j < 3; ++j) pm[j] = o[j] + s * d[j] - mu[j];
  for (int rr = 0; rr < 3; ++rr) q[rr] = A[0][rr] * pm[0] + A[1][rr] * pm[1] + A[2][rr] * pm[2];
  return exp(-(pm[0] * q[0] + pm[1] * q[1] + pm[2] * q[2]));
}

void gto_reeval_f64(const gto_scene* s, const gto_params* p, const uint32_t* ray_list, uint32_t n_list,
                    const uint32_t* n_rounds, uint32_t max_rounds, const uint32_t* hit_ids,
                    const uint32_t* accept, double* rgb_out, double* T_out) {
  uint32_t n = ray_list ? n_list : p->width * p->height;
  const float* tw = p->to_world;
  double X[3] = {-(double)tw[0], -(double)tw[1], -(double)tw[2]};
  double Y[3] = {tw[4], tw[5], tw[6]};
  double Z[3] = {-(double)tw[8], -(double)tw[9], -(double)tw[10]};
  double o[3] = {tw[12], tw[13], tw[14]};
  for (uint32_t li = 0; li < n; ++li) {
    uint32_t ray = ray_list ? ray_list[li] : li;
    uint32_t px = ray % p->width, py = ray / p->width;
    float ox, oy;
    pixel_offset(p, px, py, &ox, &oy);
    double nx = ((double)px + ox) / p->width * 2.0 - 1.0;
    double ny = ((double)py + oy) / p->height * 2.0 - 1.0;
    double d[3];
    for (int j = 0; j < 3; ++j) d[j] = nx * p->aspect[0] * X[j] + ny * p->aspect[1] * Y[j] + Z[j];
    double dl = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int j = 0; j < 3; ++j) d[j] /= dl;
    double L[3] = {0, 0, 0}, T = 1.0;
    uint32_t R = n_rounds[li] < max_rounds ? n_rounds[li] : max_rounds;
    for (uint32_t r = 0; r < R; ++r) {
      uint32_t mask = accept[(size_t)li * max_rounds + r];
      for (uint32_t i = 0; i < p->k; ++i) {
        uint32_t id = hit_ids[((size_t)li * max_rounds + r) * p->k + i];
        if (id == GT_SENTINEL) break;
        if (!((mask >> i) & 1u)) continue;
        const float* rec = s->packed + (size_t)GTO_PACKED_FLOATS * id;
        double alpha = response64(rec, o, d) * (double)rec[12];
        double dir[3] = {rec[0] - o[0], rec[1] - o[1], rec[2] - o[2]};
        double l = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        for (int j = 0; j < 3; ++j) dir[j] /= l;
        double Lh[3];
        radiance64(rec, p->sh_degree_max, dir, Lh);
        for (int c = 0; c < 3; ++c) L[c] += T * alpha * Lh[c];
        T *= 1.0 - alpha;
      }
    }
    for (int c = 0; c < 3; ++c) rgb_out[3 * (size_t)li + c] = L[c];
    if (T_out) T_out[li] = T;
  }
}
