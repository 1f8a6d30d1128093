/*
 * gt_fp32.h - deterministic FP32 primitives of the Gaussian ray-tracing path.
 *
 * Every function here is written with an explicit operation order, no fused
 * multiply-add (all users compile with -ffp-contract=off) and IEEE-754
 * correctly rounded division and square root, so that the same source gives
 * bit-identical results in host C (gcc, SSE) and in gfx950 device code
 * (hipcc, -fhip-fp32-correctly-rounded-divide-sqrt, f32 denormals kept).
 * That is what makes per-ray hit lists and round counts of the HIP kernel
 * reproducible on the CPU.
 *
 * The formulas restate the reference's GLSL:
 *   ray generation          src/GS_shader.ts:277-285
 *   object-space ray        webrtx/src/glsl/trace.glsl:102-105
 *   triangle test           webrtx/src/glsl/intersect.glsl:73-92
 *   facing / back-face cull webrtx/src/glsl/trace.glsl:153-164
 *   Gaussian response       src/GS_shader.ts:111-136
 *   relaunch epsilon        src/GS_shader.ts:197, 248
 *   icosahedron proxy       src/main.ts:12-25
 * GLSL/WGSL leave dot-product summation order and exp accuracy to the
 * implementation; the orders fixed below ((x+y)+z, columns left to right)
 * and gt_expf are this library's definition of them.
 *
 * Plain C99 + HIP.  Include from C, C++ or .hip translation units.
 */
#ifndef GT_FP32_H
#define GT_FP32_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GT_HD __host__ __device__ static inline
#else
#include <math.h>
#include <string.h>
#define GT_HD static inline
#endif

#ifdef __cplusplus
extern "C++" {
#endif

typedef struct gt_f3 { float x, y, z; } gt_f3;

GT_HD gt_f3 gt_f3_make(float x, float y, float z) { gt_f3 r; r.x = x; r.y = y; r.z = z; return r; }
GT_HD gt_f3 gt_add(gt_f3 a, gt_f3 b) { return gt_f3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
GT_HD gt_f3 gt_sub(gt_f3 a, gt_f3 b) { return gt_f3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
GT_HD gt_f3 gt_scale(float s, gt_f3 a) { return gt_f3_make(s * a.x, s * a.y, s * a.z); }
GT_HD gt_f3 gt_neg(gt_f3 a) { return gt_f3_make(-a.x, -a.y, -a.z); }
/* dot: (x*x' + y*y') + z*z' */
GT_HD float gt_dot(gt_f3 a, gt_f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
/* cross: (a.y b.z - a.z b.y, a.z b.x - a.x b.z, a.x b.y - a.y b.x) */
GT_HD gt_f3 gt_cross(gt_f3 a, gt_f3 b) {
  return gt_f3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

GT_HD float gt_sqrtf(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_sqrtf(x); /* correctly rounded under -fhip-fp32-correctly-rounded-divide-sqrt */
#else
  return sqrtf(x);
#endif
}

/* normalize(v) = v / sqrt(dot(v, v)), each component divided (IEEE). */
GT_HD gt_f3 gt_normalize(gt_f3 v) {
  float l = gt_sqrtf(gt_dot(v, v));
  return gt_f3_make(v.x / l, v.y / l, v.z / l);
}

GT_HD float gt_bits_to_f32(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(u);
#else
  float f; memcpy(&f, &u, 4); return f;
#endif
}
GT_HD uint32_t gt_f32_to_bits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(f);
#else
  uint32_t u; memcpy(&u, &f, 4); return u;
#endif
}

GT_HD float gt_rintf(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_rintf(x);
#else
  return rintf(x); /* round-to-nearest-even in the default rounding mode */
#endif
}

/*
 * gt_expf: exp(x) in FP32 by Cody-Waite range reduction followed by a
 * degree-7 Taylor polynomial in Horner form (the classic scheme of
 * W. J. Cody and W. Waite, "Software Manual for the Elementary Functions",
 * 1980, ch. 6).
 *   n  = rint(x * log2(e))
 *   r  = (x - n*LN2_HI) - n*LN2_LO     LN2_HI has 16 significant bits, so
 *                                        n*LN2_HI is exact for |n| < 2^8
 *   e^r = 1 + r + r^2/2! + ... + r^7/7! on |r| <= ln2/2 (truncation < 8e-9 rel.)
 *   exp(x) = e^r * 2^n, with 2^n built from its exponent bits.
 * x <= -86 returns 0 (the result would be within a factor 2 of the smallest
 * normal; returning 0 keeps every result normal or zero so flush-to-zero
 * settings cannot change it), x >= 88.5 returns +inf, NaN returns 0.
 * Accuracy: within 2 ulp of exp over (-86, 88.5).
 */
#define GT_EXP_LOG2E 1.44269504088896341f
#define GT_EXP_LN2_HI 0.693145751953125f           /* 0x3f317200 */
#define GT_EXP_LN2_LO 1.42860682030941723212e-06f  /* ln2 - LN2_HI */
GT_HD float gt_expf(float x) {
  if (!(x > -86.0f)) return 0.0f;
  if (x >= 88.5f) return gt_bits_to_f32(0x7f800000u);
  float fn = gt_rintf(x * GT_EXP_LOG2E);
  int n = (int)fn;
  float r = x - fn * GT_EXP_LN2_HI;
  r = r - fn * GT_EXP_LN2_LO;
  float p = 1.0f / 5040.0f;
  p = p * r + 1.0f / 720.0f;
  p = p * r + 1.0f / 120.0f;
  p = p * r + 1.0f / 24.0f;
  p = p * r + 1.0f / 6.0f;
  p = p * r + 0.5f;
  p = p * r + 1.0f;
  p = p * r + 1.0f;
  /* n in [-124, 128]; split 2^n in two factors so both stay normal */
  int n1 = n / 2;
  int n2 = n - n1;
  float s1 = gt_bits_to_f32((uint32_t)(n1 + 127) << 23);
  float s2 = gt_bits_to_f32((uint32_t)(n2 + 127) << 23);
  return (p * s1) * s2;
}

/* ---------------------------------------------------------------------- */
/* Icosahedron proxy (src/main.ts:12-25): 12 vertices, 20 faces, inradius 1,
 * circumradius 1.2584.  Vertex coordinates are the f32 roundings of
 *   s = 1/r, t = (1+sqrt 5)/(2r), r = (3+sqrt 5)/(2 sqrt 3).            */
#define GT_ICO_S 0x1.52bb36p-1f   /* 0.661584556 */
#define GT_ICO_T 0x1.120a14p+0f   /* 1.070466280 */
#define GT_ICO_CIRCUMRADIUS 1.2584085723648188f
#define GT_NUM_FACES 20

#if defined(__HIP_DEVICE_COMPILE__)
#define GT_CONST_TABLE __device__ static const
#else
#define GT_CONST_TABLE static const
#endif

GT_CONST_TABLE float gt_ico_vertices[36] = {
    -GT_ICO_S, GT_ICO_T, 0.f, GT_ICO_S, GT_ICO_T, 0.f, -GT_ICO_S, -GT_ICO_T, 0.f, GT_ICO_S, -GT_ICO_T, 0.f,
    0.f, -GT_ICO_S, GT_ICO_T, 0.f, GT_ICO_S, GT_ICO_T, 0.f, -GT_ICO_S, -GT_ICO_T, 0.f, GT_ICO_S, -GT_ICO_T,
    GT_ICO_T, 0.f, -GT_ICO_S, GT_ICO_T, 0.f, GT_ICO_S, -GT_ICO_T, 0.f, -GT_ICO_S, -GT_ICO_T, 0.f, GT_ICO_S};
GT_CONST_TABLE uint8_t gt_ico_indices[60] = {
    0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
    3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};

GT_HD gt_f3 gt_ico_vertex(int v) {
  return gt_f3_make(gt_ico_vertices[3 * v], gt_ico_vertices[3 * v + 1], gt_ico_vertices[3 * v + 2]);
}

/* Per-face constants of the triangle test (intersect.glsl:78-80):
 *   e0 = p1 - p0, e1 = p0 - p2, n = cross(e1, e0)                        */
typedef struct gt_face { gt_f3 p0, e0, e1, n; } gt_face;

GT_HD gt_face gt_ico_face(int f) {
  gt_face r;
  gt_f3 p0 = gt_ico_vertex(gt_ico_indices[3 * f]);
  gt_f3 p1 = gt_ico_vertex(gt_ico_indices[3 * f + 1]);
  gt_f3 p2 = gt_ico_vertex(gt_ico_indices[3 * f + 2]);
  r.p0 = p0;
  r.e0 = gt_sub(p1, p0);
  r.e1 = gt_sub(p0, p2);
  r.n = gt_cross(r.e1, r.e0);
  return r;
}

/* ---------------------------------------------------------------------- */
/* Ray generation (src/GS_shader.ts:277-285).  to_world is the column-major
 * camera matrix (camera.ts: columns right, up, back, position).
 *   X = -col0, Y = col1, Z = -col2
 *   ndc = (id + 0.5) / size * 2 - 1
 *   dir = normalize(ndc.x*a.x*X + ndc.y*a.y*Y + Z)                       */
GT_HD gt_f3 gt_ray_dir(const float* to_world, float aspect_x, float aspect_y,
                       uint32_t px, uint32_t py, uint32_t width, uint32_t height) {
  gt_f3 X = gt_f3_make(-to_world[0], -to_world[1], -to_world[2]);
  gt_f3 Y = gt_f3_make(to_world[4], to_world[5], to_world[6]);
  gt_f3 Z = gt_f3_make(-to_world[8], -to_world[9], -to_world[10]);
  float nx = ((float)px + 0.5f) / (float)width * 2.0f - 1.0f;
  float ny = ((float)py + 0.5f) / (float)height * 2.0f - 1.0f;
  gt_f3 v = gt_add(gt_add(gt_scale(nx * aspect_x, X), gt_scale(ny * aspect_y, Y)), Z);
  return gt_normalize(v);
}

/* Object-space ray (trace.glsl:102-105): w2o is the 4x3 column-major
 * world-to-object matrix (12 floats: c0, c1, c2, c3).
 *   o_obj = ((c0*o.x + c1*o.y) + c2*o.z) + c3*1
 *   d_obj = ((c0*d.x + c1*d.y) + c2*d.z) + c3*0                          */
GT_HD gt_f3 gt_xform_point(const float* m, gt_f3 o) {
  return gt_f3_make(((m[0] * o.x + m[3] * o.y) + m[6] * o.z) + m[9] * 1.0f,
                    ((m[1] * o.x + m[4] * o.y) + m[7] * o.z) + m[10] * 1.0f,
                    ((m[2] * o.x + m[5] * o.y) + m[8] * o.z) + m[11] * 1.0f);
}
GT_HD gt_f3 gt_xform_dir(const float* m, gt_f3 d) {
  return gt_f3_make(((m[0] * d.x + m[3] * d.y) + m[6] * d.z) + m[9] * 0.0f,
                    ((m[1] * d.x + m[4] * d.y) + m[7] * d.z) + m[10] * 0.0f,
                    ((m[2] * d.x + m[5] * d.y) + m[8] * d.z) + m[11] * 0.0f);
}

/* Triangle test, intersect_triangle_branchless0 (intersect.glsl:73-92):
 *   e2 = (1/dot(n,d)) * (p0 - o); t = dot(n, e2)
 *   reject t <= tmin or tmax <= t
 *   i = cross(d, e2); beta = dot(i, e1); gamma = dot(i, e0)
 *   hit iff beta >= 0 && gamma >= 0 && beta + gamma <= 1
 * Returns 1 on hit and writes t.                                         */
GT_HD int gt_tri_hit(gt_f3 o, gt_f3 d, float tmin, float tmax, gt_face f, float* t_out) {
  float rcp = 1.0f / gt_dot(f.n, d);
  gt_f3 e2 = gt_scale(rcp, gt_sub(f.p0, o));
  float t = gt_dot(f.n, e2);
  if (t <= tmin || tmax <= t) return 0;
  gt_f3 i = gt_cross(d, e2);
  float beta = gt_dot(i, f.e1);
  float gamma = gt_dot(i, f.e0);
  *t_out = t;
  return (beta >= 0.0f) && (gamma >= 0.0f) && (beta + gamma <= 1.0f);
}

/* Facing (trace.glsl:154-164): n_w = n * W2O (row vector times the 4x3
 * matrix, xyz part), front iff dot(n_w, d_world) < 0; back faces are culled. */
GT_HD int gt_front_facing(const float* w2o, gt_f3 n, gt_f3 d_world) {
  gt_f3 nw = gt_f3_make(gt_dot(n, gt_f3_make(w2o[0], w2o[1], w2o[2])),
                        gt_dot(n, gt_f3_make(w2o[3], w2o[4], w2o[5])),
                        gt_dot(n, gt_f3_make(w2o[6], w2o[7], w2o[8])));
  return gt_dot(nw, d_world) < 0.0f;
}

/* ---------------------------------------------------------------------- */
/* Inverse covariance of computeResponse (src/GS_shader.ts:115-130).
 *   R (column-major GLSL mat3 of the quaternion (r,x,y,z)),
 *   M = invS * R  ->  M[c][r] = (1/s_r) * R[c][r]
 *   A = transpose(M) * M  ->  A[c][r] = (M[r][0]M[c][0] + M[r][1]M[c][1]) + M[r][2]M[c][2]
 * A is symmetric bit for bit; the six distinct entries are returned as
 * (a00, a01, a02, a11, a12, a22) with a_ij = A[i][j].                     */
GT_HD void gt_inv_cov(const float* scale, const float* quat, float* a6) {
  float r = quat[0], x = quat[1], y = quat[2], z = quat[3];
  float R[3][3];
  R[0][0] = 1.f - 2.f * (y * y + z * z); R[0][1] = 2.f * (x * y - r * z); R[0][2] = 2.f * (x * z + r * y);
  R[1][0] = 2.f * (x * y + r * z); R[1][1] = 1.f - 2.f * (x * x + z * z); R[1][2] = 2.f * (y * z - r * x);
  R[2][0] = 2.f * (x * z - r * y); R[2][1] = 2.f * (y * z + r * x); R[2][2] = 1.f - 2.f * (x * x + y * y);
  float is[3] = {1.0f / scale[0], 1.0f / scale[1], 1.0f / scale[2]};
  float M[3][3];
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) M[c][rr] = is[rr] * R[c][rr];
  float A[3][3];
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) A[c][rr] = (M[rr][0] * M[c][0] + M[rr][1] * M[c][1]) + M[rr][2] * M[c][2];
  a6[0] = A[0][0]; a6[1] = A[0][1]; a6[2] = A[0][2];
  a6[3] = A[1][1]; a6[4] = A[1][2]; a6[5] = A[2][2];
}

/* A * v for the symmetric A of gt_inv_cov: columns left to right,
 *   (A*v)_r = (A[0][r]*v.x + A[1][r]*v.y) + A[2][r]*v.z                   */
GT_HD gt_f3 gt_sym_mul(const float* a, gt_f3 v) {
  return gt_f3_make((a[0] * v.x + a[1] * v.y) + a[2] * v.z,
                    (a[1] * v.x + a[3] * v.y) + a[4] * v.z,
                    (a[2] * v.x + a[4] * v.y) + a[5] * v.z);
}

/* Max response of the Gaussian along the ray (computeResponse, 132-135):
 *   temp = A d; p = o + (dot(mu - o, temp) / dot(d, temp)) d
 *   return exp(dot(mu - p, A (p - mu)))     (no 1/2 factor)               */
GT_HD float gt_response(const float* a6, gt_f3 mu, gt_f3 o, gt_f3 d) {
  gt_f3 temp = gt_sym_mul(a6, d);
  float s = gt_dot(gt_sub(mu, o), temp) / gt_dot(d, temp);
  gt_f3 p = gt_add(o, gt_scale(s, d));
  gt_f3 q = gt_sym_mul(a6, gt_sub(p, mu));
  return gt_expf(gt_dot(gt_sub(mu, p), q));
}

/* Relaunch step (GS_shader.ts:197, 248): t_curr = t_k + 1e-4 in FP32. */
#define GT_RELAUNCH_EPS 1e-4f
#define GT_SENTINEL 0xffffffffu

#ifdef __cplusplus
}
#endif

#endif /* GT_FP32_H */
