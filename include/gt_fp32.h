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
 *   Gaussian response       src/GS_shader.ts:111-136 (evaluated in FP64, below)
 *   relaunch epsilon        src/GS_shader.ts:197, 248
 *   icosahedron proxy       src/main.ts:12-25
 * GLSL/WGSL leave dot-product summation order and exp accuracy to the
 * implementation; the orders fixed below ((x+y)+z, columns left to right)
 * and gt_exp64 are this library's definition of them.
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

GT_HD double gt_rint64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_rint(x);
#else
  return rint(x); /* round-to-nearest-even in the default rounding mode */
#endif
}
GT_HD double gt_bits_to_f64(uint64_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double((long long)u);
#else
  double f; memcpy(&f, &u, 8); return f;
#endif
}

/*
 * gt_exp64: exp(x) in FP64 by Cody-Waite range reduction followed by a
 * degree-13 Taylor polynomial in Horner form (W. J. Cody and W. Waite,
 * "Software Manual for the Elementary Functions", 1980, ch. 6):
 *   n  = rint(x * log2(e))
 *   r  = (x - n*LN2_HI) - n*LN2_LO     LN2_HI has 32 significant bits, so
 *                                        n*LN2_HI is exact for |n| < 2^20
 *   e^r = sum_{i<=13} r^i / i!  on |r| <= ln2/2   (truncation < 5e-18 rel.)
 *   exp(x) = e^r * 2^n, 2^n built from its exponent bits.
 * x <= -700 (and NaN) returns 0, x >= 700 returns +inf.  Relative error
 * below 2e-15 on (-700, 700).
 */
#define GT_EXP64_LOG2E 1.4426950408889634074
#define GT_EXP64_LN2_HI 6.93147180369123816490e-01  /* 0x3fe62e42fee00000 */
#define GT_EXP64_LN2_LO 1.90821492927058770002e-10  /* ln2 - LN2_HI */
GT_HD double gt_exp64(double x) {
  if (!(x > -700.0)) return 0.0;
  if (x >= 700.0) return gt_bits_to_f64(0x7ff0000000000000ull);
  double fn = gt_rint64(x * GT_EXP64_LOG2E);
  int n = (int)fn;
  double r = x - fn * GT_EXP64_LN2_HI;
  r = r - fn * GT_EXP64_LN2_LO;
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return p * gt_bits_to_f64((uint64_t)(n + 1023) << 52);
}

GT_HD uint64_t gt_f64_to_bits(double f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint64_t)__double_as_longlong(f);
#else
  uint64_t u; memcpy(&u, &f, 8); return u;
#endif
}

/*
 * gt_log64: natural log in FP64 for finite x > 0 (the activations' ln, used
 * identically by the host and the device scene builders).  With
 * x = m 2^e, m in [sqrt(1/2), sqrt(2)):
 *   s = (m - 1) / (m + 1),  ln m = 2 atanh(s) = 2s + 2s z P(z),  z = s^2,
 *   P(z) = sum_{i=0..10} z^i / (2i + 3)   (|s| <= 0.1716, truncation < 1e-18 rel.)
 *   ln x = e LN2_HI + (e LN2_LO + ln m)   (e LN2_HI exact, as in gt_exp64).
 * m - 1 and m + 1 are exact (Sterbenz); relative error about 1e-16.
 * x <= 0 or NaN returns NaN, +inf returns +inf; subnormals are rescaled.
 */
GT_HD double gt_log64(double x) {
  if (!(x > 0.0)) return gt_bits_to_f64(0x7ff8000000000000ull);
  if (x == gt_bits_to_f64(0x7ff0000000000000ull)) return x;
  int e = 0;
  if (x < 2.2250738585072014e-308) { x *= 18014398509481984.0; e = -54; } /* 2^54 */
  uint64_t u = gt_f64_to_bits(x);
  e += (int)((u >> 52) & 0x7ff) - 1023;
  double m = gt_bits_to_f64((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull); /* [1, 2) */
  if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 1.0 / 23.0;
  p = p * z + 1.0 / 21.0;
  p = p * z + 1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + 1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + 1.0 / 3.0;
  const double s2 = s + s;
  const double lm = s2 + s2 * (z * p);
  const double fe = (double)e;
  return (fe * GT_EXP64_LN2_LO + lm) + fe * GT_EXP64_LN2_HI;
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
/* Face pre-test.  For the ray o + t d and a directed edge a->b define the
 * Pluecker side value
 *   F(a->b) = (d x (a - o)) . (b - a) = d . (a x b) + (o x d) . (b - a).
 * For a face (p0, p1, p2) of gt_ico_face, the quantities of the triangle
 * test are beta = F(p2->p0) / dot(n, d), gamma = F(p0->p1) / dot(n, d) and
 * 1 - beta - gamma = F(p1->p2) / dot(n, d) (the three F sum to dot(n, d)),
 * so a FRONT face (dot(n, d) < 0) can pass only if all three F <= 0.
 * gt_face_candidates evaluates the 30 edge values once, without division,
 * and returns the mask of faces whose three signed values are below margin_g.
 * With margin_g far above the FP32 error of the triangle test, every face
 * the exact test + facing rule accepts is in the mask; the exact test then
 * runs only on those faces. */
#define GT_NUM_EDGES 30

/* edges: 30 x 6 floats {a x b, b - a} per undirected edge (a < b);
 * face_edges: 20 x 3 codes +-(e+1) of the directed edges (i0->i1),
 * (i1->i2), (i2->i0) of each face (negative: edge stored as b->a). */
static inline void gt_ico_edge_tables(float* edges, int8_t* face_edges) {
  int ea[GT_NUM_EDGES], eb[GT_NUM_EDGES], ne = 0;
  for (int f = 0; f < GT_NUM_FACES; ++f) {
    for (int j = 0; j < 3; ++j) {
      int a = gt_ico_indices[3 * f + j], b = gt_ico_indices[3 * f + (j + 1) % 3];
      int lo = a < b ? a : b, hi = a < b ? b : a, e = -1;
      for (int k = 0; k < ne; ++k)
        if (ea[k] == lo && eb[k] == hi) e = k;
      if (e < 0) { ea[ne] = lo; eb[ne] = hi; e = ne++; }
      face_edges[3 * f + j] = (int8_t)((a == lo ? 1 : -1) * (e + 1));
    }
  }
  for (int e = 0; e < GT_NUM_EDGES; ++e) {
    gt_f3 A = gt_ico_vertex(ea[e]), B = gt_ico_vertex(eb[e]);
    gt_f3 K = gt_cross(A, B), D = gt_sub(B, A);
    edges[6 * e + 0] = K.x; edges[6 * e + 1] = K.y; edges[6 * e + 2] = K.z;
    edges[6 * e + 3] = D.x; edges[6 * e + 4] = D.y; edges[6 * e + 5] = D.z;
  }
}

/* Margins of the pre-test for an object-space ray (oo, od) and the
 * world-to-object matrix w2o.  margin_g covers, about 7 times over, the
 * FP32 error of the triangle test's barycentrics and of the edge values in
 * F units, together at most ~26 ulp * |d|_1 (|o|_1 + 2) (|e| <= 1.33,
 * |p0| <= 1.26); margin_n is the facing pre-test bound
 * 1e-5 * max|n|_1 * sum|W2O| (the world-space facing dot and dot(n, d_obj)
 * differ by less than 8 ulp * |n|_1 * sum|W2O|). */
GT_HD void gt_face_margins(gt_f3 oo, gt_f3 od, const float* w2o, float* margin_g, float* margin_n) {
  float o1 = (fabsf(oo.x) + fabsf(oo.y)) + fabsf(oo.z);
  float d1 = (fabsf(od.x) + fabsf(od.y)) + fabsf(od.z);
  *margin_g = 1.2e-5f * d1 * (o1 + 2.0f);
  float ws = 0.0f;
  for (int i = 0; i < 9; ++i) ws += fabsf(w2o[i]);
  *margin_n = 3e-5f * ws;
}

/* The face_edges table of gt_ico_edge_tables as a literal, so that the
 * unrolled pre-test indexes the edge values statically (checked equal to
 * the generated table by the CPU tests). */
#define GT_ICO_FACE_EDGES                                                  \
  {1, -2, -3, 3, -4, -5, 5, 6, -7, 7, 8, -9, 9, 10, -1,                   \
   4, 11, -12, 2, -13, 14, -10, -15, 16, -8, -17, 18, -6, 19, -20,        \
   21, -22, -23, 23, -24, 25, -25, 26, -27, 27, 28, -29, 29, 30, -21,     \
   22, -11, -14, 24, 13, -16, -26, 15, -18, -28, 17, 20, -30, -19, 12}
GT_CONST_TABLE int8_t gt_ico_face_edges[60] = GT_ICO_FACE_EDGES;

#if defined(__HIP_DEVICE_COMPILE__)
#define GT_UNROLL _Pragma("unroll")
#else
#define GT_UNROLL
#endif

/* The edge constants {a x b, b - a} of gt_ico_edge_tables and the face
 * normals of gt_ico_face as literals (checked equal to the computed tables
 * by the CPU tests and by the library at its first launch): in the
 * unrolled pre-test each one becomes an instruction literal rather than a
 * value held in a register for the whole kernel. */
#define GT_ICO_EDGE_VALUES \
  {0x1.6a99b6p-1, 0x1.c032e4p-2, 0x1.255994p+0, -0x1.a2b1e4p-2, -0x1.120a14p+0, 0x1.52bb36p-1, \
   0x1.c032e4p-2, -0x1.255994p+0, 0x1.6a99b6p-1, -0x1.120a14p+0, -0x1.52bb36p-1, -0x1.a2b1e4p-2, \
   0x1.255994p+0, 0x1.6a99b6p-1, -0x1.c032e4p-2, 0x1.52bb36p-1, -0x1.a2b1e4p-2, 0x1.120a14p+0, \
   0x1.255994p+0, -0x1.6a99b6p-1, 0x1.c032e4p-2, -0x1.52bb36p-1, -0x1.a2b1e4p-2, 0x1.120a14p+0, \
   0x0p+0, 0x0p+0, -0x1.6a99b6p+0, 0x1.52bb36p+0, 0x0p+0, 0x0p+0, \
   -0x1.255994p+0, 0x1.6a99b6p-1, 0x1.c032e4p-2, -0x1.52bb36p-1, -0x1.a2b1e4p-2, -0x1.120a14p+0, \
   -0x1.255994p+0, -0x1.6a99b6p-1, -0x1.c032e4p-2, 0x1.52bb36p-1, -0x1.a2b1e4p-2, -0x1.120a14p+0, \
   -0x1.c032e4p-2, 0x1.255994p+0, 0x1.6a99b6p-1, -0x1.120a14p+0, -0x1.52bb36p-1, 0x1.a2b1e4p-2, \
   -0x1.6a99b6p-1, -0x1.c032e4p-2, 0x1.255994p+0, -0x1.a2b1e4p-2, -0x1.120a14p+0, -0x1.52bb36p-1, \
   0x0p+0, 0x1.6a99b6p+0, 0x0p+0, 0x0p+0, 0x0p+0, 0x1.52bb36p+0, \
   0x1.c032e4p-2, 0x1.255994p+0, -0x1.6a99b6p-1, 0x1.120a14p+0, -0x1.52bb36p-1, -0x1.a2b1e4p-2, \
   0x1.6a99b6p-1, -0x1.c032e4p-2, -0x1.255994p+0, 0x1.a2b1e4p-2, -0x1.120a14p+0, 0x1.52bb36p-1, \
   -0x1.c032e4p-2, -0x1.255994p+0, -0x1.6a99b6p-1, -0x1.120a14p+0, 0x1.52bb36p-1, -0x1.a2b1e4p-2, \
   -0x1.6a99b6p+0, 0x0p+0, 0x0p+0, 0x0p+0, 0x1.52bb36p+0, 0x0p+0, \
   0x1.6a99b6p-1, -0x1.c032e4p-2, -0x1.255994p+0, -0x1.a2b1e4p-2, 0x1.120a14p+0, -0x1.52bb36p-1, \
   -0x1.6a99b6p-1, 0x1.c032e4p-2, -0x1.255994p+0, -0x1.a2b1e4p-2, 0x1.120a14p+0, 0x1.52bb36p-1, \
   0x1.6a99b6p+0, 0x0p+0, 0x0p+0, 0x0p+0, 0x1.52bb36p+0, 0x0p+0, \
   0x1.c032e4p-2, 0x1.255994p+0, -0x1.6a99b6p-1, -0x1.120a14p+0, 0x1.52bb36p-1, 0x1.a2b1e4p-2, \
   -0x1.6a99b6p-1, 0x1.c032e4p-2, -0x1.255994p+0, 0x1.a2b1e4p-2, -0x1.120a14p+0, -0x1.52bb36p-1, \
   -0x1.c032e4p-2, -0x1.255994p+0, -0x1.6a99b6p-1, 0x1.120a14p+0, -0x1.52bb36p-1, 0x1.a2b1e4p-2, \
   -0x1.6a99b6p-1, -0x1.c032e4p-2, 0x1.255994p+0, 0x1.a2b1e4p-2, 0x1.120a14p+0, 0x1.52bb36p-1, \
   -0x1.c032e4p-2, 0x1.255994p+0, 0x1.6a99b6p-1, 0x1.120a14p+0, 0x1.52bb36p-1, -0x1.a2b1e4p-2, \
   -0x1.255994p+0, -0x1.6a99b6p-1, -0x1.c032e4p-2, -0x1.52bb36p-1, 0x1.a2b1e4p-2, 0x1.120a14p+0, \
   -0x1.255994p+0, 0x1.6a99b6p-1, 0x1.c032e4p-2, 0x1.52bb36p-1, 0x1.a2b1e4p-2, 0x1.120a14p+0, \
   0x0p+0, 0x0p+0, 0x1.6a99b6p+0, 0x1.52bb36p+0, 0x0p+0, 0x0p+0, \
   0x1.255994p+0, -0x1.6a99b6p-1, 0x1.c032e4p-2, 0x1.52bb36p-1, 0x1.a2b1e4p-2, -0x1.120a14p+0, \
   0x1.255994p+0, 0x1.6a99b6p-1, -0x1.c032e4p-2, -0x1.52bb36p-1, 0x1.a2b1e4p-2, -0x1.120a14p+0, \
   0x1.c032e4p-2, -0x1.255994p+0, 0x1.6a99b6p-1, 0x1.120a14p+0, 0x1.52bb36p-1, 0x1.a2b1e4p-2, \
   0x1.6a99b6p-1, 0x1.c032e4p-2, 0x1.255994p+0, 0x1.a2b1e4p-2, 0x1.120a14p+0, -0x1.52bb36p-1, \
   0x0p+0, -0x1.6a99b6p+0, 0x0p+0, 0x0p+0, 0x0p+0, 0x1.52bb36p+0}
#define GT_ICO_FACE_NORMALS \
  {-0x1.c032e4p-1, 0x1.c032e4p-1, 0x1.c032e4p-1, \
   0x0p+0, 0x1.6a99b6p+0, 0x1.150088p-1, \
   0x0p+0, 0x1.6a99b6p+0, -0x1.150088p-1, \
   -0x1.c032e4p-1, 0x1.c032e4p-1, -0x1.c032e4p-1, \
   -0x1.6a99b6p+0, 0x1.150088p-1, 0x0p+0, \
   0x1.c032e4p-1, 0x1.c032e4p-1, 0x1.c032e4p-1, \
   -0x1.150088p-1, 0x0p+0, 0x1.6a99b6p+0, \
   -0x1.6a99b6p+0, -0x1.150088p-1, -0x0p+0, \
   -0x1.150088p-1, 0x0p+0, -0x1.6a99b6p+0, \
   0x1.c032e4p-1, 0x1.c032e4p-1, -0x1.c032e4p-1, \
   0x1.c032e4p-1, -0x1.c032e4p-1, 0x1.c032e4p-1, \
   0x0p+0, -0x1.6a99b6p+0, 0x1.150088p-1, \
   -0x0p+0, -0x1.6a99b6p+0, -0x1.150088p-1, \
   0x1.c032e4p-1, -0x1.c032e4p-1, -0x1.c032e4p-1, \
   0x1.6a99b6p+0, -0x1.150088p-1, 0x0p+0, \
   0x1.150088p-1, 0x0p+0, 0x1.6a99b6p+0, \
   -0x1.c032e4p-1, -0x1.c032e4p-1, 0x1.c032e4p-1, \
   -0x1.c032e4p-1, -0x1.c032e4p-1, -0x1.c032e4p-1, \
   0x1.150088p-1, 0x0p+0, -0x1.6a99b6p+0, \
   0x1.6a99b6p+0, 0x1.150088p-1, 0x0p+0}

/* Faces of each edge: edge e is used with + orientation by face
 * GT_ICO_EDGE_FACES[2e] and with - orientation by face [2e+1] (the inverse
 * of GT_ICO_FACE_EDGES; every edge borders two faces). */
#define GT_ICO_EDGE_FACES \
  {0, 4, 6, 0, 1, 0, 5, 1, 2, 1, 2, 9, 3, 2, 3, 8, 4, 3, 4, 7, 5, 15, 19, 5, 16, 6, 6, 15, 17, 7, \
   7, 16, 18, 8, 8, 17, 9, 19, 18, 9, 10, 14, 15, 10, 11, 10, 16, 11, 11, 12, 12, 17, 13, 12, 13, 18, 14, 13, 14, 19}

/* Undirected edges of gt_ico_edge_tables as vertex index pairs (a, b), a < b. */
#define GT_ICO_EDGE_VERTS                                                                         \
  {0, 11, 5, 11, 0, 5, 1, 5, 0, 1, 1, 7, 0, 7, 7, 10, 0, 10, 10, 11, 5, 9, 1, 9, 4, 11, 4, 5, 2, 10, \
   2, 11, 6, 7, 6, 10, 1, 8, 7, 8, 3, 9, 4, 9, 3, 4, 2, 4, 2, 3, 2, 6, 3, 6, 6, 8, 3, 8, 8, 9}
#define GT_ICO_VERTEX_VALUES                                                                      \
  {-GT_ICO_S, GT_ICO_T, 0.f, GT_ICO_S, GT_ICO_T, 0.f, -GT_ICO_S, -GT_ICO_T, 0.f, GT_ICO_S, -GT_ICO_T, 0.f, \
   0.f, -GT_ICO_S, GT_ICO_T, 0.f, GT_ICO_S, GT_ICO_T, 0.f, -GT_ICO_S, -GT_ICO_T, 0.f, GT_ICO_S, -GT_ICO_T, \
   GT_ICO_T, 0.f, -GT_ICO_S, GT_ICO_T, 0.f, GT_ICO_S, -GT_ICO_T, 0.f, -GT_ICO_S, -GT_ICO_T, 0.f, GT_ICO_S}

/* Sum of the nonzero terms a[j] * c[j] in order j = 0, 1, 2 (c are
 * literals: after unrolling the zero tests fold away). */
GT_HD float gt_sparse_dot(gt_f3 a, float c0, float c1, float c2) {
  float acc = 0.0f;
  int first = 1;
  if (c0 != 0.0f) { acc = a.x * c0; first = 0; }
  if (c1 != 0.0f) { const float p = a.y * c1; acc = first ? p : acc + p; first = 0; }
  if (c2 != 0.0f) { const float p = a.z * c2; acc = first ? p : acc + p; first = 0; }
  return acc;
}

/* Face-major evaluation through vertex terms.  With w = o x d,
 *   F(a->b) = d . (a x b) + w . b - w . a,
 * so the vertex terms W_v = w . v and the products of d with the few
 * distinct a x b components are formed once, and an edge value costs a few
 * adds.  The icosahedron is centrally symmetric: each vertex v has the
 * antipode -v (GT_ICO_VERTEX_BASE / GT_ICO_VERTEX_SIGN: v = sign * base
 * vertex), so W_{-v} = -W_v exactly, and the edge (a, b) and its antipodal
 * edge (-a, -b) share a x b.  For an antipodal pair (GT_ICO_EDGE_PAIRS:
 * edge e, edge e', s)
 *   F(e)  = d . K + D,   D = W_b - W_a,
 *   F(e') = s (d . K - D),
 * with s = -1 where the table stores e' as (-b, -a).  So 6 vertex terms,
 * 15 dot products with a x b and 15 differences D give the 30 edge values.
 * A face (edges e1 e2 e3 with orientations) is a candidate iff
 * max(+-F_e1, +-F_e2, +-F_e3) - margin_g < 0; the sign bit of that value
 * is shifted into the mask, from the last face to the first (face f at
 * bit f).  The rounding error of this evaluation is below 25 ulp * |d|_1
 * (|o|_1 + 1), inside margin_g with the triangle test's own error.  A NaN
 * edge value makes fmaxf ignore it; a face whose three values are all NaN
 * may come out either way, and the exact test rejects it.  No facing
 * pre-test: a back face passes only when all three signed values lie in
 * [0, margin_g] (its three F sum to dot(n, d) > 0 with the opposite sign
 * pattern), i.e. for faces nearly parallel to the ray, and the exact
 * test's facing rule rejects it; margin_n is kept for the interface. */
#define GT_ICO_VERTEX_BASE {0, 1, 1, 0, 2, 3, 3, 2, 4, 5, 5, 4}
#define GT_ICO_VERTEX_SIGN {1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1}
#define GT_ICO_BASE_VERTS {0, 1, 4, 5, 8, 9}
#define GT_ICO_EDGE_PAIRS                                                                                  \
  {0, 28, 1, 1, 27, 1, 2, 26, 1, 3, 25, 1, 4, 24, -1, 5, 23, 1, 6, 22, 1, 7, 21, 1, 8, 20, 1, 9, 29, -1, \
   10, 17, 1, 11, 14, 1, 12, 19, 1, 13, 16, -1, 15, 18, 1}
GT_HD uint32_t gt_face_candidates(gt_f3 oo, gt_f3 od, float margin_g, float margin_n) {
  const float edges[6 * GT_NUM_EDGES] = GT_ICO_EDGE_VALUES;
  const uint8_t ev[2 * GT_NUM_EDGES] = GT_ICO_EDGE_VERTS;
  const float vtx[36] = GT_ICO_VERTEX_VALUES;
  const int8_t fe[3 * GT_NUM_FACES] = GT_ICO_FACE_EDGES;
  const int8_t vbase[12] = GT_ICO_VERTEX_BASE, vsign[12] = GT_ICO_VERTEX_SIGN;
  const uint8_t bverts[6] = GT_ICO_BASE_VERTS;
  const int8_t pairs[45] = GT_ICO_EDGE_PAIRS;
  const gt_f3 w = gt_cross(oo, od);
  float W6[6], F[GT_NUM_EDGES];
  (void)margin_n;
  GT_UNROLL
  for (int j = 0; j < 6; ++j) {
    const int v = bverts[j];
    W6[j] = gt_sparse_dot(w, vtx[3 * v], vtx[3 * v + 1], vtx[3 * v + 2]);
  }
  GT_UNROLL
  for (int p = 0; p < 15; ++p) {
    const int e = pairs[3 * p], e2 = pairs[3 * p + 1], sg = pairs[3 * p + 2];
    const int a = ev[2 * e], b = ev[2 * e + 1];
    const float* K = edges + 6 * e;
    const float kd = gt_sparse_dot(od, K[0], K[1], K[2]);
    const float wb = vsign[b] > 0 ? W6[vbase[b]] : -W6[vbase[b]];
    const float wa = vsign[a] > 0 ? W6[vbase[a]] : -W6[vbase[a]];
    const float D = wb - wa;
    F[e] = kd + D;
    const float f2 = kd - D;
    F[e2] = sg > 0 ? f2 : -f2;
  }
  uint32_t cand = 0;
  GT_UNROLL
  for (int f = GT_NUM_FACES - 1; f >= 0; --f) {
    float val[3];
    GT_UNROLL
    for (int j = 0; j < 3; ++j) {
      const int c = fe[3 * f + j], e = (c > 0 ? c : -c) - 1;
      val[j] = c > 0 ? F[e] : -F[e];
    }
    const float m = fmaxf(fmaxf(val[0], val[1]), val[2]) - margin_g;
#if defined(__HIP_DEVICE_COMPILE__)
    cand = __builtin_amdgcn_alignbit(cand, __float_as_uint(m), 31u); /* (cand << 1) | sign bit of m */
#else
    uint32_t mb;
    memcpy(&mb, &m, 4);
    cand = (cand << 1) | (mb >> 31);
#endif
  }
  return cand;
}

/* 1 if the literal tables equal the computed ones (host check). */
static inline int gt_ico_literals_match(void) {
  float e[6 * GT_NUM_EDGES];
  int8_t fe[3 * GT_NUM_FACES];
  const float le[6 * GT_NUM_EDGES] = GT_ICO_EDGE_VALUES;
  const float ln[3 * GT_NUM_FACES] = GT_ICO_FACE_NORMALS;
  const int8_t lf[3 * GT_NUM_FACES] = GT_ICO_FACE_EDGES;
  gt_ico_edge_tables(e, fe);
  for (int i = 0; i < 6 * GT_NUM_EDGES; ++i)
    if (e[i] != le[i]) return 0;
  for (int f = 0; f < GT_NUM_FACES; ++f) {
    gt_face fc = gt_ico_face(f);
    if (fc.n.x != ln[3 * f] || fc.n.y != ln[3 * f + 1] || fc.n.z != ln[3 * f + 2]) return 0;
  }
  for (int i = 0; i < 3 * GT_NUM_FACES; ++i)
    if (fe[i] != lf[i]) return 0;
  {
    const uint8_t lv[2 * GT_NUM_EDGES] = GT_ICO_EDGE_VERTS;
    const float vv[36] = GT_ICO_VERTEX_VALUES;
    for (int i = 0; i < 36; ++i)
      if (vv[i] != gt_ico_vertices[i]) return 0;
    for (int e = 0; e < GT_NUM_EDGES; ++e) {
      const gt_f3 A = gt_ico_vertex(lv[2 * e]), B = gt_ico_vertex(lv[2 * e + 1]);
      const gt_f3 K = gt_cross(A, B), D = gt_sub(B, A);
      if (lv[2 * e] >= lv[2 * e + 1] || K.x != le[6 * e] || K.y != le[6 * e + 1] || K.z != le[6 * e + 2] ||
          D.x != le[6 * e + 3] || D.y != le[6 * e + 4] || D.z != le[6 * e + 5])
        return 0;
    }
  }
  {
    const uint8_t ef[2 * GT_NUM_EDGES] = GT_ICO_EDGE_FACES;
    for (int f = 0; f < GT_NUM_FACES; ++f)
      for (int j = 0; j < 3; ++j) {
        const int c = lf[3 * f + j], e = (c > 0 ? c : -c) - 1;
        if (ef[2 * e + (c > 0 ? 0 : 1)] != f) return 0;
      }
  }
  {
    /* antipodes: vertex v = sign * base vertex, and each edge pair maps onto
     * itself under v -> -v with the stated orientation */
    const int8_t vb[12] = GT_ICO_VERTEX_BASE, vs[12] = GT_ICO_VERTEX_SIGN;
    const uint8_t bv[6] = GT_ICO_BASE_VERTS;
    const int8_t pr[45] = GT_ICO_EDGE_PAIRS;
    const uint8_t lv[2 * GT_NUM_EDGES] = GT_ICO_EDGE_VERTS;
    int anti[12], seen[GT_NUM_EDGES] = {0};
    for (int v = 0; v < 12; ++v) {
      if (vb[v] < 0 || vb[v] > 5 || (vs[v] != 1 && vs[v] != -1)) return 0;
      const gt_f3 P = gt_ico_vertex(v), B = gt_ico_vertex(bv[vb[v]]);
      if (P.x != vs[v] * B.x || P.y != vs[v] * B.y || P.z != vs[v] * B.z) return 0;
      anti[v] = -1;
      for (int u = 0; u < 12; ++u) {
        const gt_f3 Q = gt_ico_vertex(u);
        if (Q.x == -P.x && Q.y == -P.y && Q.z == -P.z) anti[v] = u;
      }
      if (anti[v] < 0) return 0;
    }
    for (int p = 0; p < 15; ++p) {
      const int e = pr[3 * p], e2 = pr[3 * p + 1], sg = pr[3 * p + 2];
      if (e < 0 || e >= GT_NUM_EDGES || e2 < 0 || e2 >= GT_NUM_EDGES || seen[e] || seen[e2]) return 0;
      seen[e] = seen[e2] = 1;
      const int a = anti[lv[2 * e]], b = anti[lv[2 * e + 1]];
      if (sg == 1 ? (lv[2 * e2] != a || lv[2 * e2 + 1] != b) : (sg != -1 || lv[2 * e2] != b || lv[2 * e2 + 1] != a))
        return 0;
    }
  }
  return 1;
}

/* ---------------------------------------------------------------------- */
/* Ray generation (src/GS_shader.ts:277-285).  to_world is the column-major
 * camera matrix (camera.ts: columns right, up, back, position).
 *   X = -col0, Y = col1, Z = -col2
 *   ndc = (id + 0.5) / size * 2 - 1
 *   dir = normalize(ndc.x*a.x*X + ndc.y*a.y*Y + Z)                       */
GT_HD gt_f3 gt_ray_dir_off(const float* to_world, float aspect_x, float aspect_y, uint32_t px, uint32_t py,
                           uint32_t width, uint32_t height, float off_x, float off_y) {
  gt_f3 X = gt_f3_make(-to_world[0], -to_world[1], -to_world[2]);
  gt_f3 Y = gt_f3_make(to_world[4], to_world[5], to_world[6]);
  gt_f3 Z = gt_f3_make(-to_world[8], -to_world[9], -to_world[10]);
  float nx = ((float)px + off_x) / (float)width * 2.0f - 1.0f;
  float ny = ((float)py + off_y) / (float)height * 2.0f - 1.0f;
  gt_f3 v = gt_add(gt_add(gt_scale(nx * aspect_x, X), gt_scale(ny * aspect_y, Y)), Z);
  return gt_normalize(v);
}
GT_HD gt_f3 gt_ray_dir(const float* to_world, float aspect_x, float aspect_y,
                       uint32_t px, uint32_t py, uint32_t width, uint32_t height) {
  return gt_ray_dir_off(to_world, aspect_x, aspect_y, px, py, width, height, 0.5f, 0.5f);
}

/* Antialiasing jitter of the reference's ray generation, compiled out there
 * (antialiasing = false, GS_shader.ts:54) and opt-in here.  The pixel seed
 * is getNewSeed(y*W + x, accumulatedFrames, 8) (GS_shader.ts:8-22): rounds
 * of the Tiny Encryption Algorithm (Wheeler & Needham 1994) with the
 * shader's four 32-bit key words, returning the first word.  rng
 * (GS_shader.ts:24-28) is the LCG seed = 1664525 seed + 1013904223 with the
 * low 24 bits scaled to [0, 1) (exact in FP32).  All integer arithmetic:
 * identical on host and device. */
GT_HD uint32_t gt_aa_seed(uint32_t v0, uint32_t v1, uint32_t rounds) {
  uint32_t sum = 0;
  for (uint32_t i = 0; i < rounds; ++i) {
    sum += 0x9e3779b9u;
    v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xc8013ea4u);
    v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7e95761eu);
  }
  return v0;
}
GT_HD float gt_aa_rng(uint32_t* seed) {
  *seed = 1664525u * *seed + 1013904223u;
  return (float)(*seed & 0x00ffffffu) / (float)0x01000000u;
}
/* pixel offset of pixel (px, py) in frame `frames` (GS_shader.ts:282-283) */
GT_HD void gt_aa_offset(uint32_t px, uint32_t py, uint32_t width, uint32_t frames, float* off_x, float* off_y) {
  uint32_t seed = gt_aa_seed(width * py + px, frames, 8u);
  *off_x = gt_aa_rng(&seed);
  *off_y = gt_aa_rng(&seed);
}
/* running average of the accumulated frames (GS_shader.ts:303-307):
 * frames == 0 ? c : mix(prev, c, 1 / (frames + 1)), mix(x, y, a) = x (1 - a) + y a */
GT_HD float gt_aa_mix(float prev, float c, uint32_t frames) {
  if (frames == 0) return c;
  const float a = 1.0f / ((float)frames + 1.0f);
  return prev * (1.0f - a) + c * a;
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
/* Max response of the Gaussian along the ray, computeResponse
 * (src/GS_shader.ts:111-136): exp(-min_t (o + t d - mu)^T Sigma^-1 (o + t d - mu))
 * with Sigma^-1 = (S^-1 R)^T (S^-1 R), R the shader's column-major mat3 of the
 * quaternion (r, x, y, z), and no 1/2 in the exponent.
 *
 * Evaluated in FP64 from the FP32 inputs, in the whitened frame:
 *   a = S^-1 R (o - mu),  b = S^-1 R d,  min = |a x b|^2 / |b|^2.
 * This is the same function as the reference's p = o + (dot(mu-o, Ad) /
 * dot(d, Ad)) d, exp(dot(mu - p, A (p - mu))), but the literal FP32 form
 * loses up to 3e-3 in the response on flat Gaussians (scale ratios of 500 in
 * the reference's data/pc_short.ply), because forming A squares the
 * condition number; FP64 keeps it within 1e-12 of the exact value.
 * Returns the response in double; alpha = (float)(response * opacity). */
GT_HD double gt_response64(const float* mu, const float* scale, const float* quat, gt_f3 o, gt_f3 d) {
  const double r = quat[0], x = quat[1], y = quat[2], z = quat[3];
  /* R[c][row] as GS_shader.ts:119-122 */
  const double R00 = 1.0 - 2.0 * (y * y + z * z), R01 = 2.0 * (x * y - r * z), R02 = 2.0 * (x * z + r * y);
  const double R10 = 2.0 * (x * y + r * z), R11 = 1.0 - 2.0 * (x * x + z * z), R12 = 2.0 * (y * z - r * x);
  const double R20 = 2.0 * (x * z - r * y), R21 = 2.0 * (y * z + r * x), R22 = 1.0 - 2.0 * (x * x + y * y);
  const double i0 = 1.0 / (double)scale[0], i1 = 1.0 / (double)scale[1], i2 = 1.0 / (double)scale[2];
  const double u0 = (double)o.x - (double)mu[0], u1 = (double)o.y - (double)mu[1], u2 = (double)o.z - (double)mu[2];
  const double d0 = d.x, d1 = d.y, d2 = d.z;
  /* (invS * R) v: row k = (1/s_k) * ((R[0][k] v0 + R[1][k] v1) + R[2][k] v2) */
  const double a0 = i0 * ((R00 * u0 + R10 * u1) + R20 * u2);
  const double a1 = i1 * ((R01 * u0 + R11 * u1) + R21 * u2);
  const double a2 = i2 * ((R02 * u0 + R12 * u1) + R22 * u2);
  const double b0 = i0 * ((R00 * d0 + R10 * d1) + R20 * d2);
  const double b1 = i1 * ((R01 * d0 + R11 * d1) + R21 * d2);
  const double b2 = i2 * ((R02 * d0 + R12 * d1) + R22 * d2);
  const double c0 = a1 * b2 - a2 * b1, c1 = a2 * b0 - a0 * b2, c2 = a0 * b1 - a1 * b0;
  const double e = ((c0 * c0 + c1 * c1) + c2 * c2) / ((b0 * b0 + b1 * b1) + b2 * b2);
  return gt_exp64(-e);
}

/* Relaunch step (GS_shader.ts:197, 248): t_curr = t_k + 1e-4 in FP32. */
#define GT_RELAUNCH_EPS 1e-4f
#define GT_SENTINEL 0xffffffffu

#ifdef __cplusplus
}
#endif

#endif /* GT_FP32_H */
