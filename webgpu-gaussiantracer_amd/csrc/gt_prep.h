/*
 * gt_prep.h - per-Gaussian scene preparation, shared by the host builder
 * (gt_scene.cpp, g++) and the device builder (gt_build.hip, hipcc) so that
 * both produce bit-identical records, proxy transforms and proxy boxes.
 *
 *   activations + record layout   src/GS_ply.ts:187-196, 225-235
 *   proxy transform               src/main.ts:73-94
 *   inverse                       webrtx/bvh/src/lib.rs:501-505 (glam Affine3A::inverse)
 *   instance box                  webrtx/bvh/src/lib.rs:482-499 (padded here, see gt_prep_box)
 *
 * Double arithmetic with a fixed operation order, no contraction
 * (-ffp-contract=off on both compilers), IEEE division and square root, and
 * this library's exp/log (gt_exp64, gt_log64) in place of the JS engine's
 * Math.exp / Math.log.
 */
#ifndef GT_PREP_H
#define GT_PREP_H

#include <stdint.h>

#include "../../include/gt.h"
#include "../../include/gt_fp32.h"

#if defined(__HIP_DEVICE_COMPILE__)
#define GT_PREP_SQRT64(x) __builtin_sqrt(x)
#define GT_PREP_NEXTUP(x) nextafterf((x), __builtin_inff())
#define GT_PREP_NEXTDOWN(x) nextafterf((x), -__builtin_inff())
#else
#include <math.h>
#define GT_PREP_SQRT64(x) sqrt(x)
#define GT_PREP_NEXTUP(x) nextafterf((x), INFINITY)
#define GT_PREP_NEXTDOWN(x) nextafterf((x), -INFINITY)
#endif

#ifdef __cplusplus
extern "C++" {
#endif

/* Activations of one PLY-raw record r (59 floats) in double, as the JS
 * loader does: scale = exp(log-scale), q = raw / |raw| (w, x, y, z order),
 * opacity = sigmoid(logit).  act = {opacity, scale[3], q[4]}.  geom (12
 * floats) = {mu, f32 opacity} {f32 scale, f32 |raw q|} {f32 q};
 * sh (48 floats) = sh[j][c] with sh[0] = f_dc and
 * sh[1+j] = (f_rest_j, f_rest_{15+j}, f_rest_{30+j}). */
GT_HD void gt_prep_activate(const float* r, double* act, float* geom, float* sh) {
  const double sc0 = gt_exp64((double)r[52]), sc1 = gt_exp64((double)r[53]), sc2 = gt_exp64((double)r[54]);
  double q0 = r[55], q1 = r[56], q2 = r[57], q3 = r[58];
  const double norm = GT_PREP_SQRT64(((q0 * q0 + q1 * q1) + q2 * q2) + q3 * q3);
  q0 = q0 / norm; q1 = q1 / norm; q2 = q2 / norm; q3 = q3 / norm;
  const double op = 1.0 / (1.0 + gt_exp64(-(double)r[51]));
  act[0] = op; act[1] = sc0; act[2] = sc1; act[3] = sc2;
  act[4] = q0; act[5] = q1; act[6] = q2; act[7] = q3;
  geom[0] = r[0]; geom[1] = r[1]; geom[2] = r[2]; geom[3] = (float)op;
  geom[4] = (float)sc0; geom[5] = (float)sc1; geom[6] = (float)sc2; geom[7] = (float)norm;
  geom[8] = (float)q0; geom[9] = (float)q1; geom[10] = (float)q2; geom[11] = (float)q3;
  for (int c = 0; c < 3; ++c) sh[c] = r[3 + c];
  for (int j = 0; j < 15; ++j)
    for (int c = 0; c < 3; ++c) sh[3 * (1 + j) + c] = r[6 + c * 15 + j];
}

/* Whether PLY-raw record r, with its activations act (gt_prep_activate),
 * describes a finite Gaussian: every raw value finite, a quaternion of
 * nonzero norm (else q / |q| is 0/0) and scales exp(log-scale) finite and
 * nonzero in f32 (else the proxy transform is singular or infinite).  The
 * reference has no such check (GS_ply.ts:225-235 divides by a zero norm and
 * builds NaN proxies); here such a record is an argument error. */
GT_HD int gt_prep_valid(const float* r, const double* act) {
  int ok = 1;
  for (int j = 0; j < GT_RAW_FLOATS; ++j) ok &= __builtin_isfinite(r[j]) ? 1 : 0;
  for (int j = 1; j < 4; ++j) {
    const float sc = (float)act[j];
    ok &= (sc > 0.0f && __builtin_isfinite(sc)) ? 1 : 0;
  }
  for (int j = 4; j < 8; ++j) ok &= __builtin_isfinite(act[j]) ? 1 : 0;
  return ok;
}

/* Rotation of wgpu-matrix mat3.fromQuat([x, y, z, w]) evaluated in double
 * and stored as f32; R[3c + r] = column c, row r. */
GT_HD void gt_prep_quat_to_mat(double x, double y, double z, double w, float* R) {
  const double x2 = x + x, y2 = y + y, z2 = z + z;
  const double xx = x * x2, yx = y * x2, yy = y * y2;
  const double zx = z * x2, zy = z * y2, zz = z * z2;
  const double wx = w * x2, wy = w * y2, wz = w * z2;
  R[0] = (float)(1 - yy - zz); R[1] = (float)(yx + wz); R[2] = (float)(zx - wy);
  R[3] = (float)(yx - wz); R[4] = (float)(1 - xx - zz); R[5] = (float)(zy + wx);
  R[6] = (float)(zx + wy); R[7] = (float)(zy - wx); R[8] = (float)(1 - xx - yy);
}

/* glam Affine3A::inverse (f32, scalar build): rows of the inverse 3x3 are
 * (y x z, z x x, x x y) / det(z . (x x y)), translation = -(inv * t). */
GT_HD void gt_prep_invert_affine(const float* m, float* out) {
  const gt_f3 cx = gt_f3_make(m[0], m[1], m[2]);
  const gt_f3 cy = gt_f3_make(m[3], m[4], m[5]);
  const gt_f3 cz = gt_f3_make(m[6], m[7], m[8]);
  const gt_f3 row0 = gt_cross(cy, cz), row1 = gt_cross(cz, cx), row2 = gt_cross(cx, cy);
  const float inv_det = 1.0f / gt_dot(cz, row2);
  const gt_f3 a = gt_scale(inv_det, row0), b = gt_scale(inv_det, row1), c = gt_scale(inv_det, row2);
  const gt_f3 k0 = gt_f3_make(a.x, b.x, c.x), k1 = gt_f3_make(a.y, b.y, c.y), k2 = gt_f3_make(a.z, b.z, c.z);
  const gt_f3 it = gt_add(gt_add(gt_scale(m[9], k0), gt_scale(m[10], k1)), gt_scale(m[11], k2));
  out[0] = k0.x; out[1] = k0.y; out[2] = k0.z;
  out[3] = k1.x; out[4] = k1.y; out[5] = k1.z;
  out[6] = k2.x; out[7] = k2.y; out[8] = k2.z;
  out[9] = -it.x; out[10] = -it.y; out[11] = -it.z;
}

/* Proxy of a Gaussian (src/main.ts:79-94): none unless opacity > alpha_min
 * (both doubles); scale factor s = sqrt(2 ln(opacity / alpha_min)); O2W =
 * T(mu) R(q) S(s * scale) with each f32 product of wgpu-matrix rounded per
 * operation; W2O its f32 inverse.  Returns 1 if the Gaussian has a proxy. */
GT_HD int gt_prep_proxy(const double* act, const float* pos, double alpha_min, float* o2w, float* w2o) {
  const double opacity = act[0];
  if (!(opacity > alpha_min)) return 0;
  const double s = GT_PREP_SQRT64(2 * gt_log64(opacity / alpha_min));
  float S[3];
  for (int j = 0; j < 3; ++j) S[j] = (float)(s * act[1 + j]);
  float R[9];
  gt_prep_quat_to_mat(act[5], act[6], act[7], act[4], R);
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) o2w[c * 3 + r] = (float)((double)R[c * 3 + r] * (double)S[c]);
  o2w[9] = pos[0]; o2w[10] = pos[1]; o2w[11] = pos[2];
  gt_prep_invert_affine(o2w, w2o);
  return 1;
}

/* World box of a proxy: the 12 icosahedron vertices through O2W in double,
 * padded by far more than the FP32 error of the object-space triangle test,
 * then rounded outward to f32 (lo[3], hi[3]).  Culling with these boxes
 * never drops a hit. */
GT_HD void gt_prep_box(const float* m, float* lo_out, float* hi_out) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int v = 0; v < 12; ++v) {
    const gt_f3 p = gt_ico_vertex(v);
    for (int r = 0; r < 3; ++r) {
      const double w = (double)m[r] * p.x + (double)m[3 + r] * p.y + (double)m[6 + r] * p.z + (double)m[9 + r];
      lo[r] = lo[r] < w ? lo[r] : w;
      hi[r] = hi[r] > w ? hi[r] : w;
    }
  }
  for (int r = 0; r < 3; ++r) {
    const double pad = 1e-3 * (hi[r] - lo[r]) + 2e-5 * (1.0 + fabs(lo[r]) + fabs(hi[r]));
    float l = (float)(lo[r] - pad), u = (float)(hi[r] + pad);
    if ((double)l > lo[r] - pad) l = GT_PREP_NEXTDOWN(l);
    if ((double)u < hi[r] + pad) u = GT_PREP_NEXTUP(u);
    lo_out[r] = l;
    hi_out[r] = u;
  }
}

/* Whether a proxy can be built on: its transform, inverse, world box and
 * box centre all finite in f32.  A record that gt_prep_valid accepts can
 * still fail here: a position near the f32 limit whose padded box (or box
 * centre) rounds to inf, a scale so large that scale x sqrt(2 ln(opacity /
 * alpha_min)) overflows, or one so small that the inverse's determinant
 * underflows to 0.  Such a Gaussian is an argument error, like those
 * gt_prep_valid rejects. */
GT_HD int gt_prep_proxy_finite(const float* o2w, const float* w2o, const float* lo, const float* hi) {
  int ok = 1;
  for (int j = 0; j < 12; ++j) ok &= (__builtin_isfinite(o2w[j]) && __builtin_isfinite(w2o[j])) ? 1 : 0;
  for (int j = 0; j < 3; ++j)
    ok &= (__builtin_isfinite(lo[j]) && __builtin_isfinite(hi[j]) && __builtin_isfinite(0.5f * (lo[j] + hi[j])))
              ? 1 : 0;
  return ok;
}

#ifdef __cplusplus
}
#endif

#endif /* GT_PREP_H */
