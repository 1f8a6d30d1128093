/* Screen rectangles of proxies: a conservative pixel bounding box of every
 * ray of a pinhole frame that can hit the proxy.  For a proxy in front of
 * the camera it holds every ray that the proxy test's bounding-sphere stage
 * accepts; for one wholly behind it is empty, which the exact test's
 * t > t_min >= 0 allows.  The packet traversal compares a leaf's rectangles
 * with its tile's pixel range on the scalar unit and skips the proxies no
 * lane can hit.
 *
 * Geometry.  The sphere stage (gt_forward.hip, test_proxy) accepts a ray
 * whose FP32 object-space line (oo, od) passes within
 * R = GT_ICO_CIRCUMRADIUS 1.001 + 1e-4 |oo|_1 of the object origin.  With the
 * rounding of that test and of the FP32 transforms bounded below, the real
 * world-space line o + s d of an accepted ray passes through the ellipsoid
 *   E = { x : |M x + t| <= R_eff } = { mu + A y : |y| <= R_eff },
 * M the 3x3 part and t the translation of the world-to-object matrix,
 * A = M^-1, mu = -A t.  In camera coordinates (a, b, c) = Binv (x - o), with
 * B = [X Y Z] the basis of the ray generation (gt_ray_dir_off), a pixel ray
 * has direction (nx ax, ny ay, 1) up to rounding, so a point of E seen by
 * the ray projects to a / c = nx ax, b / c = ny ay.  When E lies in front of
 * the camera (c > 0 on E), a / c over E ranges between the roots of
 *   (m_c^2 - S_cc) u^2 - 2 (m_a m_c - S_ac) u + (m_a^2 - S_aa) = 0,
 * the planes a - u c = 0 tangent to E (m the centre and S = L L^T the shape
 * of E in camera coordinates, L = Binv A R_eff).  The discriminant is
 * evaluated as m_c^2 S_aa - 2 m_a m_c S_ac + m_a^2 S_cc + (S_ac^2 - S_aa S_cc),
 * without the cancellation of the m_a^2 m_c^2 terms.  All of it runs in FP64
 * from the FP32 inputs.  A rectangle is "full" (never culls) when any
 * quantity is not finite, E reaches c <= 0, the error bounds are not small,
 * or the frame is wider or taller than GT_RECT_MAX_DIM pixels; "empty"
 * (always culls) when E lies wholly behind the camera and t_min >= 0: hits
 * have t > t_min, and no ray reaches E with s > 0.
 *
 * Error bounds (u = 2^-24, sums over components):
 *  - fl(oo) = M o + t + e1 with |e1_i| <= 4.1u (sum_j |M_ij o_j| + |t_i|);
 *    fl(od) = M d + e2 with |e2| <= eta |fl(od)|, eta = 6u |M|_F |A|_F
 *    (|M d| >= |d| / |A|_2);
 *  - the sphere stage's FP32 cross product, dots and radius move the
 *    accepted distance by at most 4u R + 4u |oo| + 1e-8 (|oo|_1 + 2);
 *  - so the real line's distance is at most
 *    R_eff = (R (1 + 4u) + 4u |oo| + |e1| (1 + eta) + |oo| eta) / (1 - eta);
 *  - a ray direction's camera coordinates are (nx ax, ny ay, 1) + Binv D
 *    with |D| <= 16u (2 |ax| |X| + 2 |ay| |Y| + |Z|) for an on-screen pixel
 *    (|nx|, |ny| <= 1 + 2 / W), and nx itself carries 4u (|nx| + 2) of
 *    rounding; both are added to the projected range in u units.
 * The pixel of a ray at continuous position P = (nx + 1) W / 2 satisfies
 * px + off = P with off in [0, 1] (antialiasing jitter, 0.5 without), so
 * px lies in [P_lo - 1, P_hi]; one more pixel on each side absorbs the FP64
 * rounding of the bounds themselves. */
#pragma once
#include <math.h>
#include <stdint.h>

#include "../../include/gt_fp32.h"

/* A rectangle is two words of two 15-bit fields, x in the low and y in the
 * high half: lo = x0 | y0 << 16, hi = (x1 | y1 << 16) | GT_RECT_GUARD, pixel
 * indices clamped to [0, 32767] (clamping never shrinks the set of
 * on-screen pixels a rectangle covers).  With the guard bits set in hi, a
 * tile [tx0, tx1] x [ty0, ty1] (lo_t = tx0 | ty0 << 16, hi_t likewise, plus
 * the guard) meets the rectangle iff
 *   ((hi_t - lo) & (hi - lo_t) & GT_RECT_GUARD) == GT_RECT_GUARD:
 * each 16-bit difference is 0x8000 + a - b with a, b < 0x8000, so no borrow
 * crosses a field and bit 15 of a field is set iff a >= b. */
#define GT_RECT_MAX_DIM 30000
#define GT_RECT_GUARD ((uint32_t)0x80008000u)
#define GT_RECT_FULL_LO ((uint32_t)0x00000000u)  /* x0 = y0 = 0 */
#define GT_RECT_FULL_HI ((uint32_t)0xffffffffu)  /* x1 = y1 = 32767, guard */
#define GT_RECT_EMPTY_LO ((uint32_t)0x7fff7fffu) /* x0 = y0 = 32767 */
#define GT_RECT_EMPTY_HI GT_RECT_GUARD           /* x1 = y1 = 0, guard */

/* inverse of a 3x3 matrix by its adjugate; 0 when singular or not finite */
GT_HD int gt_rect_inv3(const double m[3][3], double r[3][3]) {
  r[0][0] = m[1][1] * m[2][2] - m[1][2] * m[2][1];
  r[0][1] = m[0][2] * m[2][1] - m[0][1] * m[2][2];
  r[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  r[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  r[1][1] = m[0][0] * m[2][2] - m[0][2] * m[2][0];
  r[1][2] = m[0][2] * m[1][0] - m[0][0] * m[1][2];
  r[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  r[2][1] = m[0][1] * m[2][0] - m[0][0] * m[2][1];
  r[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  const double det = m[0][0] * r[0][0] + m[0][1] * r[1][0] + m[0][2] * r[2][0];
  if (!(fabs(det) > 1e-300) || !isfinite(det)) return 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r[i][j] /= det;
  return 1;
}

GT_HD double gt_rect_frob(const double m[3][3]) {
  double s = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s += m[i][j] * m[i][j];
  return sqrt(s);
}

/* range of a / c (axis 0) or b / c (axis 1) over the ellipsoid (centre m,
 * shape S) in front of the camera: lo <= hi */
GT_HD void gt_rect_range(const double m[3], const double S[3][3], int ax, double* lo, double* hi) {
  const double a2 = m[2] * m[2] - S[2][2];
  const double b2 = m[ax] * m[2] - S[ax][2];
  double disc = m[2] * m[2] * S[ax][ax] - 2.0 * m[ax] * m[2] * S[ax][2] + m[ax] * m[ax] * S[2][2] +
                (S[ax][2] * S[ax][2] - S[ax][ax] * S[2][2]);
  if (disc < 0.0) disc = 0.0;
  const double sq = sqrt(disc);
  *lo = (b2 - sq) / a2;
  *hi = (b2 + sq) / a2;
}

/* pixel range [p0, p1] of the rays whose projected coordinate lies in
 * [u_lo, u_hi] (u = n * scale, n = (p + off) / dim * 2 - 1), widened by err_u */
GT_HD void gt_rect_pixels(double u_lo, double u_hi, double err_u, double scale, uint32_t dim, double* p0, double* p1) {
  double n0 = (u_lo - err_u) / scale, n1 = (u_hi + err_u) / scale;
  if (n0 > n1) { const double t = n0; n0 = n1; n1 = t; }
  const double P0 = (n0 + 1.0) * 0.5 * (double)dim, P1 = (n1 + 1.0) * 0.5 * (double)dim;
  *p0 = floor(P0 - 1.0) - 1.0;
  *p1 = floor(P1) + 1.0;
}

GT_HD uint32_t gt_rect_pack(double a, double b) {
  a = a < 0.0 ? 0.0 : (a > 32767.0 ? 32767.0 : a);
  b = b < 0.0 ? 0.0 : (b > 32767.0 ? 32767.0 : b);
  return (uint32_t)a | ((uint32_t)b << 16);
}

/* does tile (lo_t, hi_t, packed like a rectangle) meet rectangle (lo, hi)? */
GT_HD uint32_t gt_rect_meets(uint32_t lo_t, uint32_t hi_t, uint32_t lo, uint32_t hi) {
  return (((hi_t - lo) & (hi - lo_t) & GT_RECT_GUARD) == GT_RECT_GUARD) ? 1u : 0u;
}

/* The rectangle of proxy record w (its 3x4 world-to-object matrix, columns
 * c0..c3) for the frame (to_world, ax, ay, width, height, t_min), packed as
 * out[0] = lo, out[1] = hi (inclusive pixel indices). */
GT_HD void gt_proxy_rect(const float* w, const float* to_world, float axf, float ayf, uint32_t width,
                         uint32_t height, float t_min, uint32_t* out) {
  const double u = 5.9604644775390625e-08; /* 2^-24 */
  out[0] = GT_RECT_FULL_LO;
  out[1] = GT_RECT_FULL_HI;
  if (width > GT_RECT_MAX_DIM || height > GT_RECT_MAX_DIM || width == 0 || height == 0) return;
  double M[3][3], A[3][3], B[3][3], Bi[3][3], t[3], o[3];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) M[i][j] = (double)w[3 * j + i];
    t[i] = (double)w[9 + i];
    o[i] = (double)to_world[12 + i];
    B[i][0] = -(double)to_world[i];
    B[i][1] = (double)to_world[4 + i];
    B[i][2] = -(double)to_world[8 + i];
  }
  if (!gt_rect_inv3(M, A) || !gt_rect_inv3(B, Bi)) return;
  const double ax = (double)axf, ay = (double)ayf;
  if (!(fabs(ax) > 0.0) || !(fabs(ay) > 0.0) || !isfinite(ax) || !isfinite(ay)) return;
  /* object-space origin and the sphere stage's radius */
  double oo[3], e1sq = 0.0, o1 = 0.0, no = 0.0;
  for (int i = 0; i < 3; ++i) {
    double s = t[i], sa = fabs(t[i]);
    for (int j = 0; j < 3; ++j) {
      s += M[i][j] * o[j];
      sa += fabs(M[i][j] * o[j]);
    }
    oo[i] = s;
    const double e1 = 4.1 * u * sa;
    e1sq += e1 * e1;
    o1 += fabs(s) + e1;
  }
  const double e1n = sqrt(e1sq);
  no = sqrt(oo[0] * oo[0] + oo[1] * oo[1] + oo[2] * oo[2]) + e1n; /* bounds |fl(oo)| */
  const double eta = 6.0 * u * gt_rect_frob(M) * gt_rect_frob(A);
  if (!(eta < 0.25)) return;
  const double R = 1.0011 * (double)GT_ICO_CIRCUMRADIUS + 1.0001e-4 * o1;
  const double dist = R * (1.0 + 4.0 * u) + 4.0 * u * no + 1e-8 * (o1 + 2.0);
  const double Reff = (dist + e1n * (1.0 + eta) + no * eta) / (1.0 - eta) * (1.0 + 1e-9);
  /* centre and shape in camera coordinates */
  double mu[3], m[3], L[3][3], S[3][3];
  for (int i = 0; i < 3; ++i) mu[i] = -(A[i][0] * t[0] + A[i][1] * t[1] + A[i][2] * t[2]);
  for (int i = 0; i < 3; ++i) {
    m[i] = Bi[i][0] * (mu[0] - o[0]) + Bi[i][1] * (mu[1] - o[1]) + Bi[i][2] * (mu[2] - o[2]);
    for (int j = 0; j < 3; ++j) L[i][j] = (Bi[i][0] * A[0][j] + Bi[i][1] * A[1][j] + Bi[i][2] * A[2][j]) * Reff;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) S[i][j] = L[i][0] * L[j][0] + L[i][1] * L[j][1] + L[i][2] * L[j][2];
  const double sc = sqrt(S[2][2]);
  if (!isfinite(m[2]) || !isfinite(sc)) return;
  if (m[2] + sc * (1.0 + 1e-9) < 0.0) { /* wholly behind the camera */
    if (!(t_min >= 0.0f)) return;
    out[0] = GT_RECT_EMPTY_LO;
    out[1] = GT_RECT_EMPTY_HI;
    return;
  }
  if (!(m[2] > sc * (1.0 + 1e-6) + 1e-300)) return; /* reaches the camera plane */
  double ulo, uhi, vlo, vhi;
  gt_rect_range(m, S, 0, &ulo, &uhi);
  gt_rect_range(m, S, 1, &vlo, &vhi);
  if (!isfinite(ulo) || !isfinite(uhi) || !isfinite(vlo) || !isfinite(vhi)) return;
  /* ray-generation rounding, in projected units */
  double nX = 0.0, nY = 0.0, nZ = 0.0;
  for (int i = 0; i < 3; ++i) {
    nX += B[i][0] * B[i][0];
    nY += B[i][1] * B[i][1];
    nZ += B[i][2] * B[i][2];
  }
  const double beta = gt_rect_frob(Bi) * 16.0 * u * (2.0 * fabs(ax) * sqrt(nX) + 2.0 * fabs(ay) * sqrt(nY) + sqrt(nZ));
  if (!(beta < 0.25)) return;
  const double gx = 1.0 + 2.0 / (double)width, gy = 1.0 + 2.0 / (double)height;
  const double err_x = beta * (1.0 + fabs(ax) * gx) / (1.0 - beta) + fabs(ax) * 4.0 * u * (gx + 2.0);
  const double err_y = beta * (1.0 + fabs(ay) * gy) / (1.0 - beta) + fabs(ay) * 4.0 * u * (gy + 2.0);
  double x0, x1, y0, y1;
  gt_rect_pixels(ulo - 1e-9 * (fabs(ulo) + 1.0), uhi + 1e-9 * (fabs(uhi) + 1.0), err_x, ax, width, &x0, &x1);
  gt_rect_pixels(vlo - 1e-9 * (fabs(vlo) + 1.0), vhi + 1e-9 * (fabs(vhi) + 1.0), err_y, ay, height, &y0, &y1);
  if (!isfinite(x0) || !isfinite(x1) || !isfinite(y0) || !isfinite(y1)) return;
  out[0] = gt_rect_pack(x0, y0);
  out[1] = gt_rect_pack(x1, y1) | GT_RECT_GUARD;
}
