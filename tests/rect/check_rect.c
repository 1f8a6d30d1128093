/* CPU check of the proxy screen rectangles (csrc/gt_rect.h) against the
 * bounding-sphere stage of the proxy test, the only stage the rectangles
 * stand in for.  For random proxies and pinhole frames, every pixel whose
 * ray (gt_ray_dir_off, with random antialiasing offsets or the pixel centre)
 * the FP32 sphere stage accepts must lie inside the proxy's rectangle.  The
 * sphere stage is written here as the kernel evaluates it (test_proxy in
 * gt_forward.hip): oo = gt_xform_point, od = gt_xform_dir, reject iff
 * dot(cr, cr) > R R dot(od, od) with cr = oo x od and
 * R = GT_ICO_CIRCUMRADIUS 1.001f + 1e-4f |oo|_1, all in FP32 without
 * contraction.  The sphere stage tests the ray's whole line, so for a
 * proxy wholly behind the camera (an "empty" rectangle) it may accept; such
 * a rectangle stands in for the exact test instead (hits have t > t_min >=
 * 0), and the check runs the exact triangle test (gt_tri_hit over the 20
 * faces, either facing) on the pixels the sphere stage accepts and requires
 * no hit.
 *
 * It also checks gt_rect_meets against plain comparisons.  Prints one line
 * "cases C accepted A misses M inside I shrunk S full F empty E behind B
 * swar_bad W" (W: disagreements of gt_rect_meets, must be 0): A accepted (pixel, proxy) pairs, M of them outside the
 * rectangle (must be 0; for empty rectangles: exact hits, must be 0), B
 * the accepted pairs of empty rectangles checked by the exact test, I
 * rectangle pixels in all (the culling's tightness), S accepted pairs
 * outside the rectangles shrunk by 3 pixels on every side (must be > 0:
 * the check sees the boundary), F and E the rectangles that are full
 * (never cull) and empty. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "gt_fp32.h"
#include "gt_rect.h"

static uint64_t g_rng = 0x9e3779b97f4a7c15ull;
static double urand(void) { /* splitmix64 */
  uint64_t z = (g_rng += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
static double nrand(void) { return sqrt(-2.0 * log(urand() + 1e-300)) * cos(6.283185307179586 * urand()); }

static void unit(double* v) {
  double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int i = 0; i < 3; ++i) v[i] /= n;
}
static void cross3(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* camera: to_world with columns -X, Y, -Z, origin (gt_ray_dir_off's basis) */
static void make_camera(float* tw, const double* o, double skew) {
  double z[3] = {nrand(), nrand(), nrand()}, up[3] = {nrand(), nrand(), nrand()}, x[3], y[3];
  unit(z);
  cross3(up, z, x);
  unit(x);
  cross3(z, x, y);
  for (int i = 0; i < 3; ++i) y[i] += skew * x[i]; /* a non-orthogonal basis when skew != 0 */
  for (int i = 0; i < 16; ++i) tw[i] = 0.0f;
  for (int i = 0; i < 3; ++i) {
    tw[i] = (float)-x[i];
    tw[4 + i] = (float)y[i];
    tw[8 + i] = (float)-z[i];
    tw[12 + i] = (float)o[i];
  }
  tw[15] = 1.0f;
}

/* a proxy: world-to-object = diag(1 / s) R^T (x - mu), as FP32 columns */
static void make_proxy(float* w, const double* mu, const double* s) {
  double q[4] = {nrand(), nrand(), nrand(), nrand()}, n = 0.0;
  for (int i = 0; i < 4; ++i) n += q[i] * q[i];
  n = sqrt(n);
  for (int i = 0; i < 4; ++i) q[i] /= n;
  const double a = q[0], b = q[1], c = q[2], d = q[3];
  const double R[3][3] = {{1 - 2 * (c * c + d * d), 2 * (b * c - a * d), 2 * (b * d + a * c)},
                          {2 * (b * c + a * d), 1 - 2 * (b * b + d * d), 2 * (c * d - a * b)},
                          {2 * (b * d - a * c), 2 * (c * d + a * b), 1 - 2 * (b * b + c * c)}};
  double M[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M[i][j] = R[j][i] / s[i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) w[3 * j + i] = (float)M[i][j];
  for (int i = 0; i < 3; ++i) {
    double t = 0.0;
    for (int j = 0; j < 3; ++j) t -= (double)w[3 * j + i] * mu[j];
    w[9 + i] = (float)t;
  }
}

static int exact_hit(const float* w, gt_f3 o, gt_f3 d, float t_min) {
  const gt_f3 oo = gt_xform_point(w, o);
  const gt_f3 od = gt_xform_dir(w, d);
  for (int f = 0; f < GT_NUM_FACES; ++f) {
    float t;
    if (gt_tri_hit(oo, od, t_min, INFINITY, gt_ico_face(f), &t)) return 1;
  }
  return 0;
}

static int sphere_accepts(const float* w, gt_f3 o, gt_f3 d) {
  const gt_f3 oo = gt_xform_point(w, o);
  const gt_f3 od = gt_xform_dir(w, d);
  const gt_f3 cr = gt_cross(oo, od);
  const float o1 = fabsf(oo.x) + fabsf(oo.y) + fabsf(oo.z);
  const float R = GT_ICO_CIRCUMRADIUS * 1.001f + 1e-4f * o1;
  return !(gt_dot(cr, cr) > R * R * gt_dot(od, od));
}

/* gt_rect_meets against plain comparisons: fields at 0, 1, 32766, 32767
 * and random 15-bit values; returns the number of disagreements */
static long check_meets(long n) {
  const uint32_t edge[6] = {0u, 1u, 2u, 16384u, 32766u, 32767u};
  long bad = 0;
  for (long i = 0; i < n; ++i) {
    uint32_t v[8];
    for (int j = 0; j < 8; ++j) v[j] = (i & 1) ? edge[(int)(urand() * 6)] : (uint32_t)(urand() * 32768.0);
    const uint32_t x0 = v[0], y0 = v[1], x1 = v[2], y1 = v[3], tx0 = v[4], ty0 = v[5], tx1 = v[6], ty1 = v[7];
    const uint32_t want = (x0 <= tx1 && tx0 <= x1 && y0 <= ty1 && ty0 <= y1) ? 1u : 0u;
    const uint32_t got = gt_rect_meets(tx0 | ty0 << 16, (tx1 | ty1 << 16) | GT_RECT_GUARD, x0 | y0 << 16,
                                       (x1 | y1 << 16) | GT_RECT_GUARD);
    bad += got != want;
  }
  return bad;
}

int main(int argc, char** argv) {
  const int n_cams = argc > 1 ? atoi(argv[1]) : 24;
  const int n_prox = argc > 2 ? atoi(argv[2]) : 400;
  const long swar_bad = check_meets(2000000);
  long cases = 0, accepted = 0, misses = 0, inside = 0, shrunk = 0, full = 0, empty = 0, behind = 0;
  for (int c = 0; c < n_cams; ++c) {
    const uint32_t W = 48 + (uint32_t)(urand() * 64), H = 32 + (uint32_t)(urand() * 48);
    const double hfov = (c % 4 == 0) ? 1.0 + 4.0 * urand() : 20.0 + 100.0 * urand();
    const float ax = (float)tan(0.5 * hfov * M_PI / 180.0);
    const float ay = (float)((double)ax * ((double)H / (double)W));
    const double far_scene = (c % 3 == 0) ? 1e4 : 1.0; /* coordinates far from the world origin */
    const double o[3] = {far_scene + 5 * nrand(), 5 * nrand(), 5 * nrand()};
    float tw[16];
    make_camera(tw, o, (c % 5 == 4) ? 0.3 * nrand() : 0.0);
    const int aa = c % 2;
    const float t_min = (c % 6 == 5) ? 0.0f : 1e-3f;
    for (int p = 0; p < n_prox; ++p) {
      /* a centre near the view (in front, behind, or around the camera) */
      double dir[3] = {-(double)tw[8] + 0.8 * nrand(), -(double)tw[9] + 0.8 * nrand(), -(double)tw[10] + 0.8 * nrand()};
      if (urand() < 0.1) for (int i = 0; i < 3; ++i) dir[i] = -dir[i];
      unit(dir);
      const double dist = exp(log(0.05) + urand() * log(2000.0 / 0.05));
      const double mu[3] = {o[0] + dist * dir[0], o[1] + dist * dir[1], o[2] + dist * dir[2]};
      const double base = dist * (hfov * M_PI / 180.0) * exp(-4.0 + 5.0 * urand());
      const double s[3] = {base * exp(2.0 * nrand()), base * exp(2.0 * nrand()), base * exp(2.0 * nrand())};
      float w[12];
      make_proxy(w, mu, s);
      uint32_t r[2];
      gt_proxy_rect(w, tw, ax, ay, W, H, t_min, r);
      const int is_empty = r[0] == GT_RECT_EMPTY_LO && r[1] == GT_RECT_EMPTY_HI;
      full += (r[0] == GT_RECT_FULL_LO && r[1] == GT_RECT_FULL_HI);
      empty += is_empty;
      const int x0 = (int)(r[0] & 0xffffu), y0 = (int)(r[0] >> 16);
      const int x1 = (int)(r[1] & 0x7fffu), y1 = (int)((r[1] >> 16) & 0x7fffu);
      const gt_f3 og = gt_f3_make(tw[12], tw[13], tw[14]);
      for (uint32_t py = 0; py < H; ++py)
        for (uint32_t px = 0; px < W; ++px) {
          float offx = 0.5f, offy = 0.5f;
          if (aa) {
            offx = (float)(urand() * 0.99999994);
            offy = (float)(urand() * 0.99999994);
            if (urand() < 0.05) offx = 0.0f;
            if (urand() < 0.05) offy = 0.99999994f;
          }
          const gt_f3 d = gt_ray_dir_off(tw, ax, ay, px, py, W, H, offx, offy);
          ++cases;
          const uint32_t lo_t = px | py << 16, hi_t = (px | py << 16) | GT_RECT_GUARD;
          const int in = (int)gt_rect_meets(lo_t, hi_t, r[0], r[1]);
          inside += in;
          if (!sphere_accepts(w, og, d)) continue;
          if (is_empty) {
            ++behind;
            if (exact_hit(w, og, d, t_min)) {
              ++misses;
              if (misses <= 5) fprintf(stderr, "hit behind: cam %d proxy %d pixel (%u, %u)\n", c, p, px, py);
            }
            continue;
          }
          ++accepted;
          if (!in) {
            ++misses;
            if (misses <= 5)
              fprintf(stderr, "miss: cam %d proxy %d pixel (%u, %u) rect [%d, %d] x [%d, %d]\n", c, p, px, py, x0, x1,
                      y0, y1);
          }
          if (!((int)px >= x0 + 3 && (int)px <= x1 - 3 && (int)py >= y0 + 3 && (int)py <= y1 - 3)) ++shrunk;
        }
    }
  }
  printf("cases %ld accepted %ld misses %ld inside %ld shrunk %ld full %ld empty %ld behind %ld swar_bad %ld\n", cases,
         accepted, misses, inside, shrunk, full, empty, behind, swar_bad);
  return (misses == 0 && swar_bad == 0) ? 0 : 1;
}
