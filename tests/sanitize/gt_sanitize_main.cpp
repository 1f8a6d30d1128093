/*
 * gt_sanitize_main.cpp - a driver of the library's host code (the PLY
 * reader, the host scene builder, the PNG writer) and of the oracle, built
 * by tests/test_sanitize_host.py with g++ / gcc under AddressSanitizer +
 * UndefinedBehaviorSanitizer (or ThreadSanitizer).  Test code only: it
 * links no HIP and runs without a GPU.
 *
 *   ply FILE...                     gt_ply_read on each file; a file that
 *                                   parses runs the host builder (and the
 *                                   oracle's activations and proxies) on
 *                                   its records.  Prints "FILE rc n deg".
 *   scene N SEED CLUSTERS ALPHA R   a generated scene through the host
 *                                   builder with the tree checks; R = 1
 *                                   also renders it with the oracle (BVH
 *                                   and brute force, 2 threads), checks the
 *                                   two agree bit for bit and against the
 *                                   FP64 re-evaluation, and writes a PNG.
 *   png DIR                         gt_write_png edge cases.
 *
 * Exit status 0 when every check holds, 1 otherwise; a sanitizer report
 * aborts the process (-fno-sanitize-recover).
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gt.h"
#include "../../oracle/gt_oracle.h"
#include "../../webgpu-gaussiantracer_amd/csrc/gt_host_build.h"
#include "../../webgpu-gaussiantracer_amd/csrc/gt_internal.h"
#include "../../webgpu-gaussiantracer_amd/csrc/gt_prep.h"

static int g_fail = 0;

#define CHECK(cond, ...)                          \
  do {                                            \
    if (!(cond)) {                                \
      std::fprintf(stderr, "CHECK failed: " __VA_ARGS__); \
      std::fprintf(stderr, "  [%s:%d]\n", __FILE__, __LINE__); \
      ++g_fail;                                   \
    }                                             \
  } while (0)

static bool same_bits(const void* a, const void* b, size_t bytes) { return bytes == 0 || std::memcmp(a, b, bytes) == 0; }

static bool box_in(const float* lo, const float* hi, const float* LO, const float* HI) {
  for (int j = 0; j < 3; ++j)
    if (!(lo[j] >= LO[j] && hi[j] <= HI[j])) return false;
  return true;
}

/* Structure of a built tree: every node reached once from the root, filled
 * slots first, every leaf position in exactly one leaf, each slot box
 * holding what lies below it, the stack need and depth as recorded, the
 * proxy records in leaf order. */
static void check_tree(const HostScene& h, const HostTree& t) {
  const size_t np = h.proxy_ids.size();
  if (np == 0) {
    CHECK(t.root == GT_NODE_EMPTY && t.nodes.empty(), "empty scene has a tree\n");
    return;
  }
  CHECK(t.root < t.nodes.size(), "root %u of %zu nodes\n", t.root, t.nodes.size());
  CHECK(t.order.size() == np && t.prox.size() == np * 16, "order/prox sizes\n");
  if (g_fail) return;
  std::vector<uint8_t> seen(t.nodes.size(), 0);
  std::vector<uint32_t> cover(np, 0);
  std::vector<uint32_t> need(t.nodes.size(), 0);
  /* depth-first order, children after parents; needs filled in reverse */
  std::vector<uint32_t> order_nodes, level(t.nodes.size(), 0);
  std::vector<uint32_t> stack{t.root};
  level[t.root] = 1;
  while (!stack.empty()) {
    const uint32_t ni = stack.back();
    stack.pop_back();
    if (seen[ni]) { CHECK(false, "node %u reached twice\n", ni); return; }
    seen[ni] = 1;
    order_nodes.push_back(ni);
    const gt_node4& nd = t.nodes[ni];
    bool ended = false;
    for (int i = 0; i < 4; ++i) {
      const uint32_t c = nd.child[i];
      const float lo[3] = {nd.lo_x[i], nd.lo_y[i], nd.lo_z[i]};
      const float hi[3] = {nd.hi_x[i], nd.hi_y[i], nd.hi_z[i]};
      if (c == GT_NODE_EMPTY) {
        ended = true;
        CHECK(lo[0] == INFINITY && hi[0] == -INFINITY, "empty slot %d of node %u has a box\n", i, ni);
        continue;
      }
      CHECK(!ended, "node %u: filled slot %d after an empty one\n", ni, i);
      if (c & GT_LEAF_BIT) {
        const uint32_t first = c & GT_LEAF_FIRST_MASK;
        const uint32_t count = ((c >> GT_LEAF_COUNT_SHIFT) & GT_LEAF_COUNT_MASK) + 1u;
        CHECK((size_t)first + count <= np, "leaf [%u, +%u) beyond %zu proxies\n", first, count, np);
        if ((size_t)first + count > np) continue;
        for (uint32_t j = first; j < first + count; ++j) {
          ++cover[j];
          float plo[3], phi[3];
          gt_prep_box(&h.o2w[(size_t)12 * t.order[j]], plo, phi);
          CHECK(box_in(plo, phi, lo, hi), "proxy at leaf position %u outside its slot box\n", j);
        }
      } else {
        CHECK(c < t.nodes.size(), "child %u of node %u out of range\n", c, ni);
        if (c >= t.nodes.size()) continue;
        const gt_node4& cn = t.nodes[c];
        for (int k = 0; k < 4; ++k) {
          if (cn.child[k] == GT_NODE_EMPTY) continue;
          const float clo[3] = {cn.lo_x[k], cn.lo_y[k], cn.lo_z[k]};
          const float chi[3] = {cn.hi_x[k], cn.hi_y[k], cn.hi_z[k]};
          CHECK(box_in(clo, chi, lo, hi), "node %u slot %d not inside its parent's slot\n", c, k);
        }
        level[c] = level[ni] + 1;
        stack.push_back(c);
      }
    }
  }
  for (size_t i = 0; i < t.nodes.size(); ++i) CHECK(seen[i], "node %zu unreachable\n", i);
  for (size_t j = 0; j < np; ++j) CHECK(cover[j] == 1, "leaf position %zu covered %u times\n", j, cover[j]);
  std::vector<uint8_t> perm(np, 0);
  for (size_t j = 0; j < np; ++j) {
    const uint32_t p = t.order[j];
    CHECK(p < np && !perm[p], "order is not a permutation at %zu\n", j);
    if (p >= np || perm[p]) return;
    perm[p] = 1;
    CHECK(same_bits(&t.prox[16 * j], &h.w2o[(size_t)12 * p], 48), "proxy record %zu: matrix\n", j);
    CHECK(same_bits(&t.prox[16 * j + 12], &h.proxy_ids[p], 4), "proxy record %zu: id\n", j);
  }
  uint32_t depth = 0;
  for (auto it = order_nodes.rbegin(); it != order_nodes.rend(); ++it) {
    const gt_node4& nd = t.nodes[*it];
    uint32_t m = 0, cneed = 0;
    for (int i = 0; i < 4; ++i) {
      const uint32_t c = nd.child[i];
      if (c == GT_NODE_EMPTY) continue;
      ++m;
      if (!(c & GT_LEAF_BIT)) cneed = std::max(cneed, need[c]);
    }
    need[*it] = std::max(m, (m - 1) + cneed);
    depth = std::max(depth, level[*it]);
  }
  CHECK(need[t.root] == t.need, "stack need %u, recorded %u\n", need[t.root], t.need);
  CHECK(t.need <= GT_STACK_MAX, "stack need %u above %u\n", t.need, (unsigned)GT_STACK_MAX);
  CHECK(depth == t.depth, "depth %u, recorded %u\n", depth, t.depth);
}

/* The host builder on n raw records; with the oracle's activations and
 * proxies compared bit for bit.  Returns the status of the activations and
 * proxies (records that describe no finite Gaussian are GT_E_INVALID). */
static int build_and_check(const float* raw, uint32_t n, double alpha_min, HostScene& h, HostTree& t) {
  h = HostScene{};
  h.n = n;
  h.raw.assign(raw, raw + (size_t)n * GT_RAW_FLOATS);
  int rc = gt_host_activate_raw(h);
  if (rc != GT_OK) return rc;
  if ((rc = gt_host_build_proxies(h, alpha_min)) != GT_OK) return rc;
  const int trc = gt_host_build_tree(h, t);
  CHECK(trc == GT_OK, "tree build failed: %s\n", gt_last_error());
  if (trc == GT_OK) check_tree(h, t);
  std::vector<float> geom, sh;
  gt_host_shading_records(h, geom, sh);
  CHECK(geom.size() == (size_t)n * 12 && sh.size() == (size_t)n * 48, "shading record sizes\n");
  for (uint32_t i = 0; i < n; ++i)
    CHECK(geom[(size_t)12 * i + 3] == h.packed[(size_t)80 * i + 12], "record %u opacity\n", i);

  std::vector<float> packed((size_t)n * GTO_PACKED_FLOATS + 1);
  std::vector<double> act((size_t)n * 8 + 1);
  gto_activate(raw, n, packed.data(), act.data());
  CHECK(same_bits(packed.data(), h.packed.data(), h.packed.size() * 4), "oracle packed records differ\n");
  CHECK(same_bits(act.data(), h.act.data(), h.act.size() * 8), "oracle activations differ\n");
  std::vector<uint32_t> ids((size_t)n + 1);
  std::vector<float> o2w((size_t)n * 12 + 1), w2o((size_t)n * 12 + 1);
  const uint32_t np = gto_build_proxies(raw, act.data(), n, alpha_min, ids.data(), o2w.data(), w2o.data());
  CHECK(np == h.proxy_ids.size(), "oracle proxies %u, host %zu\n", np, h.proxy_ids.size());
  if (np == h.proxy_ids.size()) {
    CHECK(same_bits(ids.data(), h.proxy_ids.data(), (size_t)np * 4), "proxy ids differ\n");
    CHECK(same_bits(o2w.data(), h.o2w.data(), (size_t)np * 48), "o2w differ\n");
    CHECK(same_bits(w2o.data(), h.w2o.data(), (size_t)np * 48), "w2o differ\n");
  }
  return GT_OK;
}

static int cmd_ply(int argc, char** argv) {
  for (int a = 0; a < argc; ++a) {
    float* raw = nullptr;
    uint32_t n = 0, deg = 0;
    const int rc = gt_ply_read(argv[a], &raw, &n, &deg);
    int brc = 0;
    if (rc == GT_OK) {
      CHECK(deg <= 3, "%s: degree %u\n", argv[a], deg);
      HostScene h;
      HostTree t;
      brc = build_and_check(raw, n, 0.01, h, t);
      gt_free(raw);
    } else {
      CHECK(raw == nullptr, "%s: records returned with an error\n", argv[a]);
      CHECK(std::strlen(gt_last_error()) > 0, "%s: error without a message\n", argv[a]);
    }
    std::printf("%s %d %u %u %d\n", argv[a], rc, rc == GT_OK ? n : 0u, rc == GT_OK ? deg : 0u, brc);
  }
  return 0;
}

static int cmd_scene(int argc, char** argv) {
  if (argc < 5) return 2;
  gt_gen_params g{};
  g.n = (uint32_t)std::strtoul(argv[0], nullptr, 10);
  g.seed = std::strtoull(argv[1], nullptr, 10);
  g.n_clusters = (uint32_t)std::strtoul(argv[2], nullptr, 10);
  const double alpha_min = std::strtod(argv[3], nullptr);
  const bool render = std::atoi(argv[4]) != 0;
  g.box_half = 2.0f;
  g.cluster_sigma = g.n_clusters ? 0.4f : 0.0f;
  g.log_scale_mean = -3.0f;
  g.log_scale_std = 0.5f;
  g.logit_mean = 0.0f;
  g.logit_std = 1.5f;
  g.dc_std = 0.5f;
  g.rest_std = 0.05f;
  std::vector<float> raw((size_t)g.n * GT_RAW_FLOATS + 1);
  CHECK(gt_generate_scene(&g, raw.data()) == GT_OK, "generator: %s\n", gt_last_error());
  HostScene h;
  HostTree t;
  CHECK(build_and_check(raw.data(), g.n, alpha_min, h, t) == GT_OK, "activation: %s\n", gt_last_error());
  std::printf("scene n=%u proxies=%zu nodes=%zu need=%u depth=%u\n", g.n, h.proxy_ids.size(), t.nodes.size(),
              t.need, t.depth);
  if (!render || g_fail) return 0;

  gt_uniforms u;
  float m[16];
  const float pos[3] = {0.3f, 0.8f, 6.0f}, target[3] = {0.0f, 0.0f, 0.0f};
  gt_camera_look_at(pos, target, m);
  const uint32_t W = 40, H = 30, K = 8, MR = 96;
  gt_uniforms_init(&u, m, 50.0, W, H, 0.0f, 1e30f, 0.001f, (float)alpha_min, K, 3, 0, 0);
  gto_params p{};
  std::memcpy(p.to_world, u.to_world, sizeof(p.to_world));
  p.aspect[0] = u.camera_aspect[0];
  p.aspect[1] = u.camera_aspect[1];
  p.t_min = u.t_min; p.t_max = u.t_max; p.T_min = u.T_min; p.alpha_min = u.alpha_min;
  p.k = u.k; p.sh_degree_max = u.sh_degree_max; p.accumulated_frames = 0;
  p.early_stop = u.early_stop; p.debug_mode = 0; p.width = W; p.height = H;
  gto_scene s{};
  s.n = h.n;
  s.packed = h.packed.data();
  s.n_proxies = (uint32_t)h.proxy_ids.size();
  s.proxy_ids = h.proxy_ids.data();
  s.o2w = h.o2w.data();
  s.w2o = h.w2o.data();
  const size_t R = (size_t)W * H;
  struct Out {
    std::vector<float> rgba, T;
    std::vector<uint32_t> rounds, hits, acc;
    uint64_t st[GTO_NSTATS];
  } o[2];
  for (int b = 0; b < 2; ++b) {
    o[b].rgba.assign(R * 4, 0.0f);
    o[b].T.assign(R, 0.0f);
    o[b].rounds.assign(R, 0u);
    o[b].hits.assign(R * MR * K, 0u);
    o[b].acc.assign(R * MR, 0u);
    std::memset(o[b].st, 0, sizeof(o[b].st));
    void* bvh = b ? gto_bvh_build(&s) : nullptr;
    s.bvh = bvh;
    const int rc = gto_render(&s, &p, b, 2, nullptr, 0, o[b].rgba.data(), o[b].T.data(), o[b].rounds.data(), MR,
                              o[b].hits.data(), o[b].acc.data(), o[b].st);
    CHECK(rc == 0, "oracle render (bvh %d) returned %d\n", b, rc);
    s.bvh = nullptr;
    gto_bvh_free(bvh);
  }
  CHECK(o[0].st[7] == 0, "hit lists overflowed %llu times\n", (unsigned long long)o[0].st[7]);
  CHECK(o[0].rgba == o[1].rgba && o[0].T == o[1].T, "BVH and brute-force pixels differ\n");
  CHECK(o[0].rounds == o[1].rounds && o[0].hits == o[1].hits && o[0].acc == o[1].acc,
        "BVH and brute-force hit lists differ\n");
  CHECK((o[0].st[3] > 0) == !h.proxy_ids.empty(), "composited %llu hits with %zu proxies\n",
        (unsigned long long)o[0].st[3], h.proxy_ids.size());
  if (h.proxy_ids.empty())
    for (size_t r = 0; r < R; ++r) CHECK(o[0].T[r] == 1.0f, "ray %zu of an empty scene has T %g\n", r, o[0].T[r]);
  std::vector<double> rgb(R * 3), T64(R);
  gto_reeval_f64(&s, &p, nullptr, (uint32_t)R, o[1].rounds.data(), MR, o[1].hits.data(), o[1].acc.data(),
                 rgb.data(), T64.data());
  double worst = 0.0;
  for (size_t r = 0; r < R; ++r)
    for (int c = 0; c < 3; ++c) worst = std::max(worst, std::fabs(rgb[3 * r + c] - (double)o[1].rgba[4 * r + c]));
  CHECK(worst <= 1e-4, "pixels %.3g from the FP64 re-evaluation\n", worst);
  std::vector<uint8_t> img(R * 4);
  for (size_t i = 0; i < R * 4; ++i) {
    const float v = std::fmin(std::fmax(o[1].rgba[i], 0.0f), 1.0f);
    img[i] = (uint8_t)std::lround(v * 255.0f);
  }
  if (argc > 5) CHECK(gt_write_png(argv[5], img.data(), W, H) == GT_OK, "png: %s\n", gt_last_error());
  std::printf("render composited=%llu worst=%.3g\n", (unsigned long long)o[1].st[3], worst);
  return 0;
}

static int cmd_png(const char* dir) {
  const std::string d(dir);
  std::vector<uint8_t> px(4096 * 4 * 3, 0x7f);
  CHECK(gt_write_png(nullptr, px.data(), 1, 1) == GT_E_INVALID, "NULL path accepted\n");
  CHECK(gt_write_png((d + "/a.png").c_str(), nullptr, 1, 1) == GT_E_INVALID, "NULL pixels accepted\n");
  CHECK(gt_write_png((d + "/b.png").c_str(), px.data(), 0, 1) == GT_E_INVALID, "width 0 accepted\n");
  CHECK(gt_write_png((d + "/c.png").c_str(), px.data(), 1, 0) == GT_E_INVALID, "height 0 accepted\n");
  CHECK(gt_write_png((d + "/no/such/dir/d.png").c_str(), px.data(), 1, 1) == GT_E_IO, "bad path not GT_E_IO\n");
  CHECK(gt_write_png((d + "/e.png").c_str(), px.data(), 1, 1) == GT_OK, "1x1: %s\n", gt_last_error());
  CHECK(gt_write_png((d + "/f.png").c_str(), px.data(), 4096, 3) == GT_OK, "4096x3: %s\n", gt_last_error());
  CHECK(gt_write_png((d + "/g.png").c_str(), px.data(), 3, 4096) == GT_OK, "3x4096: %s\n", gt_last_error());
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string cmd = argv[1];
  int rc = 2;
  if (cmd == "ply") rc = cmd_ply(argc - 2, argv + 2);
  else if (cmd == "scene") rc = cmd_scene(argc - 2, argv + 2);
  else if (cmd == "png" && argc > 2) rc = cmd_png(argv[2]);
  if (rc != 0) return rc;
  std::printf("failures %d\n", g_fail);
  return g_fail ? 1 : 0;
}
