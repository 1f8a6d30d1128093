/*
 * gt_host_build.cpp - the host scene builder: activations, proxies, the
 * binned-SAH tree and its 4-wide collapse, shading records (plain C++, no
 * HIP; gt_scene.cpp uploads the result).
 *
 * Replaces, for the Gaussian path:
 *   PackedGaussians activations + packing   src/GS_ply.ts:187-252
 *   gaussianSceneToBvh (filter, proxies)    src/main.ts:30-111
 *   TLAS build                              webrtx/src/wasm_bvh_builder.ts:260-377,
 *                                           webrtx/bvh/src/lib.rs:253-347, 482-603
 */
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "../../include/gt.h"
#include "../../include/gt_fp32.h"
#include "gt_host_build.h"
#include "gt_internal.h"
#include "gt_prep.h"

/* ---------------- activations (src/GS_ply.ts:187-196, 225-235) ---------------- */

/* Activations of every raw record; returns GT_OK, or GT_E_INVALID naming
 * the first record that gt_prep_valid rejects. */
int gt_host_activate_raw(HostScene& h) {
  h.act.assign((size_t)h.n * 8, 0.0);
  h.packed.assign((size_t)h.n * 80, 0.0f);
  const unsigned chunks = gt_host_threads(h.n);
  std::vector<uint32_t> bad(chunks, 0xffffffffu);
  gt_parallel_chunks(h.n, chunks, [&](size_t i0, size_t i1, unsigned c) {
  for (size_t i = i0; i < i1; ++i) {
    float g[12], sh[48];
    gt_prep_activate(&h.raw[(size_t)i * GT_RAW_FLOATS], &h.act[(size_t)i * 8], g, sh);
    if (bad[c] == 0xffffffffu && !gt_prep_valid(&h.raw[(size_t)i * GT_RAW_FLOATS], &h.act[(size_t)i * 8]))
      bad[c] = (uint32_t)i;
    float* o = &h.packed[(size_t)i * 80];
    o[0] = g[0]; o[1] = g[1]; o[2] = g[2];
    o[4] = g[4]; o[5] = g[5]; o[6] = g[6];
    o[8] = g[8]; o[9] = g[9]; o[10] = g[10]; o[11] = g[11];
    o[12] = g[3];
    for (int j = 0; j < 16; ++j)
      for (int cc = 0; cc < 3; ++cc) o[16 + 4 * j + cc] = sh[3 * j + cc];
  }
  });
  for (unsigned c = 0; c < chunks; ++c)
    if (bad[c] != 0xffffffffu) {
      gt_set_error("Gaussian %u: non-finite parameter, zero quaternion or a scale that is 0 or inf in f32", bad[c]);
      return GT_E_INVALID;
    }
  return GT_OK;
}

/* The activated values of packed 320-B records in double (the JS doubles
 * are not part of the buffer).  A record must be finite in every field the
 * path reads (position, scale, quaternion, opacity, the 16 SH vec3), with
 * scales > 0 and a nonzero quaternion; GT_E_INVALID names the first that
 * is not. */
int gt_host_activate_packed(HostScene& h) {
  h.act.assign((size_t)h.n * 8, 0.0);
  for (uint32_t i = 0; i < h.n; ++i) {
    const float* o = &h.packed[(size_t)i * 80];
    double* a = &h.act[(size_t)i * 8];
    bool ok = std::isfinite(o[0]) && std::isfinite(o[1]) && std::isfinite(o[2]) && std::isfinite(o[12]);
    for (int j = 0; j < 3; ++j) ok = ok && std::isfinite(o[4 + j]) && o[4 + j] > 0.0f;
    for (int j = 0; j < 4; ++j) ok = ok && std::isfinite(o[8 + j]);
    ok = ok && (o[8] != 0.0f || o[9] != 0.0f || o[10] != 0.0f || o[11] != 0.0f);
    for (int j = 0; j < 16; ++j)
      for (int c = 0; c < 3; ++c) ok = ok && std::isfinite(o[16 + 4 * j + c]);
    if (!ok) {
      gt_set_error("Gaussian %u: non-finite field, scale <= 0 or zero quaternion in the packed record", i);
      return GT_E_INVALID;
    }
    a[0] = o[12];
    for (int j = 0; j < 3; ++j) a[1 + j] = o[4 + j];
    for (int j = 0; j < 4; ++j) a[4 + j] = o[8 + j];
  }
  return GT_OK;
}

/* ---------------- proxies (src/main.ts:73-94, webrtx/bvh/src/lib.rs:501-505) ---------------- */
int gt_host_build_proxies(HostScene& h, double alpha_min) {
  const unsigned chunks = gt_host_threads(h.n);
  std::vector<std::vector<uint32_t>> ids(chunks);
  std::vector<std::vector<float>> o2w(chunks), w2o(chunks);
  std::vector<uint32_t> bad(chunks, 0xffffffffu);
  gt_parallel_chunks(h.n, chunks, [&](size_t i0, size_t i1, unsigned c) {
    for (size_t i = i0; i < i1; ++i) {
      float m[12], inv[12], lo[3], hi[3];
      if (!gt_prep_proxy(&h.act[i * 8], &h.packed[i * 80], alpha_min, m, inv)) continue;
      gt_prep_box(m, lo, hi);
      if (bad[c] == 0xffffffffu && !gt_prep_proxy_finite(m, inv, lo, hi)) bad[c] = (uint32_t)i;
      ids[c].push_back((uint32_t)i);
      o2w[c].insert(o2w[c].end(), m, m + 12);
      w2o[c].insert(w2o[c].end(), inv, inv + 12);
    }
  });
  for (unsigned c = 0; c < chunks; ++c)
    if (bad[c] != 0xffffffffu) {
      gt_set_error("Gaussian %u: proxy transform or box not finite in f32 (position or scale too large, or a "
                   "scale too small to invert)", bad[c]);
      return GT_E_INVALID;
    }
  h.proxy_ids.clear();
  h.o2w.clear();
  h.w2o.clear();
  for (unsigned c = 0; c < chunks; ++c) { /* in Gaussian order, as one pass would make them */
    h.proxy_ids.insert(h.proxy_ids.end(), ids[c].begin(), ids[c].end());
    h.o2w.insert(h.o2w.end(), o2w[c].begin(), o2w[c].end());
    h.w2o.insert(h.w2o.end(), w2o[c].begin(), w2o[c].end());
  }
  return GT_OK;
}

/* ---------------- BVH ---------------- */
namespace {


struct Box {
  float lo[3], hi[3];
  void empty() {
    for (int j = 0; j < 3; ++j) { lo[j] = INFINITY; hi[j] = -INFINITY; }
  }
  void grow(const Box& b) {
    for (int j = 0; j < 3; ++j) { lo[j] = std::min(lo[j], b.lo[j]); hi[j] = std::max(hi[j], b.hi[j]); }
  }
  double area() const {
    double dx = std::max(0.0f, hi[0] - lo[0]), dy = std::max(0.0f, hi[1] - lo[1]), dz = std::max(0.0f, hi[2] - lo[2]);
    return 2.0 * (dx * dy + dy * dz + dz * dx);
  }
};

struct B2Node {
  Box box;
  int32_t left = -1, right = -1;  // leaf if left < 0
  uint32_t first = 0, count = 0;
};

static Box proxy_box(const float* m) {
  Box b;
  gt_prep_box(m, b.lo, b.hi);
  return b;
}


struct Builder {
  const std::vector<Box>& boxes;
  std::vector<float> cent;  // 3 per prim
  std::vector<uint32_t> order;
  std::vector<B2Node> nodes;
  uint32_t max_depth = 0;

  explicit Builder(const std::vector<Box>& b) : boxes(b) {
    const size_t n = b.size();
    cent.resize(3 * n);
    order.resize(n);
    for (size_t i = 0; i < n; ++i) {
      order[i] = (uint32_t)i;
      for (int j = 0; j < 3; ++j) cent[3 * i + j] = 0.5f * (b[i].lo[j] + b[i].hi[j]);
    }
  }

  struct Task { int32_t node; uint32_t first, count, depth; };

  /* One node of the binned-SAH build: its box, then a leaf or a split whose
   * two children are appended to ns and pushed on work.  It reads only the
   * node's own range of order, so subtrees may be built concurrently. */
  static constexpr int NB = 32;

  /* bin of centroid c in [lo, lo + NB / scale), clamped to 0..NB-1; a
   * difference that overflows f32 (boxes near +-FLT_MAX, scale 0) gives
   * NaN or inf, which go to bin 0 / NB-1 instead of an undefined cast */
  static int bin_of(float c, float lo, float scale) {
    const float f = (c - lo) * scale;
    if (f >= (float)(NB - 1)) return NB - 1;
    return f > 0.0f ? (int)f : 0;
  }

  /* Box and centroid bounds of a range, and its proxies binned by centroid
   * on the three axes (axes with no extent are left empty).  Large ranges
   * run in chunks on threads; min, max and counts are exact, so the result
   * does not depend on the chunking. */
  void bounds(uint32_t first, uint32_t count, bool spawn, Box& bb, Box& cb) const {
    const unsigned nth = spawn && count >= (1u << 16) ? gt_host_threads(count) : 1u;
    std::vector<Box> pb(nth), pc(nth);
    gt_parallel_chunks(count, nth, [&](size_t a, size_t b, unsigned c) {
      Box x, y;
      x.empty();
      y.empty();
      for (size_t i = first + a; i < first + b; ++i) {
        const uint32_t p = order[i];
        x.grow(boxes[p]);
        Box q;
        for (int j = 0; j < 3; ++j) q.lo[j] = q.hi[j] = cent[3 * p + j];
        y.grow(q);
      }
      pb[c] = x;
      pc[c] = y;
    });
    bb.empty();
    cb.empty();
    for (unsigned c = 0; c < nth; ++c) { bb.grow(pb[c]); cb.grow(pc[c]); }
  }

  void bin(uint32_t first, uint32_t count, bool spawn, const Box& cb, Box (&bin_box)[3][NB],
           uint32_t (&bin_cnt)[3][NB]) const {
    const unsigned nth = spawn && count >= (1u << 16) ? gt_host_threads(count) : 1u;
    struct Bins { Box box[3][NB]; uint32_t cnt[3][NB]; };
    std::vector<Bins> part(nth);
    float scale[3];
    bool use[3];
    for (int ax = 0; ax < 3; ++ax) {
      const float ext = cb.hi[ax] - cb.lo[ax];
      use[ax] = ext > 0.0f;
      scale[ax] = use[ax] ? NB / ext : 0.0f;
    }
    gt_parallel_chunks(count, nth, [&](size_t a, size_t b, unsigned c) {
      Bins& B = part[c];
      for (int ax = 0; ax < 3; ++ax)
        for (int k = 0; k < NB; ++k) { B.box[ax][k].empty(); B.cnt[ax][k] = 0; }
      for (size_t i = first + a; i < first + b; ++i) {
        const uint32_t p = order[i];
        for (int ax = 0; ax < 3; ++ax) {
          if (!use[ax]) continue;
          const int k = bin_of(cent[3 * p + ax], cb.lo[ax], scale[ax]);
          B.cnt[ax][k]++;
          B.box[ax][k].grow(boxes[p]);
        }
      }
    });
    for (int ax = 0; ax < 3; ++ax)
      for (int k = 0; k < NB; ++k) {
        bin_box[ax][k].empty();
        bin_cnt[ax][k] = 0;
        for (unsigned c = 0; c < nth; ++c) {
          bin_box[ax][k].grow(part[c].box[ax][k]);
          bin_cnt[ax][k] += part[c].cnt[ax][k];
        }
      }
  }

  void step(const Task& t, std::vector<B2Node>& ns, std::vector<Task>& work, uint32_t& maxd, bool spawn) {
    const double C_TRAV = 1.0, C_ISECT = GT_SAH_ISECT;
    const double lam = GT_SAH_LEAF / GT_SAH_ISECT; /* leaf cost in proxy tests */
    maxd = std::max(maxd, t.depth);
    Box bb, cb;
    bounds(t.first, t.count, spawn, bb, cb);
    ns[t.node].box = bb;
    auto make_leaf = [&]() {
      ns[t.node].left = -1;
      ns[t.node].first = t.first;
      ns[t.node].count = t.count;
    };
    if (t.count <= 1) { make_leaf(); return; }
    // best binned SAH split over the three axes
    int best_axis = -1, best_bin = -1;
    double best_cost = INFINITY;
    if (t.depth < 48) {
      Box bins[3][NB];
      uint32_t counts[3][NB];
      bin(t.first, t.count, spawn, cb, bins, counts);
      for (int ax = 0; ax < 3; ++ax) {
        const float ext = cb.hi[ax] - cb.lo[ax];
        if (!(ext > 0.0f)) continue;
        const Box* bin_box = bins[ax];
        const uint32_t* bin_cnt = counts[ax];
        double right_area[NB];
        uint32_t right_cnt[NB];
        Box acc;
        acc.empty();
        uint32_t cnt = 0;
        for (int b = NB - 1; b > 0; --b) {
          acc.grow(bin_box[b]);
          cnt += bin_cnt[b];
          right_area[b] = acc.area();
          right_cnt[b] = cnt;
        }
        acc.empty();
        cnt = 0;
        for (int b = 0; b < NB - 1; ++b) {
          acc.grow(bin_box[b]);
          cnt += bin_cnt[b];
          if (cnt == 0 || right_cnt[b + 1] == 0) continue;
          const double c = acc.area() * (cnt + lam) + right_area[b + 1] * (right_cnt[b + 1] + lam);
          if (c < best_cost) { best_cost = c; best_axis = ax; best_bin = b; }
        }
      }
    }
    const double parent_area = std::max(bb.area(), 1e-30);
    const double split_cost = C_TRAV + C_ISECT * best_cost / parent_area;
    const double leaf_cost = C_ISECT * (t.count + lam);
    if (t.count <= GT_LEAF_MAX && (best_axis < 0 || leaf_cost <= split_cost)) { make_leaf(); return; }
    uint32_t mid;
    if (best_axis >= 0) {
      const int ax = best_axis;
      const float ext = cb.hi[ax] - cb.lo[ax];
      const float scale = NB / ext;
      auto it = std::partition(order.begin() + t.first, order.begin() + t.first + t.count, [&](uint32_t p) {
        return bin_of(cent[3 * p + ax], cb.lo[ax], scale) <= best_bin;
      });
      mid = (uint32_t)(it - order.begin());
    } else {
      mid = t.first + t.count;  // force the median fallback below
    }
    if (mid == t.first || mid == t.first + t.count) {
      // degenerate or too deep: object median on the widest centroid axis
      int ax = 0;
      for (int j = 1; j < 3; ++j)
        if (cb.hi[j] - cb.lo[j] > cb.hi[ax] - cb.lo[ax]) ax = j;
      mid = t.first + t.count / 2;
      std::nth_element(order.begin() + t.first, order.begin() + mid, order.begin() + t.first + t.count,
                       [&](uint32_t a, uint32_t b) {
                         if (cent[3 * a + ax] != cent[3 * b + ax]) return cent[3 * a + ax] < cent[3 * b + ax];
                         return a < b;
                       });
    }
    const int32_t l = (int32_t)ns.size();
    ns.emplace_back();
    const int32_t r = (int32_t)ns.size();
    ns.emplace_back();
    ns[t.node].left = l;
    ns[t.node].right = r;
    work.push_back({r, mid, t.first + t.count - mid, t.depth + 1});
    work.push_back({l, t.first, mid - t.first, t.depth + 1});
  }

  /* The top levels run here; subtrees of at most `grain` proxies go to
   * worker threads, each into a node list of its own, merged in a fixed
   * order afterwards.  A node's split depends only on its range, so the
   * tree is the one the sequential build makes (node numbering aside). */
  void build() {
    const uint32_t n = (uint32_t)boxes.size();
    nodes.reserve(2 * (size_t)n / 2 + 8);
    nodes.emplace_back();
    std::vector<Task> work{{0, 0, n, 0}}, deferred;
    const unsigned nth = gt_host_threads(n);
    const uint32_t grain = std::max<uint32_t>(4096u, n / 256u);
    while (!work.empty()) {
      const Task t = work.back();
      work.pop_back();
      if (nth > 1 && t.count <= grain) {
        deferred.push_back(t);
        continue;
      }
      step(t, nodes, work, max_depth, nth > 1);
    }
    if (deferred.empty()) return;
    std::vector<std::vector<B2Node>> local(deferred.size());
    std::vector<uint32_t> depth(deferred.size(), 0u);
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (size_t i; (i = next.fetch_add(1)) < deferred.size();) {
        std::vector<B2Node>& ns = local[i];
        ns.emplace_back();
        std::vector<Task> w{{0, deferred[i].first, deferred[i].count, deferred[i].depth}};
        while (!w.empty()) {
          const Task t = w.back();
          w.pop_back();
          step(t, ns, w, depth[i], false);
        }
      }
    };
    std::vector<std::thread> th;
    for (unsigned k = 1; k < nth; ++k) th.emplace_back(worker);
    worker();
    for (auto& x : th) x.join();
    for (size_t i = 0; i < deferred.size(); ++i) { /* local node 0 is the deferred node itself */
      const int32_t g0 = deferred[i].node;
      const int32_t base = (int32_t)nodes.size() - 1;
      for (size_t k = 0; k < local[i].size(); ++k) {
        B2Node b = local[i][k];
        if (b.left >= 0) {
          b.left = b.left == 0 ? g0 : base + b.left;
          b.right = b.right == 0 ? g0 : base + b.right;
        }
        if (k == 0) nodes[g0] = b;
        else nodes.push_back(b);
      }
      max_depth = std::max(max_depth, depth[i]);
    }
  }
};

#ifndef GT_COLLAPSE_DP
#define GT_COLLAPSE_DP 1 /* 0: open the largest inner child until four are present */
#endif

/* Collapse the binary tree into 4-wide nodes.
 *
 * GT_COLLAPSE_DP (default): the least-SAH-cost collapse by dynamic
 * programming over the binary tree (the wide-BVH conversion of H. Ylitie,
 * T. Karras, S. Laine, "Efficient Incoherent Ray Traversal on GPUs Through
 * Compressed Wide BVHs", HPG 2017, for width 4).  With the build's costs
 * (C_T per wide node visited, C_I per proxy tested, both times the box area):
 *   D(n, 1) = min(C_I cnt(n) A(n)              [n a leaf, cnt <= 8],
 *                 C_T A(n) + min_k D(l, k) + D(r, 4 - k))   [n a wide node]
 *   D(n, j) = min(D(n, 1), min_{0<k<j} D(l, k) + D(r, j - k))  for j = 2..4
 * D(n, j) is the least cost of the subtree of n as at most j slots of its
 * wide parent; a binary leaf is one slot of cost C_I cnt A.  On the bench
 * scenes this lowers the wide tree's SAH cost by 11-13 % against the greedy
 * rule below, whose nodes always take four children.
 *
 * Otherwise: repeatedly open the inner child with the largest surface area
 * until four children are present. */
struct Collapser {
  const std::vector<B2Node>& b2;
  std::vector<gt_node4> out;
  uint32_t depth = 0;
  std::vector<uint32_t> first, count; /* proxy range of every binary node */
  std::vector<std::array<double, 5>> D;
  std::vector<std::array<uint8_t, 5>> choice; /* j = 1: 0 leaf, k wide node split k; j > 1: 0 one slot, k split */

  explicit Collapser(const std::vector<B2Node>& n) : b2(n) {}

  static uint32_t leaf_of(const B2Node& n) { return gt_leaf_ref(n.first, n.count); }

  void ranges(int32_t n) {
    const B2Node& b = b2[n];
    if (b.left < 0) { first[n] = b.first; count[n] = b.count; return; }
    ranges(b.left);
    ranges(b.right);
    first[n] = first[b.left];
    count[n] = count[b.left] + count[b.right];
  }

  void dp(int32_t n) {
    const B2Node& b = b2[n];
    const double A = b.box.area();
    if (b.left < 0) {
      for (int j = 1; j <= 4; ++j) { D[n][j] = (GT_SAH_ISECT * (double)b.count + GT_SAH_LEAF) * A; choice[n][j] = 0; }
      return;
    }
    dp(b.left);
    dp(b.right);
    double dist4 = INFINITY;
    uint8_t k4 = 1;
    for (uint8_t k = 1; k <= 3; ++k) {
      const double v = D[b.left][k] + D[b.right][4 - k];
      if (v < dist4) { dist4 = v; k4 = k; }
    }
    const double wide = 1.0 * A + dist4;
    const double leafc = count[n] <= GT_LEAF_MAX ? (GT_SAH_ISECT * (double)count[n] + GT_SAH_LEAF) * A : INFINITY;
    if (leafc <= wide) { D[n][1] = leafc; choice[n][1] = 0; }
    else { D[n][1] = wide; choice[n][1] = k4; }
    for (int j = 2; j <= 4; ++j) {
      double best = D[n][1];
      uint8_t ch = 0;
      for (uint8_t k = 1; k < j; ++k) {
        const double v = D[b.left][k] + D[b.right][j - k];
        if (v < best) { best = v; ch = k; }
      }
      D[n][j] = best;
      choice[n][j] = ch;
    }
  }

  /* the slots that subtree n fills in at most j places, left to right */
  void gather(int32_t n, int j, std::vector<int32_t>& slots) const {
    if (j == 1 || choice[n][j] == 0) { slots.push_back(n); return; }
    const int k = choice[n][j];
    gather(b2[n].left, k, slots);
    gather(b2[n].right, j - k, slots);
  }

  bool is_leaf(int32_t n) const {
#if GT_COLLAPSE_DP
    return b2[n].left < 0 || choice[n][1] == 0;
#else
    return b2[n].left < 0;
#endif
  }

  void prepare() {
#if GT_COLLAPSE_DP
    first.assign(b2.size(), 0u);
    count.assign(b2.size(), 0u);
    D.assign(b2.size(), {});
    choice.assign(b2.size(), {});
    ranges(0);
    dp(0);
#endif
  }

  // returns (node index, stack need)
  std::pair<uint32_t, uint32_t> emit(int32_t b2i, uint32_t d) {
    depth = std::max(depth, d + 1);
    std::vector<int32_t> ch;
    const B2Node& root = b2[b2i];
    if (is_leaf(b2i)) {
      ch.push_back(b2i);
    } else {
#if GT_COLLAPSE_DP
      const int k = choice[b2i][1];
      gather(root.left, k, ch);
      gather(root.right, 4 - k, ch);
#else
      ch.push_back(root.left);
      ch.push_back(root.right);
      while (ch.size() < 4) {
        int best = -1;
        double best_area = -1.0;
        for (size_t i = 0; i < ch.size(); ++i) {
          const B2Node& c = b2[ch[i]];
          if (c.left < 0) continue;
          const double a = c.box.area();
          if (a > best_area) { best_area = a; best = (int)i; }
        }
        if (best < 0) break;
        const int32_t open = ch[best];
        ch[best] = b2[open].left;
        ch.insert(ch.begin() + best + 1, b2[open].right);
      }
#endif
    }
    const uint32_t me = (uint32_t)out.size();
    out.emplace_back();
    gt_node4 node;
    std::memset(&node, 0, sizeof(node));
    uint32_t child_need = 0;
    for (int i = 0; i < 4; ++i) {
      node.lo_x[i] = node.lo_y[i] = node.lo_z[i] = INFINITY;  // empty slots never hit
      node.hi_x[i] = node.hi_y[i] = node.hi_z[i] = -INFINITY;
      node.child[i] = GT_NODE_EMPTY;
    }
    for (size_t i = 0; i < ch.size(); ++i) {
      const B2Node& c = b2[ch[i]];
      node.lo_x[i] = c.box.lo[0]; node.hi_x[i] = c.box.hi[0];
      node.lo_y[i] = c.box.lo[1]; node.hi_y[i] = c.box.hi[1];
      node.lo_z[i] = c.box.lo[2]; node.hi_z[i] = c.box.hi[2];
      if (is_leaf(ch[i])) {
#if GT_COLLAPSE_DP
        node.child[i] = gt_leaf_ref(first[ch[i]], count[ch[i]]);
#else
        node.child[i] = leaf_of(c);
#endif
      } else {
        auto sub = emit(ch[i], d + 1);
        node.child[i] = sub.first;
        child_need = std::max(child_need, sub.second);
      }
    }
    out[me] = node;
    const uint32_t m = (uint32_t)ch.size();
    const uint32_t need = std::max(m, (m - 1) + child_need);
    return {me, need};
  }
};

}  // namespace

int gt_host_build_tree(const HostScene& h, HostTree& t) {
  const uint32_t np = (uint32_t)h.proxy_ids.size();
  if (np > GT_MAX_PROXIES) {
    gt_set_error("%u proxies exceed the limit %u", np, GT_MAX_PROXIES);
    return GT_E_LIMIT;
  }
  std::vector<Box> boxes(np);
  gt_parallel_chunks(np, gt_host_threads(np), [&](size_t i0, size_t i1, unsigned) {
    for (size_t i = i0; i < i1; ++i) boxes[i] = proxy_box(&h.o2w[12 * i]);
  });
  t.nodes.clear();
  t.order.clear();
  t.prox.assign((size_t)np * 16, 0.0f);
  t.root = GT_NODE_EMPTY;
  t.need = 0;
  t.depth = 0;
  if (np > 0) {
    Builder b(boxes);
    b.build();
    Collapser c(b.nodes);
    c.prepare();
    auto r = c.emit(0, 0);
    t.root = r.first;
    t.need = r.second;
    t.depth = c.depth;
    t.nodes.swap(c.out);
    t.order.swap(b.order);
    for (uint32_t i = 0; i < np; ++i) {
      const uint32_t p = t.order[i];
      float* d = &t.prox[(size_t)16 * i];
      std::memcpy(d, &h.w2o[(size_t)12 * p], 12 * sizeof(float));
      const uint32_t gid = h.proxy_ids[p];
      std::memcpy(d + 12, &gid, 4);
    }
  }
  if (t.need > GT_STACK_MAX) {
    gt_set_error("BVH needs a traversal stack of %u entries (limit %u)", t.need, GT_STACK_MAX);
    return GT_E_LIMIT;
  }
  return GT_OK;
}

void gt_host_shading_records(const HostScene& h, std::vector<float>& geom, std::vector<float>& sh) {
  geom.assign((size_t)h.n * 12, 0.0f);
  sh.assign((size_t)h.n * 48, 0.0f);
  for (uint32_t i = 0; i < h.n; ++i) {
    const float* o = &h.packed[(size_t)i * 80];
    float* g = &geom[(size_t)i * 12];
    g[0] = o[0]; g[1] = o[1]; g[2] = o[2]; g[3] = o[12];
    /* g[7]: |raw quaternion|, for the adjoint of the normalisation (1 for packed input) */
    float qn = 1.0f;
    if (!h.raw.empty()) {
      const float* r = &h.raw[(size_t)i * GT_RAW_FLOATS + 55];
      qn = (float)std::sqrt((((double)r[0] * r[0] + (double)r[1] * r[1]) + (double)r[2] * r[2]) + (double)r[3] * r[3]);
    }
    g[4] = o[4]; g[5] = o[5]; g[6] = o[6]; g[7] = qn;
    g[8] = o[8]; g[9] = o[9]; g[10] = o[10]; g[11] = o[11];
    for (int j = 0; j < 16; ++j)
      for (int c = 0; c < 3; ++c) sh[(size_t)i * 48 + 3 * j + c] = o[16 + 4 * j + c];
  }
}

/* Host stage of gt_scene_create_raw without a device: activations, packing
 * and proxy construction.  Outputs sized for n (packed n x 80, ids n,
 * o2w/w2o n x 12); *n_proxies receives the count. */
extern "C" int gt_host_prepare(const float* raw, uint32_t n, double alpha_min, float* packed, uint32_t* proxy_ids,
                               float* o2w, float* w2o, uint32_t* n_proxies) {
  if ((n && !raw) || !n_proxies) { gt_set_error("gt_host_prepare: NULL argument"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_host_prepare: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  try {
    HostScene h;
    h.n = n;
    h.raw.assign(raw, raw + (size_t)n * GT_RAW_FLOATS);
    int rc = gt_host_activate_raw(h);
    if (rc != GT_OK) return rc;
    if ((rc = gt_host_build_proxies(h, alpha_min)) != GT_OK) return rc;
    if (packed) std::memcpy(packed, h.packed.data(), h.packed.size() * sizeof(float));
    if (proxy_ids) std::memcpy(proxy_ids, h.proxy_ids.data(), h.proxy_ids.size() * sizeof(uint32_t));
    if (o2w) std::memcpy(o2w, h.o2w.data(), h.o2w.size() * sizeof(float));
    if (w2o) std::memcpy(w2o, h.w2o.data(), h.w2o.size() * sizeof(float));
    *n_proxies = (uint32_t)h.proxy_ids.size();
    return GT_OK;
  } catch (const std::bad_alloc&) {
    gt_set_error("out of host memory");
    return GT_E_NOMEM;
  }
}

