/*
 * gt_scene.cpp - scene (acceleration container) construction and the
 * gt_render host wrapper.
 *
 * Replaces, for the Gaussian path:
 *   PackedGaussians activations + packing   src/GS_ply.ts:187-252
 *   gaussianSceneToBvh (filter, proxies)    src/main.ts:30-111
 *   TLAS/BLAS build + upload                webrtx/src/wasm_bvh_builder.ts:260-377,
 *                                           webrtx/bvh/src/lib.rs:253-347, 482-603
 * The reference builds a binary SAH TLAS over instance boxes and walks a
 * 39-node BLAS per instance; here a binned-SAH binary tree over padded
 * world boxes of the proxies is collapsed into 4-wide 128-B nodes whose
 * leaves hold runs of 64-B proxy records (world-to-object matrix + id), and
 * the 20 icosahedron faces are tested directly (no BLAS).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <thread>
#include <vector>

#include "../../include/gt.h"
#include "../../include/gt_fp32.h"
#include "gt_internal.h"
#include "gt_prep.h"
#include "gt_scene_impl.h"

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      gt_set_error("%s failed: %s", #expr, hipGetErrorString(e_));                 \
      return GT_E_HIP;                                                             \
    }                                                                              \
  } while (0)

/* ---------------- activations (src/GS_ply.ts:187-196, 225-235) ---------------- */
/* worker threads for host stages over n items (1 for small inputs) */
static unsigned host_threads(size_t n) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return (unsigned)std::min<size_t>(std::min(hw, 16u), std::max<size_t>(1, n / 16384));
}

/* f(begin, end, chunk) over [0, n) in `chunks` contiguous chunks on threads */
template <class F>
static void parallel_chunks(size_t n, unsigned chunks, F f) {
  if (chunks <= 1) { f((size_t)0, n, 0u); return; }
  std::vector<std::thread> th;
  for (unsigned c = 0; c < chunks; ++c) th.emplace_back([=, &f]() { f(n * c / chunks, n * (c + 1) / chunks, c); });
  for (auto& x : th) x.join();
}

/* Activations of every raw record; returns GT_OK, or GT_E_INVALID naming
 * the first record that gt_prep_valid rejects. */
static int activate_raw(HostScene& h) {
  h.act.assign((size_t)h.n * 8, 0.0);
  h.packed.assign((size_t)h.n * 80, 0.0f);
  const unsigned chunks = host_threads(h.n);
  std::vector<uint32_t> bad(chunks, 0xffffffffu);
  parallel_chunks(h.n, chunks, [&](size_t i0, size_t i1, unsigned c) {
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
static int activate_packed(HostScene& h) {
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
static void build_proxies(HostScene& h, double alpha_min) {
  const unsigned chunks = host_threads(h.n);
  std::vector<std::vector<uint32_t>> ids(chunks);
  std::vector<std::vector<float>> o2w(chunks), w2o(chunks);
  parallel_chunks(h.n, chunks, [&](size_t i0, size_t i1, unsigned c) {
    for (size_t i = i0; i < i1; ++i) {
      float m[12], inv[12];
      if (!gt_prep_proxy(&h.act[i * 8], &h.packed[i * 80], alpha_min, m, inv)) continue;
      ids[c].push_back((uint32_t)i);
      o2w[c].insert(o2w[c].end(), m, m + 12);
      w2o[c].insert(w2o[c].end(), inv, inv + 12);
    }
  });
  h.proxy_ids.clear();
  h.o2w.clear();
  h.w2o.clear();
  for (unsigned c = 0; c < chunks; ++c) { /* in Gaussian order, as one pass would make them */
    h.proxy_ids.insert(h.proxy_ids.end(), ids[c].begin(), ids[c].end());
    h.o2w.insert(h.o2w.end(), o2w[c].begin(), o2w[c].end());
    h.w2o.insert(h.w2o.end(), w2o[c].begin(), w2o[c].end());
  }
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

  /* Box and centroid bounds of a range, and its proxies binned by centroid
   * on the three axes (axes with no extent are left empty).  Large ranges
   * run in chunks on threads; min, max and counts are exact, so the result
   * does not depend on the chunking. */
  void bounds(uint32_t first, uint32_t count, bool spawn, Box& bb, Box& cb) const {
    const unsigned nth = spawn && count >= (1u << 16) ? host_threads(count) : 1u;
    std::vector<Box> pb(nth), pc(nth);
    parallel_chunks(count, nth, [&](size_t a, size_t b, unsigned c) {
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
    const unsigned nth = spawn && count >= (1u << 16) ? host_threads(count) : 1u;
    struct Bins { Box box[3][NB]; uint32_t cnt[3][NB]; };
    std::vector<Bins> part(nth);
    float scale[3];
    bool use[3];
    for (int ax = 0; ax < 3; ++ax) {
      const float ext = cb.hi[ax] - cb.lo[ax];
      use[ax] = ext > 0.0f;
      scale[ax] = use[ax] ? NB / ext : 0.0f;
    }
    parallel_chunks(count, nth, [&](size_t a, size_t b, unsigned c) {
      Bins& B = part[c];
      for (int ax = 0; ax < 3; ++ax)
        for (int k = 0; k < NB; ++k) { B.box[ax][k].empty(); B.cnt[ax][k] = 0; }
      for (size_t i = first + a; i < first + b; ++i) {
        const uint32_t p = order[i];
        for (int ax = 0; ax < 3; ++ax) {
          if (!use[ax]) continue;
          int k = (int)((cent[3 * p + ax] - cb.lo[ax]) * scale[ax]);
          k = std::min(std::max(k, 0), NB - 1);
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
        int b = (int)((cent[3 * p + ax] - cb.lo[ax]) * scale);
        b = std::min(std::max(b, 0), NB - 1);
        return b <= best_bin;
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
    const unsigned nth = host_threads(n);
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

static void face_table(float* f) {
  for (int i = 0; i < GT_NUM_FACES; ++i) {
    const gt_face fc = gt_ico_face(i);
    const float v[12] = {fc.p0.x, fc.p0.y, fc.p0.z, fc.e0.x, fc.e0.y, fc.e0.z,
                         fc.e1.x, fc.e1.y, fc.e1.z, fc.n.x,  fc.n.y,  fc.n.z};
    std::memcpy(f + 12 * i, v, sizeof(v));
  }
}

template <typename T>
static int upload(DevBuf& b, const T* src, size_t count) {
  if (b.p) { (void)hipFree(b.p); b.p = nullptr; }
  b.bytes = std::max<size_t>(count * sizeof(T), 256);
  hipError_t e = hipMalloc(&b.p, b.bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    gt_set_error("hipMalloc(%zu) failed: %s", b.bytes, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? GT_E_NOMEM : GT_E_HIP;
  }
  if (count) HIP_TRY(hipMemcpy(b.p, src, count * sizeof(T), hipMemcpyHostToDevice));
  return GT_OK;
}


int gt_devbuf_reserve(DevBuf& b, size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  if (b.p && b.bytes >= bytes) return GT_OK;
  if (b.p) { (void)hipFree(b.p); b.p = nullptr; b.bytes = 0; }
  const hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    gt_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? GT_E_NOMEM : GT_E_HIP;
  }
  b.bytes = bytes;
  return GT_OK;
}

int gt_scene_static_upload(gt_scene* s) {
  int rc;
  /* the render kernels' constant face table, before any render (which may
   * be captured into a graph) needs it */
  if ((rc = gt_forward_prepare()) != GT_OK) return rc;
  if (!s->faces.p) {
    float faces[12 * GT_NUM_FACES];
    face_table(faces);
    if ((rc = upload(s->faces, faces, 12 * GT_NUM_FACES)) != GT_OK) return rc;
  }
  if (!s->ctl.p) { /* the persistent renderers' work counters, one set per stream */
    const std::vector<uint32_t> zero((size_t)GT_WORK_SLOTS * GT_WORK_WORDS, 0u);
    if ((rc = upload(s->ctl, zero.data(), zero.size())) != GT_OK) return rc;
  }
  return GT_OK;
}

static int build_and_upload(gt_scene* s) {
  const auto t0 = std::chrono::steady_clock::now();
  HostScene& h = s->h;
  build_proxies(h, s->alpha_min);
  const uint32_t np = (uint32_t)h.proxy_ids.size();
  if (np > GT_MAX_PROXIES) {
    gt_set_error("%u proxies exceed the limit %u", np, GT_MAX_PROXIES);
    return GT_E_LIMIT;
  }
  std::vector<Box> boxes(np);
  parallel_chunks(np, host_threads(np), [&](size_t i0, size_t i1, unsigned) {
    for (size_t i = i0; i < i1; ++i) boxes[i] = proxy_box(&h.o2w[12 * i]);
  });

  std::vector<gt_node4> nodes;
  std::vector<float> prox((size_t)np * 16, 0.0f);
  uint32_t root = GT_NODE_EMPTY, need = 0, depth = 0;
  if (np > 0) {
    Builder b(boxes);
    b.build();
    Collapser c(b.nodes);
    c.prepare();
    auto r = c.emit(0, 0);
    root = r.first;
    need = r.second;
    depth = c.depth;
    nodes.swap(c.out);
    for (uint32_t i = 0; i < np; ++i) {
      const uint32_t p = b.order[i];
      float* d = &prox[(size_t)16 * i];
      std::memcpy(d, &h.w2o[(size_t)12 * p], 12 * sizeof(float));
      const uint32_t gid = h.proxy_ids[p];
      std::memcpy(d + 12, &gid, 4);
    }
  }
  if (need > GT_STACK_MAX) {
    gt_set_error("BVH needs a traversal stack of %u entries (limit %u)", need, GT_STACK_MAX);
    return GT_E_LIMIT;
  }
  // per-Gaussian shading records
  std::vector<float> geom((size_t)h.n * 12, 0.0f), sh((size_t)h.n * 48, 0.0f);
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
  int rc;
  if ((rc = upload(s->geom, geom.data(), geom.size())) != GT_OK) return rc;
  if ((rc = upload(s->sh, sh.data(), sh.size())) != GT_OK) return rc;
  if ((rc = upload(s->prox, prox.data(), prox.size())) != GT_OK) return rc;
  if ((rc = upload(s->nodes, nodes.data(), nodes.size())) != GT_OK) return rc;
  if ((rc = gt_scene_static_upload(s)) != GT_OK) return rc;
  s->dev.geom = (const float*)s->geom.p;
  s->dev.sh = (const float*)s->sh.p;
  s->dev.prox = (const float*)s->prox.p;
  s->dev.nodes = (const gt_node4*)s->nodes.p;
  s->dev.faces = (const float*)s->faces.p;
  s->dev.root = root;
  s->n = h.n;
  s->dev.n = h.n;
  s->dev.n_prox = np;
  s->dev.n_nodes = (uint32_t)nodes.size();
  s->n_nodes = (uint32_t)nodes.size();
  s->max_stack = need;
  s->depth = depth;
  s->rsched = false; /* depth-first numbering: the refit schedule is made on first use */
  s->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return GT_OK;
}

static int finish_create(std::unique_ptr<gt_scene>& s, gt_scene** out) {
  if (hipGetDevice(&s->device) != hipSuccess) {
    gt_set_error("no HIP device available");
    return GT_E_HIP;
  }
  const int rc = build_and_upload(s.get());
  if (rc != GT_OK) return rc;
  *out = s.release();
  return GT_OK;
}

extern "C" int gt_scene_create_raw(const float* raw, uint32_t n, double alpha_min, gt_scene** out) {
  if (!out) { gt_set_error("gt_scene_create_raw: out is NULL"); return GT_E_INVALID; }
  *out = nullptr;
  if (n && !raw) { gt_set_error("gt_scene_create_raw: raw is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_create_raw: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  std::unique_ptr<gt_scene> s(new (std::nothrow) gt_scene());
  if (!s) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  try {
    s->h.n = n;
    s->n = n;
    s->h.raw.assign(raw, raw + (size_t)n * GT_RAW_FLOATS);
    s->alpha_min = alpha_min;
    const int rc = activate_raw(s->h);
    if (rc != GT_OK) return rc;
    return finish_create(s, out);
  } catch (const std::bad_alloc&) {
    gt_set_error("out of host memory");
    return GT_E_NOMEM;
  }
}

extern "C" int gt_scene_create_packed(const void* records, uint32_t n, double alpha_min, gt_scene** out) {
  if (!out) { gt_set_error("gt_scene_create_packed: out is NULL"); return GT_E_INVALID; }
  *out = nullptr;
  if (n && !records) { gt_set_error("gt_scene_create_packed: records is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_create_packed: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  std::unique_ptr<gt_scene> s(new (std::nothrow) gt_scene());
  if (!s) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  try {
    s->h.n = n;
    s->n = n;
    const float* f = (const float*)records;
    s->h.packed.assign(f, f + (size_t)n * 80);
    s->alpha_min = alpha_min;
    const int rc = activate_packed(s->h);
    if (rc != GT_OK) return rc;
    return finish_create(s, out);
  } catch (const std::bad_alloc&) {
    gt_set_error("out of host memory");
    return GT_E_NOMEM;
  }
}

extern "C" int gt_scene_set_alpha_min(gt_scene* s, double alpha_min) {
  if (!s) { gt_set_error("gt_scene_set_alpha_min: scene is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_set_alpha_min: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  s->alpha_min = alpha_min;
  if (s->device_built) return gt_device_build(s, nullptr);
  try {
    return build_and_upload(s);
  } catch (const std::bad_alloc&) {
    gt_set_error("out of host memory");
    return GT_E_NOMEM;
  }
}

extern "C" void gt_scene_destroy(gt_scene* s) { delete s; }

/* ---------------- device build (GPU LBVH) ---------------- */
static int device_rebuild(gt_scene* s, const float* raw_dev, void* stream) {
  const size_t bytes = (size_t)s->n * GT_RAW_FLOATS * sizeof(float);
  int rc = gt_devbuf_reserve(s->raw_dev, bytes);
  if (rc != GT_OK) return rc;
  if (bytes) HIP_TRY(hipMemcpyAsync(s->raw_dev.p, raw_dev, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return gt_device_build(s, (hipStream_t)stream);
}

extern "C" int gt_scene_create_device(const float* raw_dev, uint32_t n, double alpha_min, void* stream,
                                      gt_scene** out) {
  if (!out) { gt_set_error("gt_scene_create_device: out is NULL"); return GT_E_INVALID; }
  *out = nullptr;
  if (n && !raw_dev) { gt_set_error("gt_scene_create_device: raw_dev is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_create_device: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  if (n > GT_MAX_PROXIES) {
    gt_set_error("gt_scene_create_device: %u Gaussians exceed the limit %u", n, GT_MAX_PROXIES);
    return GT_E_LIMIT;
  }
  std::unique_ptr<gt_scene> s(new (std::nothrow) gt_scene());
  if (!s) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  if (hipGetDevice(&s->device) != hipSuccess) {
    gt_set_error("no HIP device available");
    return GT_E_HIP;
  }
  s->n = n;
  s->device_built = true;
  s->alpha_min = alpha_min;
  const int rc = device_rebuild(s.get(), raw_dev, stream);
  if (rc != GT_OK) return rc;
  *out = s.release();
  return GT_OK;
}

extern "C" int gt_scene_create_raw_gpu(const float* raw, uint32_t n, double alpha_min, gt_scene** out) {
  if (!out) { gt_set_error("gt_scene_create_raw_gpu: out is NULL"); return GT_E_INVALID; }
  *out = nullptr;
  if (n && !raw) { gt_set_error("gt_scene_create_raw_gpu: raw is NULL"); return GT_E_INVALID; }
  if (!(alpha_min > 0.0 && alpha_min < 1.0)) {
    gt_set_error("gt_scene_create_raw_gpu: alpha_min %g outside (0, 1)", alpha_min);
    return GT_E_INVALID;
  }
  if (n > GT_MAX_PROXIES) {
    gt_set_error("gt_scene_create_raw_gpu: %u Gaussians exceed the limit %u", n, GT_MAX_PROXIES);
    return GT_E_LIMIT;
  }
  std::unique_ptr<gt_scene> s(new (std::nothrow) gt_scene());
  if (!s) { gt_set_error("out of host memory"); return GT_E_NOMEM; }
  if (hipGetDevice(&s->device) != hipSuccess) {
    gt_set_error("no HIP device available");
    return GT_E_HIP;
  }
  s->n = n;
  s->device_built = true;
  s->alpha_min = alpha_min;
  const size_t bytes = (size_t)n * GT_RAW_FLOATS * sizeof(float);
  int rc = gt_devbuf_reserve(s->raw_dev, bytes);
  if (rc != GT_OK) return rc;
  if (bytes) HIP_TRY(hipMemcpy(s->raw_dev.p, raw, bytes, hipMemcpyHostToDevice));
  rc = gt_device_build(s.get(), nullptr);
  if (rc != GT_OK) return rc;
  *out = s.release();
  return GT_OK;
}

extern "C" int gt_scene_update_device(gt_scene* s, const float* raw_dev, void* stream) {
  if (!s) { gt_set_error("gt_scene_update_device: scene is NULL"); return GT_E_INVALID; }
  if (!s->device_built) {
    gt_set_error("gt_scene_update_device: the scene was not created by gt_scene_create_device");
    return GT_E_INVALID;
  }
  if (s->n && !raw_dev) { gt_set_error("gt_scene_update_device: raw_dev is NULL"); return GT_E_INVALID; }
  return device_rebuild(s, raw_dev, stream);
}

extern "C" int gt_scene_refit_device(gt_scene* s, const float* raw_dev, void* stream, uint32_t* rebuilt) {
  if (!s) { gt_set_error("gt_scene_refit_device: scene is NULL"); return GT_E_INVALID; }
  if (s->n && !raw_dev) { gt_set_error("gt_scene_refit_device: raw_dev is NULL"); return GT_E_INVALID; }
  uint32_t r = 0;
  const int rc = gt_device_refit(s, raw_dev, (hipStream_t)stream, &r);
  if (rebuilt) *rebuilt = r;
  return rc;
}

extern "C" int gt_scene_get_info(const gt_scene* s, gt_scene_info* info) {
  if (!s || !info) { gt_set_error("gt_scene_get_info: NULL argument"); return GT_E_INVALID; }
  std::memset(info, 0, sizeof(*info));
  info->n_gaussians = s->n;
  info->n_proxies = s->dev.n_prox;
  info->n_nodes = s->n_nodes;
  info->max_stack = s->max_stack;
  info->bvh_depth = s->depth;
  info->device_bytes = s->geom.bytes + s->sh.bytes + s->prox.bytes + s->nodes.bytes + s->faces.bytes +
                      s->raw_dev.bytes;
  info->build_ms = s->build_ms;
  return GT_OK;
}

extern "C" int gt_scene_get_host_data(const gt_scene* s, float* packed, uint32_t* proxy_ids, float* o2w,
                                      float* w2o) {
  if (!s) { gt_set_error("gt_scene_get_host_data: scene is NULL"); return GT_E_INVALID; }
  if (s->device_built) return gt_device_host_data(s, packed, proxy_ids, o2w, w2o);
  const HostScene& h = s->h;
  if (packed) std::memcpy(packed, h.packed.data(), h.packed.size() * sizeof(float));
  if (proxy_ids) std::memcpy(proxy_ids, h.proxy_ids.data(), h.proxy_ids.size() * sizeof(uint32_t));
  if (o2w) std::memcpy(o2w, h.o2w.data(), h.o2w.size() * sizeof(float));
  if (w2o) std::memcpy(w2o, h.w2o.data(), h.w2o.size() * sizeof(float));
  return GT_OK;
}

/* ---------------- render ---------------- */
int gt_validate_uniforms(const gt_uniforms* u, uint32_t width, uint32_t height) {
  if (!u) { gt_set_error("uniforms are NULL"); return GT_E_INVALID; }
  if (width < 1 || height < 1 || width > 16384 || height > 16384) {
    gt_set_error("width(%u), height(%u) must be in 1..16384", width, height);
    return GT_E_INVALID;
  }
  if (u->k < 1 || u->k > GT_MAX_K) {
    gt_set_error("hit_array_size k=%u must be in 1..%d", u->k, GT_MAX_K);
    return GT_E_LIMIT;
  }
  if (!(u->t_min >= 0.0f) || !std::isfinite(u->t_min) || !(u->t_max > u->t_min)) {
    gt_set_error("need 0 <= t_min < t_max (t_min=%g, t_max=%g)", u->t_min, u->t_max);
    return GT_E_INVALID;
  }
  if (u->sh_degree_max > 3) { gt_set_error("sh_degree_max %u > 3", u->sh_degree_max); return GT_E_INVALID; }
  if (u->debug_mode > 2) { gt_set_error("debugMode %u not in {0,1,2}", u->debug_mode); return GT_E_INVALID; }
  for (int i = 0; i < 16; ++i)
    if (!std::isfinite(u->to_world[i])) { gt_set_error("to_world has a non-finite entry"); return GT_E_INVALID; }
  if (!std::isfinite(u->camera_aspect[0]) || !std::isfinite(u->camera_aspect[1]) || !std::isfinite(u->T_min) ||
      !std::isfinite(u->alpha_min)) {
    gt_set_error("non-finite uniform value");
    return GT_E_INVALID;
  }
  return GT_OK;
}

void gt_fill_dev_frame(const gt_uniforms* u, uint32_t width, uint32_t height, gt_dev_frame* f) {
  std::memset(f, 0, sizeof(*f));
  std::memcpy(f->to_world, u->to_world, sizeof(f->to_world));
  f->ax = u->camera_aspect[0];
  f->ay = u->camera_aspect[1];
  f->t_min = u->t_min;
  f->t_max = u->t_max;
  f->T_min = u->T_min;
  f->alpha_min = u->alpha_min;
  f->k = u->k;
  f->sh_degree = u->sh_degree_max;
  f->early_stop = u->early_stop;
  f->debug_mode = u->debug_mode;
  f->width = width;
  f->height = height;
  const float om = std::max(std::fabs(u->to_world[12]), std::max(std::fabs(u->to_world[13]), std::fabs(u->to_world[14])));
  f->slack_abs = 2e-5f * (om + 1.0f);
  f->accumulated_frames = u->accumulated_frames;
  f->tiles_x = (width + 7u) / 8u;
  f->tile_base = 0;
}

/* The per-tile kernel by default (faster on a single camera: the rays of a
 * wave stay coherent), the persistent kernel with tuning path 2. */
uint32_t gt_tune_tile_w(const gt_tuning* t) { return t->tile_w ? t->tile_w : 8u; }

/* tiles_x / tile_base of a frame for the tuning's tile shape; returns its tile count */
uint32_t gt_frame_tiles(const gt_tuning* t, gt_dev_frame* f, uint32_t tile_base) {
  const uint32_t tw = gt_tune_tile_w(t), th = 64u / tw;
  f->tiles_x = (f->width + tw - 1u) / tw;
  f->tile_base = tile_base;
  return f->tiles_x * ((f->height + th - 1u) / th);
}

/* The work counters of `stream` in this scene (assigned on the stream's
 * first render), or NULL once GT_WORK_SLOTS streams hold one. */
unsigned int* gt_scene_work(gt_scene* s, void* stream) {
  std::lock_guard<std::mutex> lk(s->ctl_mu);
  unsigned int* base = (unsigned int*)s->ctl.p;
  if (!base) return nullptr;
  for (uint32_t i = 0; i < s->ctl_used; ++i)
    if (s->ctl_streams[i] == stream) return base + (size_t)i * GT_WORK_WORDS;
  if (s->ctl_used == GT_WORK_SLOTS) return nullptr;
  s->ctl_streams[s->ctl_used] = stream;
  return base + (size_t)(s->ctl_used++) * GT_WORK_WORDS;
}

/* The persistent packet kernel by default (tuning path 0 or 4); the tile
 * kernels with paths 1 and 3; the lane-refill persistent kernel with path 2
 * (which leaves debug modes, the backward checkpoint and antialiasing to the
 * tile kernel).  A persistent kernel needs this stream's work counters; a
 * scene rendered from more than GT_WORK_SLOTS streams renders on the
 * further streams with the tile kernel (results never depend on the path). */
static int launch_frame(gt_scene* s, gt_dev_frame* f, void* stream) {
  const uint32_t tiles = gt_frame_tiles(&s->tune, f, 0);
  const uint32_t path = s->tune.path == 0 ? 4u : s->tune.path;
  unsigned int* work = (path == 2 || path == 4) ? gt_scene_work(s, stream) : nullptr;
  gt_tuning t = s->tune;
  t.path = path;
  if ((path == 2 || path == 4) && !work) t.path = 3;
  if (t.path != 2 || f->debug_mode != 0 || f->comp_ids || f->n_comp || f->antialias) {
    if (t.path == 2) t.path = 3;
    return gt_launch_forward(&s->dev, f, nullptr, 1, tiles, &t, stream, 0, work);
  }
  return gt_launch_persistent(&s->dev, f, nullptr, 1, tiles, work, (unsigned long long*)f->stats, &t, stream);
}

extern "C" int gt_scene_set_tuning(gt_scene* s, const gt_tuning* t) {
  if (!s || !t) { gt_set_error("gt_scene_set_tuning: NULL argument"); return GT_E_INVALID; }
  const bool tile_ok = t->tile_w == 0 || t->tile_w == 4 || t->tile_w == 8 || t->tile_w == 16 || t->tile_w == 32;
  const bool lds_ok = t->stack_lds == 0 || t->stack_lds == 8 || t->stack_lds == 16 || t->stack_lds == 32 ||
                      t->stack_lds == 64;
  if (t->path > 4 || t->refill_min > 64 || t->comp_min > 64 || t->blocks_per_cu > 8 || !tile_ok || !lds_ok) {
    gt_set_error("gt_scene_set_tuning: value out of range (path %u, refill %u, comp %u, blocks %u, tile_w %u, "
                 "stack_lds %u)", t->path, t->refill_min, t->comp_min, t->blocks_per_cu, t->tile_w, t->stack_lds);
    return GT_E_INVALID;
  }
  s->tune = *t;
  return GT_OK;
}

extern "C" int gt_scene_get_tuning(const gt_scene* s, gt_tuning* t) {
  if (!s || !t) { gt_set_error("gt_scene_get_tuning: NULL argument"); return GT_E_INVALID; }
  *t = s->tune;
  return GT_OK;
}

const gt_dev_scene* gt_scene_dev(const gt_scene* s) { return &s->dev; }
const gt_tuning* gt_scene_tuning(const gt_scene* s) { return &s->tune; }

static void* grow(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return b.p;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  const hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    gt_set_error("backward scratch hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    return nullptr;
  }
  b.bytes = bytes;
  return b.p;
}

/* Backward scratch of at least `bytes` (contents not kept when it grows),
 * allocated on first use and kept with the scene; NULL (with the error set)
 * when the allocation fails.  `which` 0: accumulators and colours, 1: the
 * recompute mode's band buffers. */
void* gt_scene_bwd_scratch(gt_scene* s, int which, size_t bytes) { return grow(which ? s->rc : s->bwd, bytes); }
uint32_t* gt_scene_rc_cap(gt_scene* s) { return &s->rc_cap; }

extern "C" int gt_render(gt_scene* s, const gt_uniforms* u, uint32_t width, uint32_t height, const gt_frame* out,
                         int buf_flags, void* stream) {
  if (!s || !out || !out->rgba) { gt_set_error("gt_render: scene, out and out->rgba are required"); return GT_E_INVALID; }
  int rc = gt_validate_uniforms(u, width, height);
  if (rc != GT_OK) return rc;
  if ((out->hit_ids || out->accept) && out->max_rounds == 0) {
    gt_set_error("gt_render: max_rounds must be >= 1 when hit_ids/accept are requested");
    return GT_E_INVALID;
  }
  if (out->comp_ids && (!out->n_comp || out->comp_cap == 0)) {
    gt_set_error("gt_render: the checkpoint needs comp_ids, n_comp and comp_cap >= 1");
    return GT_E_INVALID;
  }
  if ((out->comp_ids || out->n_comp) && u->debug_mode != 0) {
    gt_set_error("gt_render: the checkpoint is written in debugMode 0 only");
    return GT_E_INVALID;
  }
  if (out->antialiasing && (out->comp_ids || out->n_comp)) {
    gt_set_error("gt_render: antialiasing renders do not write the backward checkpoint");
    return GT_E_INVALID;
  }
  gt_dev_frame f;
  gt_fill_dev_frame(u, width, height, &f);
  f.antialias = out->antialiasing ? 1u : 0u;
  const size_t R = (size_t)width * height;
  const size_t MR = out->max_rounds;
  if (buf_flags == GT_BUF_DEVICE) {
    f.rgba = out->rgba;
    f.hit_ids = out->hit_ids;
    f.accept = out->accept;
    f.n_rounds = out->n_rounds;
    f.T_out = out->T;
    f.stats = (unsigned long long*)out->stats;
    f.max_rounds = (uint32_t)MR;
    f.comp_ids = out->comp_ids;
    f.n_comp = out->n_comp;
    f.comp_cap = out->comp_cap;
    if (f.antialias && f.accumulated_frames) { /* render into scratch, then average into rgba */
      if ((rc = gt_devbuf_reserve(s->aa, R * 16)) != GT_OK) return rc;
      f.rgba = (float*)s->aa.p;
      if ((rc = launch_frame(s, &f, stream)) != GT_OK) return rc;
      return gt_launch_aa_mix(f.rgba, out->rgba, (uint32_t)R, f.accumulated_frames, stream);
    }
    return launch_frame(s, &f, stream);
  }
  // host buffers: device temporaries, synchronous
  struct Tmp {
    void* p = nullptr;
    ~Tmp() { if (p) (void)hipFree(p); }
  } t_hits, t_acc, t_rounds, t_T, t_stats, t_cids, t_ncomp;
  auto alloc = [&](Tmp& t, size_t bytes) -> int {
    hipError_t e = hipMalloc(&t.p, std::max<size_t>(bytes, 4));
    if (e != hipSuccess) {
      t.p = nullptr;
      gt_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
      return GT_E_NOMEM;
    }
    return GT_OK;
  };
  /* the image goes through a buffer the scene keeps: no allocation per
   * frame once the largest frame size has been seen */
  if ((rc = gt_devbuf_reserve(s->himg, R * 16)) != GT_OK) return rc;
  f.rgba = (float*)s->himg.p;
  if (out->hit_ids) {
    if ((rc = alloc(t_hits, R * MR * u->k * 4)) != GT_OK) return rc;
    f.hit_ids = (uint32_t*)t_hits.p;
  }
  if (out->accept) {
    if ((rc = alloc(t_acc, R * MR * 4)) != GT_OK) return rc;
    f.accept = (uint32_t*)t_acc.p;
  }
  if (out->n_rounds) {
    if ((rc = alloc(t_rounds, R * 4)) != GT_OK) return rc;
    f.n_rounds = (uint32_t*)t_rounds.p;
  }
  if (out->T) {
    if ((rc = alloc(t_T, R * 4)) != GT_OK) return rc;
    f.T_out = (float*)t_T.p;
  }
  if (out->stats) {
    if ((rc = alloc(t_stats, GT_NSTATS * 8)) != GT_OK) return rc;
    HIP_TRY(hipMemset(t_stats.p, 0, GT_NSTATS * 8));
    f.stats = (unsigned long long*)t_stats.p;
  }
  if (out->comp_ids) {
    if ((rc = alloc(t_cids, R * out->comp_cap * 4)) != GT_OK) return rc;
    f.comp_ids = (uint32_t*)t_cids.p;
    f.comp_cap = out->comp_cap;
  }
  if (out->n_comp) {
    if ((rc = alloc(t_ncomp, R * 4)) != GT_OK) return rc;
    f.n_comp = (uint32_t*)t_ncomp.p;
  }
  f.max_rounds = (uint32_t)MR;
  if ((rc = launch_frame(s, &f, nullptr)) != GT_OK) return rc;
  if (f.antialias && f.accumulated_frames) { /* average the new frame into the caller's image */
    Tmp t_avg;
    if ((rc = alloc(t_avg, R * 16)) != GT_OK) return rc;
    HIP_TRY(hipMemcpy(t_avg.p, out->rgba, R * 16, hipMemcpyHostToDevice));
    if ((rc = gt_launch_aa_mix(f.rgba, (float*)t_avg.p, (uint32_t)R, f.accumulated_frames, nullptr)) != GT_OK)
      return rc;
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out->rgba, t_avg.p, R * 16, hipMemcpyDeviceToHost));
  } else {
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out->rgba, f.rgba, R * 16, hipMemcpyDeviceToHost));
  }
  if (out->hit_ids) HIP_TRY(hipMemcpy(out->hit_ids, f.hit_ids, R * MR * u->k * 4, hipMemcpyDeviceToHost));
  if (out->accept) HIP_TRY(hipMemcpy(out->accept, f.accept, R * MR * 4, hipMemcpyDeviceToHost));
  if (out->n_rounds) HIP_TRY(hipMemcpy(out->n_rounds, f.n_rounds, R * 4, hipMemcpyDeviceToHost));
  if (out->T) HIP_TRY(hipMemcpy(out->T, f.T_out, R * 4, hipMemcpyDeviceToHost));
  if (out->stats) HIP_TRY(hipMemcpy(out->stats, f.stats, GT_NSTATS * 8, hipMemcpyDeviceToHost));
  if (out->comp_ids) HIP_TRY(hipMemcpy(out->comp_ids, f.comp_ids, R * out->comp_cap * 4, hipMemcpyDeviceToHost));
  if (out->n_comp) HIP_TRY(hipMemcpy(out->n_comp, f.n_comp, R * 4, hipMemcpyDeviceToHost));
  return GT_OK;
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
    const int rc = activate_raw(h);
    if (rc != GT_OK) return rc;
    build_proxies(h, alpha_min);
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

/* Self-check of the face pre-test on the host: for n (proxy, ray) pairs,
 * cand[i] = gt_face_candidates mask, exact[i] = faces accepted by the exact
 * triangle test + facing rule over (tmin, tmax).  The kernel evaluates the
 * same shared functions; exact & ~cand must be 0. */
extern "C" int gt_debug_face_test(const float* w2o, const float* o, const float* d, uint32_t n, float tmin, float tmax,
                                  float margin_scale, uint32_t* cand, uint32_t* exact) {
  if (!w2o || !o || !d || !cand || !exact) { gt_set_error("gt_debug_face_test: NULL argument"); return GT_E_INVALID; }
  if (!gt_ico_literals_match()) { gt_set_error("icosahedron literal tables out of date"); return GT_E_LIMIT; }
  for (uint32_t i = 0; i < n; ++i) {
    const float* m = w2o + 12 * (size_t)i;
    const gt_f3 ro = gt_f3_make(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
    const gt_f3 rd = gt_f3_make(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
    const gt_f3 oo = gt_xform_point(m, ro), od = gt_xform_dir(m, rd);
    float mg, mn;
    gt_face_margins(oo, od, m, &mg, &mn);
    cand[i] = gt_face_candidates(oo, od, mg * margin_scale, mn * margin_scale);
    uint32_t ex = 0;
    for (int f = 0; f < GT_NUM_FACES; ++f) {
      const gt_face fc = gt_ico_face(f);
      float t;
      if (gt_tri_hit(oo, od, tmin, tmax, fc, &t) && gt_front_facing(m, fc.n, rd)) ex |= 1u << f;
    }
    exact[i] = ex;
  }
  return GT_OK;
}
