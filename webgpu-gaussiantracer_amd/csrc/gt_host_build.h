/*
 * gt_host_build.h - the host scene builder (C++ only, no HIP): activations
 * of the input records, proxy construction, the binned-SAH binary tree and
 * its collapse into 4-wide nodes, and the per-Gaussian shading records.
 * gt_scene.cpp uploads what these produce; tests/sanitize/ links them
 * with the PLY reader and the PNG writer into an AddressSanitizer /
 * UndefinedBehaviorSanitizer build that needs no GPU.
 */
#ifndef GT_HOST_BUILD_H
#define GT_HOST_BUILD_H

#include <stdint.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "gt_internal.h"

struct HostScene {
  uint32_t n = 0;
  std::vector<float> raw;      // n x 59 (empty for packed input)
  std::vector<double> act;     // n x 8: opacity, scale[3], quat[4] (JS doubles)
  std::vector<float> packed;   // n x 80
  // proxies, in original order
  std::vector<uint32_t> proxy_ids;
  std::vector<float> o2w, w2o;  // n' x 12
};

/* The device arrays of a host-built scene, before upload. */
struct HostTree {
  std::vector<gt_node4> nodes;  /* 4-wide nodes, root first */
  std::vector<float> prox;      /* n' x 16: world-to-object matrix, id bits at [12], in leaf order */
  std::vector<uint32_t> order;  /* order[i] = proxy index (into proxy_ids) at leaf position i */
  uint32_t root = GT_NODE_EMPTY;
  uint32_t need = 0;            /* traversal stack entries a ray can need */
  uint32_t depth = 0;           /* 4-wide levels */
};

/* worker threads for host stages over n items (1 for small inputs) */
inline unsigned gt_host_threads(size_t n) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return (unsigned)std::min<size_t>(std::min(hw, 16u), std::max<size_t>(1, n / 16384));
}

/* f(begin, end, chunk) over [0, n) in `chunks` contiguous chunks on threads */
template <class F>
void gt_parallel_chunks(size_t n, unsigned chunks, F f) {
  if (chunks <= 1) { f((size_t)0, n, 0u); return; }
  std::vector<std::thread> th;
  for (unsigned c = 0; c < chunks; ++c) th.emplace_back([=, &f]() { f(n * c / chunks, n * (c + 1) / chunks, c); });
  for (auto& x : th) x.join();
}

/* h.act and h.packed from h.raw (n x 59); GT_OK, or GT_E_INVALID naming the
 * first record that gt_prep_valid rejects (message set). */
int gt_host_activate_raw(HostScene& h);
/* h.act from h.packed (n x 80); GT_E_INVALID names the first record with a
 * non-finite field, a scale <= 0 or a zero quaternion. */
int gt_host_activate_packed(HostScene& h);
/* h.proxy_ids, h.o2w, h.w2o: the Gaussians whose proxy gt_prep_proxy keeps
 * at alpha_min, in Gaussian order.  GT_OK, or GT_E_INVALID naming the first
 * kept Gaussian whose proxy gt_prep_proxy_finite rejects. */
int gt_host_build_proxies(HostScene& h, double alpha_min);
/* The tree over h's proxies.  GT_OK, or GT_E_LIMIT (message set) when the
 * proxies or the stack need exceed GT_MAX_PROXIES / GT_STACK_MAX. */
int gt_host_build_tree(const HostScene& h, HostTree& t);
/* Shading records: geom n x 12 (mu, opacity, scale, |raw q|, q) and sh
 * n x 48 (sh[j][c]). */
void gt_host_shading_records(const HostScene& h, std::vector<float>& geom, std::vector<float>& sh);

#endif
