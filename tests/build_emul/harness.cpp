/* CPU emulation of the GPU scene build's tree kernels (radix tree, treelet
 * roots, PLOC, refit of the nodes above the treelets, final gather, level
 * collapse): the kernels' own text from gt_build.hip, one thread per lane,
 * real barriers (kern.inc, generated by tests/test_build_emulation.py).
 * Input: a file of float boxes {lo.xyz, hi.xyz}; argv[2] caps their number.
 * Checks: perm is a permutation, every proxy is in exactly one leaf, every
 * box contains its children, no overflow; prints the 4-wide SAH cost and the
 * collapse DP's own prediction of it. */
#include "kern.inc"
#include <numeric>
#include <set>
static uint32_t expand_bits10h(uint32_t v){v=(v*0x00010001u)&0xFF0000FFu;v=(v*0x00000101u)&0x0F00F00Fu;v=(v*0x00000011u)&0xC30C30C3u;v=(v*0x00000005u)&0x49249249u;return v;}
int main(int argc, char** argv) {
  if (argc < 2) { fprintf(stderr, "usage: harness boxes.bin [n]\n"); return 2; }
  FILE* f = fopen(argv[1], "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", argv[1]); return 2; }
  fseek(f, 0, SEEK_END); long sz = ftell(f); fseek(f, 0, SEEK_SET);
  uint32_t np = (uint32_t)(sz / 24); if (argc > 2) np = std::min<uint32_t>(np, atoi(argv[2]));
  if (np < 2) { fprintf(stderr, "need at least 2 boxes\n"); return 2; }
  std::vector<float> B((size_t)np * 6);
  if (fread(B.data(), 4, B.size(), f) != B.size()) { fprintf(stderr, "short read\n"); return 2; }
  fclose(f);
  std::vector<float4> box(2 * (size_t)np);
  float cmin[3] = {1e30f, 1e30f, 1e30f}, cmax[3] = {-1e30f, -1e30f, -1e30f};
  for (uint32_t i = 0; i < np; ++i) {
    box[2 * i] = {B[6 * i], B[6 * i + 1], B[6 * i + 2], 1.0f}; box[2 * i + 1] = {B[6 * i + 3], B[6 * i + 4], B[6 * i + 5], 0.0f};
    for (int j = 0; j < 3; ++j) { float c = 0.5f * (B[6 * i + j] + B[6 * i + 3 + j]); cmin[j] = std::min(cmin[j], c); cmax[j] = std::max(cmax[j], c); }
  }
  std::vector<unsigned long long> keys(np);
  for (uint32_t i = 0; i < np; ++i) { uint32_t q[3];
    for (int j = 0; j < 3; ++j) { float c = 0.5f * (B[6 * i + j] + B[6 * i + 3 + j]); float e = cmax[j] - cmin[j]; float u = e > 0 ? (c - cmin[j]) / e : 0; int v = (int)(u * 1024.f); v = v < 0 ? 0 : (v > 1023 ? 1023 : v); q[j] = v; }
    keys[i] = ((unsigned long long)((expand_bits10h(q[0]) << 2) | (expand_bits10h(q[1]) << 1) | expand_bits10h(q[2])) << 32) | i; }
  std::sort(keys.begin(), keys.end());
  std::vector<float4> lbox(2 * (size_t)np), lbox2(2 * (size_t)np), w2o(3 * (size_t)np, float4{0, 0, 0, 0}), prox(4 * (size_t)np);
  for (uint32_t p = 0; p < np; ++p) { uint32_t id = (uint32_t)keys[p]; lbox[2 * p] = box[2 * id]; lbox[2 * p + 1] = box[2 * id + 1]; }
  const uint32_t bn_cap = 2 * np;
  std::vector<BNode> bn(bn_cap); std::vector<uint32_t> lparent(np), visit(np, 0), items(np + 64), perm(np, 0xffffffffu);
  std::vector<unsigned int> ctl(CTL_WORDS, 0); ctl[CTL_NODES] = 1; ctl[CTL_LVL] = 1;
  auto blocks = [](uint64_t n) { return (uint32_t)((n + 255) / 256); };
  launch(blocks(np - 1), 256, [&]() { k_karras(keys.data(), (int)np, bn.data(), lparent.data()); });
  launch(blocks(np - 1), 256, [&]() { k_treelet_roots(np, bn.data(), items.data(), ctl.data()); });
  printf("treelets %u\n", ctl[CTL_ITEMS]);
  launch(std::min(np, 64u), PLOC_N, [&]() { k_ploc(np, bn.data(), lbox.data(), items.data(), ctl.data(), perm.data(), bn_cap); });
  printf("ploc nodes %u ovf %u\n", ctl[CTL_PNODES], ctl[CTL_OVF]);
  launch(1, 1024, [&]() { k_refit_top(bn.data(), lbox.data(), lparent.data(), items.data(), ctl.data(), visit.data()); });
  std::vector<Front> fa(np + 2), fb(np + 2);
  launch(blocks(np), 256, [&]() { k_gather_final(keys.data(), np, perm.data(), w2o.data(), box.data(), prox.data(), lbox2.data(), fa.data()); });
  /* perm is a permutation */
  { std::vector<char> seen(np, 0); for (uint32_t p = 0; p < np; ++p) { if (perm[p] >= np || seen[perm[p]]) { printf("BAD perm at %u\n", p); return 1; } seen[perm[p]] = 1; } }
  double RA = box_area(bn[0].lo, bn[0].hi);
  printf("root D1 (4-wide DP cost) %.6f (root leaf %u)\n", bn[0].D[0] / RA, bn[0].leaf);
  const uint32_t cap = np > 1 ? np : 1;
  std::vector<gt_node4> nodes(cap);
  uint32_t launched = 0;
  for (;;) { Front* fin = (launched & 1) ? fb.data() : fa.data(); Front* fout = (launched & 1) ? fa.data() : fb.data();
    launch(std::min(blocks(cap), 64u), 256, [&]() { k_collapse(fin, launched, fout, ctl.data(), bn.data(), lbox2.data(), nodes.data(), cap); });
    ++launched; if (ctl[CTL_LVL + launched] == 0 || launched > 90) break; }
  printf("collapse levels %u nodes %u ovf %u\n", launched, ctl[CTL_NODES], ctl[CTL_OVF]);
  /* validate the 4-wide tree */
  std::vector<int> cover(np, 0); double sah = 0; std::vector<uint32_t> st{0}; uint32_t seen_nodes = 0; int bad = 0;
  while (!st.empty()) { uint32_t j = st.back(); st.pop_back(); ++seen_nodes; const gt_node4& nd = nodes[j];
    for (int c = 0; c < 4; ++c) { uint32_t r = nd.child[c]; if (r == GT_NODE_EMPTY) continue;
      float lo[3] = {nd.lo_x[c], nd.lo_y[c], nd.lo_z[c]}, hi[3] = {nd.hi_x[c], nd.hi_y[c], nd.hi_z[c]};
      double A = box_area(lo, hi);
      if (r & GT_LEAF_BIT) { uint32_t f0 = r & GT_LEAF_FIRST_MASK, cnt = ((r >> GT_LEAF_COUNT_SHIFT) & GT_LEAF_COUNT_MASK) + 1;
        sah += GT_SAH_ISECT * cnt * A;
        for (uint32_t q = f0; q < f0 + cnt; ++q) { if (q >= np) { bad++; continue; } cover[q]++;
          const float4 a = lbox2[2 * q], b = lbox2[2 * q + 1];
          if (a.x < lo[0] || a.y < lo[1] || a.z < lo[2] || b.x > hi[0] || b.y > hi[1] || b.z > hi[2]) bad++; }
      } else { sah += A; if (r >= ctl[CTL_NODES]) { bad++; continue; } st.push_back(r);
        const gt_node4& k = nodes[r];
        for (int q = 0; q < 4; ++q) if (k.child[q] != GT_NODE_EMPTY && (k.lo_x[q] < lo[0] || k.hi_x[q] > hi[0] || k.lo_y[q] < lo[1] || k.hi_y[q] > hi[1] || k.lo_z[q] < lo[2] || k.hi_z[q] > hi[2])) bad++; } } }
  int miss = 0, dup = 0; for (uint32_t p = 0; p < np; ++p) { if (cover[p] == 0) miss++; if (cover[p] > 1) dup++; }
  printf("4-wide: nodes visited %u of %u, missing %d, dup %d, bad %d, SAH(4-wide, /root area) %.6f\n", seen_nodes, ctl[CTL_NODES], miss, dup, bad, sah / RA);
  return (miss || dup || bad || ctl[CTL_OVF]) ? 1 : 0;
}
