/*
 * gt_forward.hip - fused forward kernel of the Gaussian ray tracer (gfx950).
 *
 * One lane = one ray (pixel).  Per ray, rounds of
 *   traverse the 4-wide BVH over icosahedron proxies, keeping the k
 *   closest front-face hits in a register k-buffer          (any-hit, GS_shader.ts:323-368)
 *   t_curr = t_k + 1e-4                                      (GS_shader.ts:248)
 *   composite the k entries front to back with max-response
 *   opacity and SH radiance                                  (GS_shader.ts:250-269, 111-190)
 * until T <= T_min or the ray leaves the scene (GS_shader.ts:231).  This
 * replaces the ray-gen -> traceRayEXT -> any-hit chain of the reference's
 * emulated RT pipeline (webrtx/src/glsl/trace.glsl:42-245) by one kernel.
 *
 * Bit-exactness with a host C evaluation: the ray, the object-space ray, the
 * triangle test, the facing test, the response and exp are the shared
 * primitives of gt_fp32.h, compiled with -ffp-contract=off, IEEE division
 * and square root and f32 denormals kept.  Every shortcut that skips work
 * (box culling, the bounding-sphere test, the facing pre-test) is
 * conservative with margins far above FP32 rounding, so it never changes
 * which hits are found.  Exact-t ties are ordered by (t, Gaussian id).
 *
 * Launch: 256-thread blocks = 4 waves; a wave covers an 8x8 pixel tile;
 * block ids are remapped so that consecutive tiles share an XCD (and its L2).
 */
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gt.h"
#include "../../include/gt_fp32.h"
#include "gt_internal.h"
#include "gt_rect.h"

namespace {

/* Icosahedron face constants {p0, e0, e1, n} (gt_ico_face), uploaded once
 * per device and copied to LDS by each block.  The pre-test's edge and
 * normal constants are instruction literals (gt_face_candidates). */
__constant__ float c_faces[12 * GT_NUM_FACES];

#define GT_BLOCK 256
/* Waves per SIMD the register budget of k_forward is sized for (512 /
 * waves VGPRs per lane), by k-buffer size.  The packet kernel is bound by
 * per-wave latency (stack pop -> scalar node load -> box tests -> sort ->
 * push), so occupancy pays: with a 16-slot buffer five waves (96 VGPRs)
 * ran the 1080p 1M-Gaussian frame 5 % faster than four, and six spilled
 * too much.  Renders use the 32-slot buffer, whose 64 key registers spill
 * 248 B per lane at five waves (12.8 against 8.0 ms), so it keeps four;
 * at four it needs 161 VGPRs and spills 100 B per lane, yet three waves
 * (no spills) took 7.31 against 6.90 ms.  The one-slot debug-mode kernels
 * run five. */
#ifndef GT_WAVES_K16
#define GT_WAVES_K16 5
#endif
#ifndef GT_WAVES_K32
#define GT_WAVES_K32 4
#endif
template <int MAXK>
struct Waves {
  static constexpr int value = MAXK <= 16 ? GT_WAVES_K16 : GT_WAVES_K32;
};
#define GT_RENDER_WAVES 4 /* k_render (per-lane traversal, tuning path 2) */
#define GT_LDS_STACK 32 /* default LDS stack entries per lane; deeper ones spill to scratch */

/* Traversal stack: the top GT_LDS_STACK entries live in LDS as a ring
 * (entry i at slot i % GT_LDS_STACK, lane-interleaved so that a wave's
 * accesses hit distinct banks); entries below spill to a private array. */
struct Stack {
  uint32_t* lds; /* this lane's column of the block's LDS stack */
  int sp;
  int depth;     /* LDS entries per lane, a power of two */
  uint32_t spill[GT_STACK_MAX];
  __device__ __forceinline__ void push(uint32_t x) {
    const int slot = (sp & (depth - 1)) * GT_BLOCK;
    if (sp >= depth) spill[sp - depth] = lds[slot];
    lds[slot] = x;
    ++sp;
  }
  __device__ __forceinline__ uint32_t pop() {
    --sp;
    const int slot = (sp & (depth - 1)) * GT_BLOCK;
    const uint32_t x = lds[slot];
    if (sp >= depth) lds[slot] = spill[sp - depth];
    return x;
  }
};

struct RayState {
  gt_f3 o, d;
  gt_f3 inv; /* slab-test reciprocals, tiny components clamped */
};

/* k-buffer in registers, sorted by key.  k_forward uses all MAXK slots (a
 * traversal collects the MAXK smallest hits and rounds of k are served from
 * them); k_render keeps k active slots, the LAST k of MAXK, the first
 * MAXK-k holding the key (+0, 0), smaller than every hit key (hits have
 * t > t_curr >= 0), so the insertion never moves them.  Either way the
 * largest collected entry is slot MAXK-1. */
/* Max and min of two k-buffer keys with the FP64 max/min instructions.  A
 * key (t bits above, id below; t >= +0 and not NaN: hit distances, +0 in
 * the dummy slots, t_max, which may be +inf, in the sentinels) read as a
 * double has sign 0 and an exponent field that is the float's exponent and
 * top three mantissa bits, all ones only for a float NaN: a non-negative,
 * finite double.  Such doubles order like their bit patterns, including
 * denormals (the kernels keep FP64 denormals), so v_max_f64 / v_min_f64
 * return the larger / smaller key, bit for bit (they return one of their
 * operands).  Inline ALU asm keeps the compiler from adding NaN
 * canonicalisations. */
__device__ __forceinline__ uint64_t key_max(uint64_t a, uint64_t b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(__longlong_as_double((long long)a)), "v"(__longlong_as_double((long long)b)));
  return (uint64_t)__double_as_longlong(r);
}
__device__ __forceinline__ uint64_t key_min(uint64_t a, uint64_t b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(__longlong_as_double((long long)a)), "v"(__longlong_as_double((long long)b)));
  return (uint64_t)__double_as_longlong(r);
}

template <int MAXK>
struct KBuf {
  /* one 64-bit key per slot: t bits above, id below.  Hit distances are
   * positive, so the unsigned order of the keys is the (t, id) order and a
   * key compare is one 64-bit compare. */
  uint64_t key[MAXK];
  __device__ __forceinline__ float t_at(int i) const { return __uint_as_float((uint32_t)(key[i] >> 32)); }
  __device__ __forceinline__ uint32_t id_at(int i) const { return (uint32_t)key[i]; }
  __device__ __forceinline__ float kth() const { return t_at(MAXK - 1); }

  /* every slot the sentinel (t_max, GT_SENTINEL); t_max must be the same in
   * every lane of the wave.  The opaque copy lives in a scalar register: the
   * reset values cannot be hoisted out of the round loop, and no vector
   * register holds the input across the traversal. */
  __device__ __forceinline__ void reset_empty(float t_max) {
    uint32_t tb = __float_as_uint(t_max);
    asm volatile("" : "+s"(tb));
    const uint64_t sentinel = ((uint64_t)tb << 32) | GT_SENTINEL;
#pragma unroll
    for (int i = 0; i < MAXK; ++i) key[i] = sentinel;
  }
  __device__ __forceinline__ void reset(uint32_t off, float t_max) {
    /* opaque copies of the (uniform) inputs: the MAXK reset values are
     * recomputed in each round instead of hoisted out of the round loop and
     * kept live (spilled) across the traversal */
    asm volatile("" : "+v"(off), "+v"(t_max));
    const uint64_t sentinel = ((uint64_t)__float_as_uint(t_max) << 32) | GT_SENTINEL;
#pragma unroll
    for (int i = 0; i < MAXK; ++i) key[i] = (uint32_t)i < off ? 0ull : sentinel;
  }
  /* GS_shader.ts:349-353 with the key (t, id): stable for equal keys.  The
   * buffer is sorted, so the reference's swap chain shifts every entry from
   * the insertion point on by one slot: slot i becomes slot i-1 if h
   * precedes both, h if it precedes slot i only, else stays, which is
   *   slot'[i] = min(slot[i], max(slot[i-1], h)),  slot'[0] = min(slot[0], h).
   * (Where the chain stops at two equal keys the entries are identical, so
   * shifting them changes nothing.)  Written from the last slot down, every
   * slot is updated in place: one FP64 max and one min per slot, with no
   * lane masks (a compare-and-select form needs two compares, four
   * selects and two wait states on the compare's mask per slot). */
  __device__ __forceinline__ void insert(float ht, uint32_t hid) {
    const uint64_t h = ((uint64_t)__float_as_uint(ht) << 32) | hid;
    if (!(h < key[MAXK - 1])) return;
#pragma unroll
    for (int i = MAXK - 1; i > 0; --i) key[i] = key_min(key[i], key_max(key[i - 1], h));
    key[0] = key_min(key[0], h);
  }
  __device__ __forceinline__ void shift_left(int s) {
#pragma unroll
    for (int i = 0; i + 1 < MAXK; ++i)
      if (i + s < MAXK) key[i] = key[i + s];
  }
};

struct Counters {
  uint32_t rays = 0, rounds = 0, h_eval = 0, h_comp = 0, proxies = 0, hits = 0, nodes = 0, overflow = 0;
};

/* MODE_RENDER_PC renders as MODE_RENDER, compositing the frame's
 * precomputed per-Gaussian colours (F.colors) instead of evaluating each
 * composited hit's SH record: the same FP32 colour (k_colors runs the same
 * radiance() on the same origin), without the 192-B record read and the
 * basis per hit. */
enum { MODE_RENDER = 0, MODE_CLOSEST = 1, MODE_COUNT = 2, MODE_RENDER_PC = 3 };
__host__ __device__ constexpr bool is_render(int mode) { return mode == MODE_RENDER || mode == MODE_RENDER_PC; }

/* Read-only scene arrays seen through the constant address space: a load
 * whose address is the same in every lane becomes one scalar load. */
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) f4v cfloat4;
typedef const __attribute__((address_space(4))) u4v cuint4;

template <typename T>
__device__ __forceinline__ T* as_const(const void* p) {
  return (T*)(uintptr_t)p;
}


/* Exact proxy test: object-space ray (trace.glsl:102-105), the triangle
 * test of intersect.glsl:73-92 and the facing rule of trace.glsl:153-164.
 * A bounding-sphere test and the Pluecker face pre-test (gt_face_candidates)
 * select the faces that can pass; only those run the exact test, so a lane
 * normally runs one exact test and one k-buffer insertion per proxy.
 * MODE_CLOSEST keeps back faces (debugMode 1 traces with no cull flag,
 * GS_shader.ts:292-294) and tests all 20 faces. */
template <typename V4, typename P4>
__device__ __forceinline__ void load_proxy(P4* P, float* w, uint32_t* gid) {
  const V4 q0 = P[0], q1 = P[1], q2 = P[2], q3 = P[3];
  w[0] = q0.x; w[1] = q0.y; w[2] = q0.z; w[3] = q0.w;
  w[4] = q1.x; w[5] = q1.y; w[6] = q1.z; w[7] = q1.w;
  w[8] = q2.x; w[9] = q2.y; w[10] = q2.z; w[11] = q2.w;
  *gid = __float_as_uint(q3.x);
}

/* gt_xform_dir without its "+ c3 * 0" terms.  For a finite c3 the term is
 * +0 or -0, and x + (+-0) equals x except that x = -0 may become +0.  A
 * component's sign of zero changes no outcome of the proxy test: the
 * triangle test's t does not depend on the direction, dot(n, d) = +-0
 * rejects with either sign (t becomes +-inf or NaN), and beta, gamma and
 * the conservative filters compare sums against 0 and 1, where -0 and +0
 * compare equal.  So the hits are those of gt_xform_dir. */
__device__ __forceinline__ gt_f3 xform_dir_nz(const float* m, gt_f3 d) {
  return gt_f3_make((m[0] * d.x + m[3] * d.y) + m[6] * d.z, (m[1] * d.x + m[4] * d.y) + m[7] * d.z,
                    (m[2] * d.x + m[5] * d.y) + m[8] * d.z);
}

template <int MAXK, int MODE, bool UNIFORM = false>
__device__ __forceinline__ void test_proxy(const gt_dev_scene& S, uint32_t pi, const RayState& r, float tmin,
                                           float tmax, KBuf<MAXK>& kb, uint32_t& count, Counters& st,
                                           const float* s_faces) {
  float w[12];
  uint32_t gid;
  if (UNIFORM) /* pi is the same in every lane: scalar loads */
    load_proxy<f4v>(as_const<cfloat4>(S.prox) + 4 * (size_t)pi, w, &gid);
  else
    load_proxy<float4>(reinterpret_cast<const float4*>(S.prox) + 4 * (size_t)pi, w, &gid);
  st.proxies++;
  const gt_f3 oo = gt_xform_point(w, r.o);
  const gt_f3 od = xform_dir_nz(w, r.d);
  /* conservative bounding-sphere test: every face lies inside the
   * circumsphere; the margin covers the FP32 error of the cross product */
  {
    const gt_f3 cr = gt_cross(oo, od);
    const float o1 = fabsf(oo.x) + fabsf(oo.y) + fabsf(oo.z);
    const float R = GT_ICO_CIRCUMRADIUS * 1.001f + 1e-4f * o1;
    if (gt_dot(cr, cr) > R * R * gt_dot(od, od)) return;
  }
  uint32_t cand;
  if (MODE == MODE_CLOSEST) {
    cand = (1u << GT_NUM_FACES) - 1u;
  } else {
    float mg, mn;
    gt_face_margins(oo, od, w, &mg, &mn);
    cand = gt_face_candidates(oo, od, mg, mn);
  }
  while (cand) {
    const int f = __ffs(cand) - 1;
    cand &= cand - 1u;
    const float* F = s_faces + 12 * f;
    gt_face fc;
    fc.p0 = gt_f3_make(F[0], F[1], F[2]);
    fc.e0 = gt_f3_make(F[3], F[4], F[5]);
    fc.e1 = gt_f3_make(F[6], F[7], F[8]);
    fc.n = gt_f3_make(F[9], F[10], F[11]);
    float t;
    if (!gt_tri_hit(oo, od, tmin, tmax, fc, &t)) continue;
    if (MODE != MODE_CLOSEST && !gt_front_facing(w, fc.n, r.d)) continue;
    st.hits++;
    if (MODE == MODE_COUNT) {
      count++;
    } else {
      kb.insert(t, gid);
    }
  }
}

/* One traversal of the BVH for the interval (tmin, tmax): collects the
 * k smallest (t, id) hits into kb.  Children are pushed far to near and
 * culled against the current k-th entry (with conservative slack).  The
 * loop alternates "descend until this lane holds a leaf" with "test one
 * proxy", so the lanes of a wave run the expensive proxy test together. */
template <int MAXK, int MODE>
__device__ void trace(const gt_dev_scene& S, const RayState& r, float tmin, float tmax, float slack_abs,
                      KBuf<MAXK>& kb, uint32_t& count, Counters& st, Stack& stk, const float* s_faces) {
  if (S.root == GT_NODE_EMPTY) return;
  stk.sp = 0;
  stk.push(S.root);
  const float lo_t = tmin * (1.0f - 2e-5f) - slack_abs;
  uint32_t pfirst = 0, pcount = 0;
  for (;;) {
    while (pcount == 0 && stk.sp > 0) {
      const uint32_t ref = stk.pop();
      if (ref & GT_LEAF_BIT) {
        pfirst = ref & GT_LEAF_FIRST_MASK;
        pcount = ((ref >> GT_LEAF_COUNT_SHIFT) & GT_LEAF_COUNT_MASK) + 1u;
        break;
      }
      st.nodes++;
      const float4* N = reinterpret_cast<const float4*>(S.nodes + ref);
      const float4 lx = N[0], hx = N[1], ly = N[2], hy = N[3], lz = N[4], hz = N[5];
      const uint4 ch = reinterpret_cast<const uint4*>(N)[6];
      const float kth = (MODE == MODE_COUNT) ? tmax : kb.kth();
      const float hi_t = kth * (1.0f + 2e-5f) + slack_abs;
      const float clx[4] = {lx.x, lx.y, lx.z, lx.w}, chx[4] = {hx.x, hx.y, hx.z, hx.w};
      const float cly[4] = {ly.x, ly.y, ly.z, ly.w}, chy[4] = {hy.x, hy.y, hy.z, hy.w};
      const float clz[4] = {lz.x, lz.y, lz.z, lz.w}, chz[4] = {hz.x, hz.y, hz.z, hz.w};
      const uint32_t cref[4] = {ch.x, ch.y, ch.z, ch.w};
      float tn[4];
      uint32_t rr[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float ax = (clx[c] - r.o.x) * r.inv.x, bx = (chx[c] - r.o.x) * r.inv.x;
        const float ay = (cly[c] - r.o.y) * r.inv.y, by = (chy[c] - r.o.y) * r.inv.y;
        const float az = (clz[c] - r.o.z) * r.inv.z, bz = (chz[c] - r.o.z) * r.inv.z;
        const float tnear = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), lo_t));
        const float tfar = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), hi_t));
        const bool hit = tnear <= tfar && cref[c] != GT_NODE_EMPTY;
        tn[c] = hit ? tnear : INFINITY;
        rr[c] = hit ? cref[c] : GT_NODE_EMPTY;
      }
      /* sort descending by tnear (misses = +inf first), push far to near */
#define GT_CSWAP(a, b)                                   \
  if (tn[a] < tn[b]) {                                   \
    const float t_ = tn[a]; tn[a] = tn[b]; tn[b] = t_;   \
    const uint32_t r_ = rr[a]; rr[a] = rr[b]; rr[b] = r_; \
  }
      GT_CSWAP(0, 1) GT_CSWAP(2, 3) GT_CSWAP(0, 2) GT_CSWAP(1, 3) GT_CSWAP(1, 2)
#undef GT_CSWAP
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (rr[c] != GT_NODE_EMPTY) stk.push(rr[c]);
    }
    if (pcount == 0) break;
    test_proxy<MAXK, MODE>(S, pfirst, r, tmin, tmax, kb, count, st, s_faces);
    ++pfirst;
    --pcount;
  }
}

/* Wave-wide (packet) traversal for the interval (tmin, tmax) of each lane.
 * The lanes of a wave walk the BVH together from one LDS stack of
 * (node ref, lane mask) entries: a popped node is loaded once for the wave
 * (scalar loads), each lane in the mask tests the four child boxes against
 * its own ray and k-th entry, and a child is pushed with the mask of the
 * lanes that hit it.  At a leaf, the lanes in the mask test its proxies.
 * A lane tests a proxy only if its own box test of the leaf passed, and
 * may test proxies its own traversal would have culled by the k-th entry;
 * those hits lie beyond the k-th key and the insertion rejects them, so
 * every lane ends with exactly the hits of trace().  Children are pushed
 * far to near by the entry distance of the first lane that hits them. */
/* A wave's tile as a pixel rectangle, and the frame's proxy rectangles
 * (gt_rect.h; NULL: no screen-space culling) */
struct TileRect {
  const uint32_t* rects;
  uint32_t lo, hi; /* the tile packed like a rectangle (gt_rect.h) */
};

template <int MAXK, int MODE>
__device__ void trace_packet(const gt_dev_scene& S, const RayState& r, float tmin, float tmax, float slack_abs,
                             KBuf<MAXK>& kb, uint32_t& count, Counters& st, uint32_t* wst, const float* s_faces,
                             const TileRect& tr) {
  if (S.root == GT_NODE_EMPTY) return;
  const uint64_t live = __ballot(1);
  const float lo_t = __builtin_canonicalizef(tmin * (1.0f - 2e-5f) - slack_abs);
  /* The entry to visit next is held in scalar registers when it is the
   * nearest child of the node just visited (which the LDS stack would have
   * returned next anyway): no LDS store, load and read-back on a descent. */
  uint32_t next_ref = S.root;
  uint64_t next_mask = live;
  bool have_next = true;
  int sp = 0;
  float kth = (MODE == MODE_COUNT) ? tmax : kb.kth(); /* this lane's k-th distance */
  for (;;) {
    uint32_t ref;
    uint64_t mask;
    if (have_next) {
      ref = next_ref;
      mask = next_mask;
      have_next = false;
    } else {
      if (sp == 0) break;
      --sp;
      ref = __builtin_amdgcn_readfirstlane(wst[3 * sp]);
      const uint32_t mlo = __builtin_amdgcn_readfirstlane(wst[3 * sp + 1]);
      const uint32_t mhi = __builtin_amdgcn_readfirstlane(wst[3 * sp + 2]);
      mask = ((uint64_t)mhi << 32) | mlo;
    }
    const bool in = __builtin_amdgcn_inverse_ballot_w64(mask); /* this lane's bit of the entry's mask */
    if (ref & GT_LEAF_BIT) {
      const uint32_t first = ref & GT_LEAF_FIRST_MASK;
      const uint32_t cnt = ((ref >> GT_LEAF_COUNT_SHIFT) & GT_LEAF_COUNT_MASK) + 1u;
      /* uniform loop, lane predicate inside: a divergent loop around the
       * k-buffer updates made the compiler copy (and spill) the buffer */
      /* Without rectangles: one dword of each of the leaf's 64-B records,
       * all requested before the first record is needed, so that the record
       * loads of the later proxies hit the scalar cache instead of waiting
       * on L2 one after another (the value is consumed after the loop).
       * With rectangles most of them are culled and the loads do not pay. */
      typedef const __attribute__((address_space(4))) uint32_t cuint;
      cuint* P = as_const<cuint>(S.prox) + 16 * (size_t)first + 12;
      uint32_t touch[GT_LEAF_MAX];
#pragma unroll
      for (uint32_t q = 0; q < GT_LEAF_MAX; ++q) touch[q] = 0;
      if (!tr.rects) {
#pragma unroll
        for (uint32_t q = 0; q < GT_LEAF_MAX; ++q) touch[q] = P[16 * (q < cnt ? q : cnt - 1u)];
      }
      /* The leaf's proxies whose screen rectangle meets the tile; the others
       * fail the bounding-sphere stage in every lane.  The leaf's eight
       * entries come in one 64-B scalar load (the table holds GT_LEAF_MAX
       * entries past the last proxy, and scalar loads need 4-B alignment
       * only); the empty asm holds them in registers ahead of the
       * count-dependent tests, so the load is not split into the branches
       * and waited on once. */
      uint32_t todo = (2u << (cnt - 1u)) - 1u;
      if (tr.rects) {
        typedef uint32_t v16u __attribute__((ext_vector_type(16)));
        typedef const __attribute__((address_space(4))) v16u __attribute__((aligned(4))) cv16u;
        const v16u W16 = *as_const<cv16u>(tr.rects + 2 * (size_t)first);
        uint32_t w[2 * GT_LEAF_MAX];
#pragma unroll
        for (uint32_t q = 0; q < 2 * GT_LEAF_MAX; ++q) w[q] = W16[q];
        asm volatile("" : "+s"(w[0]), "+s"(w[1]), "+s"(w[2]), "+s"(w[3]), "+s"(w[4]), "+s"(w[5]), "+s"(w[6]),
                     "+s"(w[7]), "+s"(w[8]), "+s"(w[9]), "+s"(w[10]), "+s"(w[11]), "+s"(w[12]), "+s"(w[13]),
                     "+s"(w[14]), "+s"(w[15]));
        uint32_t m = 0;
#define GT_RQ(q) m |= gt_rect_meets(tr.lo, tr.hi, w[2 * q], w[2 * q + 1]) << q
        switch (cnt) { /* the leaf's own entries only (3.2 per leaf at 1M) */
          case 8: GT_RQ(7); [[fallthrough]];
          case 7: GT_RQ(6); [[fallthrough]];
          case 6: GT_RQ(5); [[fallthrough]];
          case 5: GT_RQ(4); [[fallthrough]];
          case 4: GT_RQ(3); [[fallthrough]];
          case 3: GT_RQ(2); [[fallthrough]];
          case 2: GT_RQ(1); [[fallthrough]];
          default: GT_RQ(0);
        }
#undef GT_RQ
        todo &= m;
      }
      for (; todo; todo &= todo - 1u)
        if (in) test_proxy<MAXK, MODE, true>(S, first + (uint32_t)__builtin_ctz(todo), r, tmin, tmax, kb, count, st,
                                             s_faces);
      asm volatile("" : : "s"(touch[0]), "s"(touch[1]), "s"(touch[2]), "s"(touch[3]), "s"(touch[4]),
                   "s"(touch[5]), "s"(touch[6]), "s"(touch[7]));
      if (MODE != MODE_COUNT) kth = kb.kth();
    } else {
      if (in) st.nodes++;
      typedef const __attribute__((address_space(4))) float cfloat;
      cfloat* N = as_const<cfloat>(S.nodes + ref);
      const float hi_t = __builtin_canonicalizef(kth * (1.0f + 2e-5f) + slack_abs);
      uint32_t key[4], cref[4];
      uint64_t hm[4];
      /* the lanes of the entry's mask whose slab interval for child c is
       * not empty (ballot of the compare alone, masked on the scalar
       * unit), and the push key: the entry distance of the first such
       * lane as an unsigned key that orders like the float (wave-uniform,
       * so the sort below runs on the scalar unit) */
      auto child = [&](int c, float tnear, float tfar) {
        const uint64_t filled = 0ull - (uint64_t)(cref[c] != GT_NODE_EMPTY); /* all ones or zero */
        hm[c] = __builtin_amdgcn_ballot_w64(tnear <= tfar) & mask & filled;
        const int fl = hm[c] ? __ffsll((unsigned long long)hm[c]) - 1 : 0;
        const uint32_t b = (uint32_t)__builtin_amdgcn_readlane(__float_as_int(tnear), fl);
        key[c] = hm[c] ? (b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u)) : 0xffffffffu;
      };
      /* child c: row r of the node holds axis r of the four child boxes */
#pragma unroll
      for (int c = 0; c < 4; ++c) cref[c] = __float_as_uint(N[24 + c]);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float ax = (N[c] - r.o.x) * r.inv.x, bx = (N[4 + c] - r.o.x) * r.inv.x;
        const float ay = (N[8 + c] - r.o.y) * r.inv.y, by = (N[12 + c] - r.o.y) * r.inv.y;
        const float az = (N[16 + c] - r.o.z) * r.inv.z, bz = (N[20 + c] - r.o.z) * r.inv.z;
        child(c, fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), lo_t)),
              fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), hi_t)));
      }
      /* sort (key, child) descending, then push far to near */
#define GT_OSWAP(a, b)                                                          \
    {                                                                             \
      const bool sw_ = key[a] < key[b];                                           \
      const uint32_t k_ = key[a], r_ = cref[a];                                   \
      const uint64_t m_ = hm[a];                                                  \
      key[a] = sw_ ? key[b] : k_; cref[a] = sw_ ? cref[b] : r_; hm[a] = sw_ ? hm[b] : m_; \
      key[b] = sw_ ? k_ : key[b]; cref[b] = sw_ ? r_ : cref[b]; hm[b] = sw_ ? m_ : hm[b]; \
    }
      GT_OSWAP(0, 1) GT_OSWAP(2, 3) GT_OSWAP(0, 2) GT_OSWAP(1, 3) GT_OSWAP(1, 2)
#undef GT_OSWAP
      /* misses (key ~0) sort first, so slot 3 holds the nearest hit child
       * if any child was hit: it is visited next, the others go on the
       * stack far to near */
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const uint64_t m = hm[c];
        if (m) {
          wst[3 * sp] = cref[c];
          wst[3 * sp + 1] = (uint32_t)m;
          wst[3 * sp + 2] = (uint32_t)(m >> 32);
          ++sp;
        }
      }
      if (hm[3]) {
        next_ref = cref[3];
        next_mask = hm[3];
        have_next = true;
      }
    }
  }
}

/* shToRadiance + computeRadiance (GS_shader.ts:138-190): the 16 real-SH
 * basis values with the constants of GS_shader.ts:39-52, then
 * L_c = max(0, 0.5 + sum_j b_j sh[j][c]) with the coefficients streamed
 * 16 B at a time (pixels are checked against FP64 within 1e-4, so the FP32
 * summation order is free; streaming keeps register pressure low). */
__device__ __forceinline__ gt_f3 radiance(const float* __restrict__ shp, uint32_t deg, gt_f3 dir) {
  const float4* S4 = reinterpret_cast<const float4*>(shp);
  const float x = dir.x, y = dir.y, z = dir.z;
  const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z;
  float b[16];
  b[0] = 0.28209479177387814f;
  b[1] = -0.4886025119029199f * y;
  b[2] = 0.4886025119029199f * z;
  b[3] = -0.4886025119029199f * x;
  b[4] = 1.0925484305920792f * xy;
  b[5] = -1.0925484305920792f * yz;
  b[6] = 0.31539156525252005f * ((2.f * zz - xx) - yy);
  b[7] = -1.0925484305920792f * xz;
  b[8] = 0.5462742152960396f * (xx - yy);
  b[9] = -0.5900435899266435f * y * (3.0f * xx - yy);
  b[10] = 2.890611442640554f * xy * z;
  b[11] = -0.4570457994644658f * y * ((4.0f * zz - xx) - yy);
  b[12] = 0.3731763325901154f * z * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
  b[13] = -0.4570457994644658f * x * ((4.0f * zz - xx) - yy);
  b[14] = 1.445305721320277f * z * (xx - yy);
  b[15] = -0.5900435899266435f * x * (xx - 3.0f * yy);
  const int nb = deg == 0 ? 1 : deg == 1 ? 4 : deg == 2 ? 9 : 16;
  float acc[3] = {0.5f, 0.5f, 0.5f};
#pragma unroll
  for (int i = 0; i < 12; ++i) {
    if (3 * nb > 4 * i) {
      const float4 v = S4[i];
      const float q[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int idx = 4 * i + k, j = idx / 3, c = idx % 3;
        if (j < nb) acc[c] += b[j] * q[k];
      }
    }
  }
  return gt_f3_make(fmaxf(acc[0], 0.0f), fmaxf(acc[1], 0.0f), fmaxf(acc[2], 0.0f));
}

__device__ __forceinline__ uint32_t find_camera(const gt_dev_frame* cams, uint32_t n, uint32_t tile) {
  uint32_t lo = 0, hi = n; /* last camera with tile_base <= tile */
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cams[mid].tile_base <= tile) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
  const uint32_t x = b & 7u, slot = b >> 3;
  const uint32_t q = nb >> 3, rem = nb & 7u;
  return x * q + (x < rem ? x : rem) + slot;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

/* One 64-ray tile of tile_w x (64 / tile_w) pixels, rendered by one wave
 * (GS_shader.ts:275-309 per lane).  BATCH: global tiles of all cameras are
 * numbered consecutively (camera c owns tiles from cams[c].tile_base). */
template <int MAXK, int MODE, bool DUMP, bool BATCH, bool PACKET>
__device__ __forceinline__ void render_tile(const gt_dev_scene& S, const gt_dev_frame& F0,
                                            const gt_dev_frame* __restrict__ cams, uint32_t n_cams,
                                            uint32_t total_tiles, uint32_t tile_w, uint32_t tile, uint32_t lane,
                                            Stack& stk, uint32_t* wst, const float* s_faces, Counters& st,
                                            const uint32_t* rects) {
  const uint32_t cam = (BATCH && tile < total_tiles) ? find_camera(cams, n_cams, tile) : 0u;
  const gt_dev_frame& F = BATCH ? cams[cam] : F0;
  const uint32_t tile_h = 64u / tile_w;
  const uint32_t lt = tile - (BATCH ? F.tile_base : 0u);
  const uint32_t px = (lt % F.tiles_x) * tile_w + (lane % tile_w);
  const uint32_t py = (lt / F.tiles_x) * tile_h + (lane / tile_w);
  const bool active = tile < total_tiles && px < F.width && py < F.height;
  TileRect tr;
  tr.rects = BATCH ? nullptr : rects;
  {
    const uint32_t tx0 = (lt % F.tiles_x) * tile_w, ty0 = (lt / F.tiles_x) * tile_h;
    tr.lo = tx0 | ty0 << 16;
    tr.hi = ((tx0 + tile_w - 1u) | (ty0 + tile_h - 1u) << 16) | GT_RECT_GUARD;
  }
  if (active) {
    const uint32_t ray = py * F.width + px;
    RayState r;
    r.o = gt_f3_make(F.to_world[12], F.to_world[13], F.to_world[14]);
    /* the origin is the same in every lane; held in VGPRs it is a legal
     * second operand of the box and transform arithmetic against scalar
     * node and proxy data (one scalar operand per VALU instruction), which
     * saves a copy per operation */
    asm volatile("" : "+v"(r.o.x), "+v"(r.o.y), "+v"(r.o.z));
    float ox = 0.5f, oy = 0.5f; /* pixel offset, jittered with antialiasing (GS_shader.ts:282-284) */
    if (F.antialias) gt_aa_offset(px, py, F.width, F.accumulated_frames, &ox, &oy);
    r.d = gt_ray_dir_off(F.to_world, F.ax, F.ay, px, py, F.width, F.height, ox, oy);
    {
      const float dx = fabsf(r.d.x) < 1e-20f ? copysignf(1e-20f, r.d.x) : r.d.x;
      const float dy = fabsf(r.d.y) < 1e-20f ? copysignf(1e-20f, r.d.y) : r.d.y;
      const float dz = fabsf(r.d.z) < 1e-20f ? copysignf(1e-20f, r.d.z) : r.d.z;
      r.inv = gt_f3_make(1.0f / dx, 1.0f / dy, 1.0f / dz);
    }
    gt_f3 L = gt_f3_make(0.0f, 0.0f, 0.0f);
    float T = 1.0f;
    uint32_t rounds = 0, n_comp = 0;
    const size_t MR = F.max_rounds;
    KBuf<MAXK> kb;
    uint32_t count = 0;
    if (is_render(MODE)) {
      const uint32_t k = F.k;
      float t_curr = F.t_min;
      uint32_t step = 0;
      /* A traversal collects the MAXK smallest hits beyond t_curr; a round
       * composites the k smallest hits beyond its t_curr.  The entries of
       * the last traversal not yet composited sit at the front of the buffer
       * (rem of them; the slots behind hold copies of the last one), so a
       * lane's round is served from its buffer when the first entry lies
       * beyond t_curr (the relaunch skips none of them) and either k entries
       * remain or the traversal found every hit (its last slot a sentinel,
       * which the copies then repeat).  The decision is the wave's: it
       * traverses when any active lane cannot serve the round, and then
       * every active lane refills (a traversal from t_curr gives the same k
       * smallest hits).  So rem is one scalar, the lanes whose buffer was
       * full are one ballot mask, the lanes stay in step, and no per-lane
       * state is added to the k-buffer's registers. */
      uint32_t rem = 0;
      uint64_t full = ~0ull;
      kb.reset_empty(F.t_max);
      while (F.T_min < T && t_curr < F.t_max) {
        if (F.early_stop == step++) break;
        const bool need = !((__builtin_amdgcn_inverse_ballot_w64(full) ? rem >= k : true) && kb.t_at(0) > t_curr);
        if (__ballot(need)) {
          kb.reset_empty(F.t_max);
          if (PACKET) trace_packet<MAXK, MODE>(S, r, t_curr, F.t_max, F.slack_abs, kb, count, st, wst, s_faces, tr);
          else trace<MAXK, MODE>(S, r, t_curr, F.t_max, F.slack_abs, kb, count, st, stk, s_faces);
          full = __ballot(kb.id_at(MAXK - 1) != GT_SENTINEL);
          rem = MAXK;
        }
        rem = __builtin_amdgcn_readfirstlane(rem > k ? rem - k : 0u);
        float t_k = F.t_max;
        bool stop = false;
        uint32_t mask = 0;
        /* one copy of the evaluation code: the k-buffer rotates by one slot
         * per entry (static register moves), so the loop need not unroll */
#pragma unroll 1
        for (uint32_t j = 0; j < k; ++j) {
          const uint32_t id = kb.id_at(0);
          if (DUMP && rounds < MR) F.hit_ids[((size_t)ray * MR + rounds) * k + j] = id;
          if (!stop) {
            if (id == GT_SENTINEL) {
              stop = true;
            } else {
              st.h_eval++;
              const float4* G = reinterpret_cast<const float4*>(S.geom) + 3 * (size_t)id;
              const float4 g0 = G[0], g1 = G[1], g2 = G[2];
              const float mu3[3] = {g0.x, g0.y, g0.z};
              const float sc3[3] = {g1.x, g1.y, g1.z};
              const float q4[4] = {g2.x, g2.y, g2.z, g2.w};
              const gt_f3 mu = gt_f3_make(g0.x, g0.y, g0.z);
              const float alpha = (float)(gt_response64(mu3, sc3, q4, r.o, r.d) * (double)g0.w);
              if (F.alpha_min < alpha) {
                gt_f3 c;
                if constexpr (MODE == MODE_RENDER_PC) {
                  const float4 c4 = reinterpret_cast<const float4*>(F.colors)[id];
                  c = gt_f3_make(c4.x, c4.y, c4.z);
                } else {
                  c = radiance(S.sh + 48 * (size_t)id, F.sh_degree, gt_normalize(gt_sub(mu, r.o)));
                }
                const float Ta = T * alpha;
                L = gt_add(L, gt_scale(Ta, c));
                T *= 1.0f - alpha;
                mask |= 1u << j;
                st.h_comp++;
                if (F.comp_ids && n_comp < F.comp_cap)
                  F.comp_ids[(size_t)(ray - F.comp_base) * F.comp_cap + n_comp] = id;
                ++n_comp;
              }
            }
          }
          if (j + 1u == k) t_k = kb.t_at(0); /* the round's k-th entry */
          kb.shift_left(1);
        }
        t_curr = stop ? F.t_max : t_k + GT_RELAUNCH_EPS;
        if (DUMP && rounds < MR) F.accept[(size_t)ray * MR + rounds] = mask;
        if (DUMP && rounds >= MR) st.overflow++;
        rounds++;
      }
    } else if (MODE == MODE_CLOSEST) {
      /* debugMode 1 (GS_shader.ts:289-296, 312-321) */
      kb.reset_empty(F.t_max);
      if (PACKET) trace_packet<MAXK, MODE>(S, r, F.t_min, F.t_max, F.slack_abs, kb, count, st, wst, s_faces, tr);
      else trace<MAXK, MODE>(S, r, F.t_min, F.t_max, F.slack_abs, kb, count, st, stk, s_faces);
      const uint32_t id = kb.id_at(MAXK - 1);
      if (id != GT_SENTINEL) L = radiance(S.sh + 48 * (size_t)id, F.sh_degree, r.d);
      rounds = 1;
      if (DUMP && MR > 0) {
        F.hit_ids[(size_t)ray * MR * F.k] = id;
        F.accept[(size_t)ray * MR] = 0;
      }
    } else {
      /* debugMode 2 (GS_shader.ts:201-228): front-face hits in (t_min, t_max) */
      kb.reset_empty(F.t_max);
      if (PACKET) trace_packet<MAXK, MODE>(S, r, F.t_min, F.t_max, F.slack_abs, kb, count, st, wst, s_faces, tr);
      else trace<MAXK, MODE>(S, r, F.t_min, F.t_max, F.slack_abs, kb, count, st, stk, s_faces);
      const uint32_t c = count;
      float cr, cg, cb;
      if (c == 0) { cr = 0.f; cg = 0.f; cb = 0.f; }
      else if (c < 20) { cr = 0.5f; cg = 0.5f; cb = 0.f; }
      else if (c < 40) { cr = 0.5f; cg = 0.f; cb = 0.5f; }
      else if (c < 60) { cr = 0.f; cg = 0.5f; cb = 0.5f; }
      else if (c < 80) { cr = 0.5f; cg = 0.f; cb = 0.f; }
      else if (c < 100) { cr = 0.f; cg = 0.5f; cb = 0.f; }
      else if (c < 150) { cr = 0.f; cg = 0.f; cb = 0.5f; }
      else { cr = 0.5f; cg = 0.5f; cb = 0.5f; }
      L = gt_f3_make(cr, cg, cb);
      rounds = 1;
      if (DUMP && MR > 0) {
        F.hit_ids[(size_t)ray * MR * F.k] = c;
        F.accept[(size_t)ray * MR] = 0;
      }
    }
    reinterpret_cast<float4*>(F.rgba)[ray] = make_float4(L.x, L.y, L.z, 1.0f);
    if (F.n_rounds) F.n_rounds[ray] = rounds;
    if (F.n_comp) F.n_comp[ray - F.comp_base] = n_comp;
    if (F.T_out) F.T_out[ray] = T;
    if (DUMP) {
      const uint32_t kk = F.k;
      const uint32_t first_unwritten = is_render(MODE) ? rounds : 1u;
      for (size_t rd = first_unwritten; rd < MR; ++rd) {
        for (uint32_t j = 0; j < kk; ++j) F.hit_ids[((size_t)ray * MR + rd) * kk + j] = GT_SENTINEL;
        F.accept[(size_t)ray * MR + rd] = 0;
      }
      if (!is_render(MODE) && MR > 0)
        for (uint32_t j = 1; j < kk; ++j) F.hit_ids[(size_t)ray * MR * kk + j] = GT_SENTINEL;
    }
    st.rounds += rounds;
    st.rays++;
  }
}

/* The tile kernel: one wave per tile, block ids remapped so that
 * consecutive tiles share an XCD (and its L2).  PERSIST: a grid sized to
 * the machine whose waves claim tiles from a global queue instead, one
 * counter per XCD over a contiguous eighth of the tiles (a wave whose XCD
 * range is exhausted takes tiles from the other ranges), so a launch over
 * many cameras of uneven size keeps every CU busy to its end.  Each tile
 * is rendered by the same code either way. */
template <int MAXK, int MODE, bool DUMP, bool STATS, bool BATCH, bool PACKET, bool PERSIST>
__global__ __launch_bounds__(GT_BLOCK, Waves<MAXK>::value) void k_forward(const gt_dev_scene S, const gt_dev_frame F0,
                                                                    const gt_dev_frame* __restrict__ cams,
                                                                    uint32_t n_cams, uint32_t total_tiles,
                                                                    uint32_t tile_w, uint32_t lds_depth,
                                                                    uint32_t tile_offset, unsigned int* work,
                                                                    const uint32_t* rects, const uint32_t* order) {
  extern __shared__ uint32_t s_stack[];
  __shared__ float s_faces[12 * GT_NUM_FACES];
  for (uint32_t i = threadIdx.x; i < 12 * GT_NUM_FACES; i += GT_BLOCK) s_faces[i] = c_faces[i];
  __syncthreads();
  Stack stk;
  stk.lds = s_stack + threadIdx.x;
  stk.sp = 0;
  /* PACKET: this wave's stack.  The wave index goes through readfirstlane
   * so that the base is a scalar: as a per-lane value the compiler kept it
   * as a 64-bit pointer, spilled it, and reloaded it from scratch on every
   * stack pop. */
  uint32_t* const wst = s_stack + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * (3u * GT_STACK_MAX);
  stk.depth = (int)lds_depth;
  const uint32_t lane = threadIdx.x & 63u;
  Counters st;
  if (!PERSIST) {
    const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
    /* the same in every lane of the wave; readfirstlane tells the compiler,
     * so that the tile's values live in scalar registers */
    const uint32_t tile = __builtin_amdgcn_readfirstlane(tile_offset + blk * 4u + (threadIdx.x >> 6));
    render_tile<MAXK, MODE, DUMP, BATCH, PACKET>(S, F0, cams, n_cams, total_tiles, tile_w, tile, lane, stk, wst,
                                                 s_faces, st, rects);
  } else {
    const uint32_t n = total_tiles - tile_offset;
    const uint32_t x0 = blockIdx.x & (GT_XCDS - 1u);
    for (uint32_t j = 0; j < GT_XCDS; ++j) {
      const uint32_t x = (x0 + j) & (GT_XCDS - 1u);
      const uint32_t lo = tile_offset + (uint32_t)((uint64_t)n * x / GT_XCDS);
      const uint32_t hi = tile_offset + (uint32_t)((uint64_t)n * (x + 1u) / GT_XCDS);
      for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(work + GT_WORK_STRIDE * x, 1u);
        t = __builtin_amdgcn_readfirstlane(t);
        if (t >= hi - lo) break; /* every wave of the grid reaches this exit once the ranges are claimed */
        /* with an order, the range's tiles come most expensive first (the
         * order permutes each XCD range within itself) */
        const uint32_t tile = order ? *as_const<const __attribute__((address_space(4))) uint32_t>(order + lo + t)
                                    : lo + t;
        render_tile<MAXK, MODE, DUMP, BATCH, PACKET>(S, F0, cams, n_cams, total_tiles, tile_w, tile, lane, stk,
                                                     wst, s_faces, st, rects);
      }
    }
  }
  if (STATS) {
    const uint32_t v[8] = {st.rays, st.rounds, st.h_eval, st.h_comp, st.proxies, st.hits, st.nodes, st.overflow};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t sum = wave_sum(v[i]);
      if (lane == 0 && sum) atomicAdd(F0.stats + i, (unsigned long long)sum);
    }
  }
}

/* ------------------------------------------------------------------------
 * Persistent renderer.  A grid sized to the machine; every lane runs a
 * small state machine over rays taken from a global work counter:
 *   TRAV  descend the BVH until it holds a leaf, then test one proxy
 *   COMP  round traversed; waits, then composites its k entries
 *   IDLE  takes the next ray (idle lanes of a wave take consecutive ray ids
 *         with one atomic, so a wave stays on neighbouring pixels)
 * Each iteration the wave votes which phase to run: a refill once
 * REFILL_MIN lanes are idle, a composite pass once COMP_MIN lanes wait (or
 * nothing traverses), else a traversal step.  Lanes never wait for the
 * slowest ray of their wave: a lane whose ray ends starts another.  Ray id g -> tile g/64 (8x8 pixels), pixel g%64;
 * tiles of all cameras of a batch are numbered consecutively.
 * Per-ray semantics are those of k_forward (GS_shader.ts:192-273).
 * --------------------------------------------------------------------- */
enum { PH_IDLE = 0, PH_TRAV = 1, PH_COMP = 2 };


template <int MAXK>
__device__ __forceinline__ void visit_node(const gt_dev_scene& S, uint32_t ref, const RayState& r, float lo_t,
                                           float hi_t, Stack& stk) {
  const float4* N = reinterpret_cast<const float4*>(S.nodes + ref);
  const float4 lx = N[0], hx = N[1], ly = N[2], hy = N[3], lz = N[4], hz = N[5];
  const uint4 ch = reinterpret_cast<const uint4*>(N)[6];
  const float clx[4] = {lx.x, lx.y, lx.z, lx.w}, chx[4] = {hx.x, hx.y, hx.z, hx.w};
  const float cly[4] = {ly.x, ly.y, ly.z, ly.w}, chy[4] = {hy.x, hy.y, hy.z, hy.w};
  const float clz[4] = {lz.x, lz.y, lz.z, lz.w}, chz[4] = {hz.x, hz.y, hz.z, hz.w};
  const uint32_t cref[4] = {ch.x, ch.y, ch.z, ch.w};
  float tn[4];
  uint32_t rr[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float ax = (clx[c] - r.o.x) * r.inv.x, bx = (chx[c] - r.o.x) * r.inv.x;
    const float ay = (cly[c] - r.o.y) * r.inv.y, by = (chy[c] - r.o.y) * r.inv.y;
    const float az = (clz[c] - r.o.z) * r.inv.z, bz = (chz[c] - r.o.z) * r.inv.z;
    const float tnear = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), lo_t));
    const float tfar = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), hi_t));
    const bool hit = tnear <= tfar && cref[c] != GT_NODE_EMPTY;
    tn[c] = hit ? tnear : INFINITY;
    rr[c] = hit ? cref[c] : GT_NODE_EMPTY;
  }
#define GT_CSWAP(a, b)                                   \
  if (tn[a] < tn[b]) {                                   \
    const float t_ = tn[a]; tn[a] = tn[b]; tn[b] = t_;   \
    const uint32_t r_ = rr[a]; rr[a] = rr[b]; rr[b] = r_; \
  }
  GT_CSWAP(0, 1) GT_CSWAP(2, 3) GT_CSWAP(0, 2) GT_CSWAP(1, 3) GT_CSWAP(1, 2)
#undef GT_CSWAP
#pragma unroll
  for (int c = 0; c < 4; ++c)
    if (rr[c] != GT_NODE_EMPTY) stk.push(rr[c]);
}

template <int MAXK, bool BATCH, bool DUMP, bool STATS>
__global__ __launch_bounds__(GT_BLOCK, GT_RENDER_WAVES) void k_render(const gt_dev_scene S, const gt_dev_frame cam0,
                                                                   const gt_dev_frame* __restrict__ cams,
                                                                   uint32_t n_cams, uint32_t total_rays,
                                                                   unsigned int* work,
                                                                   unsigned long long* stats,
                                                                   uint32_t REFILL_MIN, uint32_t COMP_MIN,
                                                                   uint32_t tile_w, uint32_t lds_depth) {
  extern __shared__ uint32_t s_stack[];
  __shared__ float s_faces[12 * GT_NUM_FACES];
  for (uint32_t i = threadIdx.x; i < 12 * GT_NUM_FACES; i += GT_BLOCK) s_faces[i] = c_faces[i];
  __syncthreads();
#define CAM(field) (BATCH ? cams[cam].field : cam0.field)
  Stack stk;
  stk.lds = s_stack + threadIdx.x;
  stk.sp = 0;
  stk.depth = (int)lds_depth;
  const uint32_t tile_h = 64u / tile_w;
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  const uint32_t k = cam0.k;
  const uint32_t off = (uint32_t)MAXK - k;
  Counters st;
  uint32_t n_rays = 0;
  uint32_t phase = PH_IDLE, cam = 0, pix = 0;
  RayState r;
  r.o = gt_f3_make(0.f, 0.f, 0.f); r.d = r.o; r.inv = r.o;
  gt_f3 L = r.o;
  float T = 1.0f, t_curr = 0.0f, t_next = 0.0f, t_max = cam0.t_max, slack = cam0.slack_abs;
  uint32_t step = 0, rounds = 0, pfirst = 0, pcount = 0;
  bool exhausted = false;
  KBuf<MAXK> kb;
  kb.reset(off, t_max);
  uint32_t dummy_count = 0;

  /* starts the next round of the current ray or finishes it (GS_shader.ts:231-234) */
  auto next_round_or_finish = [&]() {
    bool go = CAM(T_min) < T && t_curr < t_max;
    if (go && CAM(early_stop) == step++) go = false;
    if (go) {
      kb.reset(off, t_max);
      stk.sp = 0;
      if (S.root != GT_NODE_EMPTY) stk.push(S.root);
      pcount = 0;
      phase = PH_TRAV;
      return;
    }
    float4* out = reinterpret_cast<float4*>(CAM(rgba));
    out[pix] = make_float4(L.x, L.y, L.z, 1.0f);
    uint32_t* nr = CAM(n_rounds);
    if (nr) nr[pix] = rounds;
    float* To = CAM(T_out);
    if (To) To[pix] = T;
    if (DUMP) {
      const size_t MR = CAM(max_rounds);
      uint32_t* H = CAM(hit_ids);
      uint32_t* A = CAM(accept);
      for (size_t rd = rounds; rd < MR; ++rd) {
        for (uint32_t jj = 0; jj < k; ++jj) H[((size_t)pix * MR + rd) * k + jj] = GT_SENTINEL;
        A[(size_t)pix * MR + rd] = 0;
      }
    }
    st.rounds += rounds;
    phase = PH_IDLE;
  };

  for (;;) {
    const uint64_t mI = __ballot(phase == PH_IDLE);
    const uint64_t mT = __ballot(phase == PH_TRAV);
    const uint64_t mC = __ballot(phase == PH_COMP);
    /* 1. idle lanes take new rays, in batches */
    if (!exhausted && mI && ((uint32_t)__popcll(mI) >= REFILL_MIN || (mT | mC) == 0)) {
      const uint32_t n = (uint32_t)__popcll(mI);
      const int leader = __ffsll((long long)mI) - 1;
      uint32_t base = 0;
      if ((int)lane == leader) base = atomicAdd(work, n);
      base = __shfl(base, leader, 64);
      if (base + n >= total_rays) exhausted = true;
      if (phase == PH_IDLE) {
        const uint32_t g = base + (uint32_t)__popcll(mI & lt_mask);
        if (g < total_rays) {
          const uint32_t tile = g >> 6, sub = g & 63u;
          if (BATCH) cam = find_camera(cams, n_cams, tile);
          const uint32_t lt = tile - CAM(tile_base), tx = CAM(tiles_x);
          const uint32_t px = (lt % tx) * tile_w + (sub % tile_w), py = (lt / tx) * tile_h + (sub / tile_w);
          const uint32_t W = CAM(width), H = CAM(height);
          if (px < W && py < H) {
            pix = py * W + px;
            r.o = gt_f3_make(CAM(to_world[12]), CAM(to_world[13]), CAM(to_world[14]));
            float tw[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) tw[i] = CAM(to_world[i]);
            r.d = gt_ray_dir(tw, CAM(ax), CAM(ay), px, py, W, H);
            const float dx = fabsf(r.d.x) < 1e-20f ? copysignf(1e-20f, r.d.x) : r.d.x;
            const float dy = fabsf(r.d.y) < 1e-20f ? copysignf(1e-20f, r.d.y) : r.d.y;
            const float dz = fabsf(r.d.z) < 1e-20f ? copysignf(1e-20f, r.d.z) : r.d.z;
            r.inv = gt_f3_make(1.0f / dx, 1.0f / dy, 1.0f / dz);
            t_max = CAM(t_max);
            slack = CAM(slack_abs);
            L = gt_f3_make(0.f, 0.f, 0.f);
            T = 1.0f;
            t_curr = CAM(t_min);
            step = 0;
            rounds = 0;
            n_rays++;
            next_round_or_finish();
          }
        }
      }
      continue;
    }
    if ((mT | mC) == 0) {
      if (exhausted) break;
      continue;
    }

    /* 2. lanes whose round traversal ended composite the whole round together
     *    (GS_shader.ts:248-269) once enough of them wait or nothing else runs */
    if (mC && ((uint32_t)__popcll(mC) >= COMP_MIN || mT == 0)) {
      if (phase == PH_COMP) {
        bool stop = false;
        uint32_t mask = 0;
        const size_t MR = DUMP ? CAM(max_rounds) : 0;
#pragma unroll 1
        for (uint32_t jj = 0; jj < k; ++jj) {
          const uint32_t id = kb.id_at(0);
          if (DUMP && rounds < MR) CAM(hit_ids)[((size_t)pix * MR + rounds) * k + jj] = id;
          if (!stop) {
            if (id == GT_SENTINEL) {
              stop = true;
            } else {
              st.h_eval++;
              const float4* G = reinterpret_cast<const float4*>(S.geom) + 3 * (size_t)id;
              const float4 g0 = G[0], g1 = G[1], g2 = G[2];
              const float mu3[3] = {g0.x, g0.y, g0.z};
              const float sc3[3] = {g1.x, g1.y, g1.z};
              const float q4[4] = {g2.x, g2.y, g2.z, g2.w};
              const float alpha = (float)(gt_response64(mu3, sc3, q4, r.o, r.d) * (double)g0.w);
              if (CAM(alpha_min) < alpha) {
                const gt_f3 dir = gt_normalize(gt_sub(gt_f3_make(g0.x, g0.y, g0.z), r.o));
                const gt_f3 c = radiance(S.sh + 48 * (size_t)id, CAM(sh_degree), dir);
                const float Ta = T * alpha;
                L = gt_add(L, gt_scale(Ta, c));
                T *= 1.0f - alpha;
                mask |= 1u << jj;
                st.h_comp++;
              }
            }
          }
          kb.shift_left(1);
        }
        if (DUMP) {
          if (rounds < MR) CAM(accept)[(size_t)pix * MR + rounds] = mask;
          else st.overflow++;
        }
        rounds++;
        t_curr = stop ? t_max : t_next;
        next_round_or_finish();
      }
      continue;
    }

    /* 3. traversal: descend to a leaf, then test one proxy */
    if (phase == PH_TRAV) {
      const float lo_t = t_curr * (1.0f - 2e-5f) - slack;
      while (pcount == 0 && stk.sp > 0) {
        const uint32_t ref = stk.pop();
        if (ref & GT_LEAF_BIT) {
          pfirst = ref & GT_LEAF_FIRST_MASK;
          pcount = ((ref >> GT_LEAF_COUNT_SHIFT) & GT_LEAF_COUNT_MASK) + 1u;
          break;
        }
        st.nodes++;
        visit_node<MAXK>(S, ref, r, lo_t, kb.kth() * (1.0f + 2e-5f) + slack, stk);
      }
      if (pcount > 0) {
        test_proxy<MAXK, MODE_RENDER>(S, pfirst, r, t_curr, t_max, kb, dummy_count, st, s_faces);
        ++pfirst;
        --pcount;
      } else {
        /* round traversal done: t_curr = t_k + 1e-4 (GS_shader.ts:248) */
        t_next = kb.kth() + GT_RELAUNCH_EPS;
        if (off & 1u) kb.shift_left(1);
        if (off & 2u) kb.shift_left(2);
        if (off & 4u) kb.shift_left(4);
        if (off & 8u) kb.shift_left(8);
        if (off & 16u) kb.shift_left(16);
        phase = PH_COMP;
      }
    }
  }
#undef CAM
  if (STATS) {
    const uint32_t v[8] = {n_rays, st.rounds, st.h_eval, st.h_comp, st.proxies, st.hits, st.nodes, st.overflow};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const uint32_t sum = wave_sum(v[i]);
      if (lane == 0 && sum) atomicAdd(stats + i, (unsigned long long)sum);
    }
  }
}

__global__ void k_reset_counter(unsigned int* work) { *work = 0u; }

/* Per-Gaussian radiance seen from one camera origin (the direction
 * normalize(mu - o) of GS_shader.ts:264 depends on the Gaussian and the
 * origin only): the same FP32 evaluation as the render kernels. */
/* One thread per Gaussian.  The block's SH records (256 x 192 B,
 * contiguous) are first copied to LDS with coalesced 16-B loads, padded to
 * 13 float4 per record so that the per-thread reads do not conflict; a
 * thread reading its own record straight from HBM made each load
 * instruction touch 64 records 192 B apart. */
#define COLORS_PAD 13u
__global__ __launch_bounds__(256) void k_colors(const gt_dev_scene S, float ox, float oy, float oz, uint32_t deg,
                                                float4* __restrict__ out) {
  __shared__ float4 s_sh[256u * COLORS_PAD];
  const uint32_t base = blockIdx.x * 256u;
  const uint32_t cnt = S.n - base < 256u ? S.n - base : 256u;
  const float4* src = reinterpret_cast<const float4*>(S.sh) + 12u * (size_t)base;
  for (uint32_t j = threadIdx.x; j < cnt * 12u; j += 256u) s_sh[(j / 12u) * COLORS_PAD + j % 12u] = src[j];
  __syncthreads();
  const uint32_t i = base + threadIdx.x;
  if (i >= S.n) return;
  const float4 g0 = reinterpret_cast<const float4*>(S.geom)[3 * (size_t)i];
  const gt_f3 dir = gt_normalize(gt_sub(gt_f3_make(g0.x, g0.y, g0.z), gt_f3_make(ox, oy, oz)));
  const gt_f3 c = radiance(reinterpret_cast<const float*>(s_sh + threadIdx.x * COLORS_PAD), deg, dir);
  out[i] = make_float4(c.x, c.y, c.z, 0.0f);
}

__global__ void k_reset_work(unsigned int* work) {
  if (threadIdx.x < GT_WORK_WORDS) work[threadIdx.x] = 0u;
}

int device_cus();

template <int MAXK, int MODE, bool BATCH, bool PACKET, bool PERSIST>
hipError_t launch_tiles_p(const gt_dev_scene& S, const gt_dev_frame& F, const gt_dev_frame* cams, uint32_t n_cams,
                          uint32_t total_tiles, uint32_t tile_w, uint32_t lds_depth, hipStream_t st, bool dump,
                          bool stats, uint32_t tile_offset, unsigned int* work, uint32_t bpc,
                          const uint32_t* rects, const uint32_t* order) {
  dim3 grid((total_tiles - tile_offset + 3u) / 4u);
  if (PERSIST) {
    hipLaunchKernelGGL(k_reset_work, dim3(1), dim3(GT_WORK_WORDS), 0, st, work);
    const uint32_t g = (uint32_t)device_cus() * (bpc ? bpc : (uint32_t)Waves<MAXK>::value);
    grid = dim3(g < grid.x ? g : grid.x);
  }
  const dim3 block(GT_BLOCK);
  const size_t lds = PACKET ? (size_t)(GT_BLOCK / 64) * 3u * GT_STACK_MAX * sizeof(uint32_t)
                            : (size_t)lds_depth * GT_BLOCK * sizeof(uint32_t);
#define GT_ARGS S, F, cams, n_cams, total_tiles, tile_w, lds_depth, tile_offset, work, rects, order
  if (dump && stats)
    hipLaunchKernelGGL((k_forward<MAXK, MODE, true, true, BATCH, PACKET, PERSIST>), grid, block, lds, st, GT_ARGS);
  else if (dump)
    hipLaunchKernelGGL((k_forward<MAXK, MODE, true, false, BATCH, PACKET, PERSIST>), grid, block, lds, st, GT_ARGS);
  else if (stats)
    hipLaunchKernelGGL((k_forward<MAXK, MODE, false, true, BATCH, PACKET, PERSIST>), grid, block, lds, st, GT_ARGS);
  else
    hipLaunchKernelGGL((k_forward<MAXK, MODE, false, false, BATCH, PACKET, PERSIST>), grid, block, lds, st, GT_ARGS);
#undef GT_ARGS
  return hipGetLastError();
}

/* path 1: per-lane traversal; 3 (default): packet traversal; 4: packet
 * traversal in the persistent kernel */
template <int MAXK, int MODE, bool BATCH>
hipError_t launch_tiles(const gt_dev_scene& S, const gt_dev_frame& F, const gt_dev_frame* cams, uint32_t n_cams,
                        uint32_t total_tiles, uint32_t tile_w, uint32_t lds_depth, hipStream_t st, bool dump,
                        bool stats, uint32_t tile_offset, uint32_t path, unsigned int* work, uint32_t bpc,
                        const uint32_t* rects, const uint32_t* order) {
  if (path == 1)
    return launch_tiles_p<MAXK, MODE, BATCH, false, false>(S, F, cams, n_cams, total_tiles, tile_w, lds_depth, st,
                                                           dump, stats, tile_offset, work, bpc, nullptr, nullptr);
  if (path == 4)
    return launch_tiles_p<MAXK, MODE, BATCH, true, true>(S, F, cams, n_cams, total_tiles, tile_w, lds_depth, st,
                                                         dump, stats, tile_offset, work, bpc, rects, order);
  return launch_tiles_p<MAXK, MODE, BATCH, true, false>(S, F, cams, n_cams, total_tiles, tile_w, lds_depth, st, dump,
                                                        stats, tile_offset, work, bpc, rects, nullptr);
}

template <int MAXK, bool BATCH>
hipError_t launch_persistent(const gt_dev_scene& S, const gt_dev_frame& c0, const gt_dev_frame* cams, uint32_t n,
                             uint32_t total_rays, unsigned int* work, unsigned long long* stats, hipStream_t st,
                             bool dump, int grid, uint32_t refill, uint32_t comp, uint32_t tile_w,
                             uint32_t lds_depth) {
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, st, work);
  const dim3 g(grid), b(GT_BLOCK);
  const size_t lds = (size_t)lds_depth * GT_BLOCK * sizeof(uint32_t);
#define GT_ARGS S, c0, cams, n, total_rays, work, stats, refill, comp, tile_w, lds_depth
  if (dump && stats) hipLaunchKernelGGL((k_render<MAXK, BATCH, true, true>), g, b, lds, st, GT_ARGS);
  else if (dump) hipLaunchKernelGGL((k_render<MAXK, BATCH, true, false>), g, b, lds, st, GT_ARGS);
  else if (stats) hipLaunchKernelGGL((k_render<MAXK, BATCH, false, true>), g, b, lds, st, GT_ARGS);
  else hipLaunchKernelGGL((k_render<MAXK, BATCH, false, false>), g, b, lds, st, GT_ARGS);
#undef GT_ARGS
  return hipGetLastError();
}


/* ---- cost-ordered tile queue of a one-camera persistent frame ----
 * A tile's cost follows the number of proxy rectangles that cover it
 * (1080p 1M-Gaussian frame: rank correlation 0.87-0.91 with the measured
 * tile times).  Coverage is counted on a grid of cells of cell x cell tiles:
 * each of GT_COVER_PARTS workgroups adds the four corners of every
 * rectangle of its contiguous run of proxies (neighbours in the tree, so
 * few distinct cells) into a difference grid in LDS and stores it; one
 * workgroup sums the partial grids, prefix-sums them along rows and
 * columns and orders each XCD's tile range by decreasing coverage within
 * the range.  No global atomics (device-scope atomics from eight XCDs on
 * the same cells cost 1.8 ms at 1M).  The persistent kernel's ranges and its
 * stealing between them stay as they are; only the order in which tiles
 * are claimed changes. */
/* cells cx x cy; a rectangle covering cells [a, b) x [c, d) adds +1 at
 * (c, a) and (d, b), -1 at (c, b) and (d, a); corners in row cy or column
 * cx reach no cell and are skipped */
__global__ __launch_bounds__(1024) void k_cover_part(const uint2* __restrict__ rects, uint32_t n, uint32_t width,
                                                     uint32_t height, uint32_t cpw, uint32_t cph, uint32_t cx,
                                                     uint32_t cy, int* __restrict__ part) {
  __shared__ int g[GT_COVER_CELLS];
  const uint32_t cells = cx * cy;
  for (uint32_t i = threadIdx.x; i < cells; i += 1024u) g[i] = 0;
  __syncthreads();
  const uint32_t lo = (uint32_t)((uint64_t)n * blockIdx.x / gridDim.x);
  const uint32_t hi = (uint32_t)((uint64_t)n * (blockIdx.x + 1u) / gridDim.x);
  for (uint32_t i = lo + threadIdx.x; i < hi; i += 1024u) {
    const uint2 r = rects[i];
    const uint32_t x0 = r.x & 0x7fffu, y0 = r.x >> 16, x1 = r.y & 0x7fffu, y1 = (r.y >> 16) & 0x7fffu;
    if (x0 > x1 || y0 > y1 || x0 >= width || y0 >= height) continue;
    const uint32_t a = x0 / cpw, b = (x1 < width ? x1 : width - 1u) / cpw + 1u;
    const uint32_t c = y0 / cph, d = (y1 < height ? y1 : height - 1u) / cph + 1u;
    atomicAdd(&g[c * cx + a], 1);
    if (b < cx) atomicAdd(&g[c * cx + b], -1);
    if (d < cy) {
      atomicAdd(&g[d * cx + a], -1);
      if (b < cx) atomicAdd(&g[d * cx + b], 1);
    }
  }
  __syncthreads();
  int* out = part + (size_t)blockIdx.x * cells;
  for (uint32_t i = threadIdx.x; i < cells; i += 1024u) out[i] = g[i];
}

/* One workgroup: the partial grids summed into LDS, prefix sums along
 * rows and columns there, then a counting sort of the tiles on (XCD range,
 * log-scale coverage) buckets: histogram, exclusive scan, scatter, all in
 * LDS.  Tile i belongs to XCD range x = min(7, (8 i + 7) / n), the largest
 * x with n x / 8 <= i, the persistent kernel's partition. */
__global__ __launch_bounds__(1024) void k_order(const int* __restrict__ part, uint32_t cx, uint32_t cy,
                                                uint32_t cell, uint32_t tiles_x, uint32_t n,
                                                uint32_t* __restrict__ order) {
  __shared__ int g[GT_COVER_CELLS];
  __shared__ uint32_t h[GT_ORDER_BUCKETS];
  const uint32_t cells = cx * cy, t = threadIdx.x;
  for (uint32_t i = t; i < cells; i += 1024u) {
    int v = 0;
    for (uint32_t p = 0; p < GT_COVER_PARTS; ++p) v += part[(size_t)p * cells + i];
    g[i] = v;
  }
  for (uint32_t i = t; i < GT_ORDER_BUCKETS; i += 1024u) h[i] = 0u;
  __syncthreads();
  for (uint32_t y = t; y < cy; y += 1024u) {
    int acc = 0;
    for (uint32_t x = 0; x < cx; ++x) g[y * cx + x] = (acc += g[y * cx + x]);
  }
  __syncthreads();
  for (uint32_t x = t; x < cx; x += 1024u) {
    int acc = 0;
    for (uint32_t y = 0; y < cy; ++y) g[y * cx + x] = (acc += g[y * cx + x]);
  }
  __syncthreads();
  auto bucket = [&](uint32_t i) {
    const uint32_t r = (8u * i + 7u) / n;
    const int cov = g[((i / tiles_x) / cell) * cx + (i % tiles_x) / cell];
    const float l = log2f((float)(cov > 0 ? cov : 0) + 1.0f) * (float)(GT_ORDER_LEVELS / 16u);
    const uint32_t q = l < (float)(GT_ORDER_LEVELS - 1u) ? (uint32_t)l : GT_ORDER_LEVELS - 1u;
    return (r < GT_XCDS ? r : GT_XCDS - 1u) * GT_ORDER_LEVELS + (GT_ORDER_LEVELS - 1u - q);
  };
  for (uint32_t i = t; i < n; i += 1024u) atomicAdd(&h[bucket(i)], 1u);
  __syncthreads();
  { /* exclusive scan of the buckets: PER per thread, then the partial sums */
    constexpr uint32_t PER = GT_ORDER_BUCKETS / 1024u;
    __shared__ uint32_t ps[1024];
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
      v[j] = h[t * PER + j];
      sum += v[j];
    }
    ps[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
      const uint32_t add = t >= off ? ps[t - off] : 0u;
      __syncthreads();
      ps[t] += add;
      __syncthreads();
    }
    uint32_t base = ps[t] - sum;
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
      h[t * PER + j] = base;
      base += v[j];
    }
  }
  __syncthreads();
  for (uint32_t i = t; i < n; i += 1024u) order[atomicAdd(&h[bucket(i)], 1u)] = i;
}

/* ws: gt_order_words(tiles_x, tiles_y) words; returns the order in *order */
hipError_t order_tiles(uint32_t n_prox, const gt_dev_frame& F, uint32_t tile_w, uint32_t tiles_y, uint32_t n,
                       const uint2* rects, uint32_t* ws, hipStream_t st, const uint32_t** order) {
  const uint32_t tiles_x = F.tiles_x;
  const uint32_t cell = gt_cover_cell(tiles_x, tiles_y);
  const uint32_t cx = (tiles_x + cell - 1u) / cell, cy = (tiles_y + cell - 1u) / cell, cells = cx * cy;
  int* part = reinterpret_cast<int*>(ws);
  uint32_t* ord = ws + (size_t)GT_COVER_PARTS * cells;
  hipLaunchKernelGGL(k_cover_part, dim3(GT_COVER_PARTS), dim3(1024), 0, st, rects, n_prox, F.width, F.height,
                     tile_w * cell, (64u / tile_w) * cell, cx, cy, part);
  hipLaunchKernelGGL(k_order, dim3(1), dim3(1024), 0, st, part, cx, cy, cell, tiles_x, n, ord);
  *order = ord;
  return hipGetLastError();
}

int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  return cus;
}

std::mutex g_init_mu;
bool g_faces_ready[64];

/* The frame's proxy rectangles (gt_rect.h), GT_LEAF_MAX "empty" entries past
 * the last proxy so that a leaf's eight entries are read without a bound. */
__global__ __launch_bounds__(256) void k_rects(const float* __restrict__ prox, uint32_t n, const gt_dev_frame F,
                                               uint2* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= (uint32_t)GT_RECT_ENTRIES(n)) return;
  uint32_t r[2] = {GT_RECT_EMPTY_LO, GT_RECT_EMPTY_HI};
  if (i < n) {
    const float4* P = reinterpret_cast<const float4*>(prox) + 4 * (size_t)i;
    const float4 a = P[0], b = P[1], c = P[2];
    const float w[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
    gt_proxy_rect(w, F.to_world, F.ax, F.ay, F.width, F.height, F.t_min, r);
  }
  out[i] = make_uint2(r[0], r[1]);
}


hipError_t ensure_faces() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (dev >= 0 && dev < 64 && g_faces_ready[dev]) return hipSuccess;
  float f[12 * GT_NUM_FACES];
  for (int i = 0; i < GT_NUM_FACES; ++i) {
    const gt_face fc = gt_ico_face(i);
    const float v[12] = {fc.p0.x, fc.p0.y, fc.p0.z, fc.e0.x, fc.e0.y, fc.e0.z,
                         fc.e1.x, fc.e1.y, fc.e1.z, fc.n.x,  fc.n.y,  fc.n.z};
    for (int j = 0; j < 12; ++j) f[12 * i + j] = v[j];
  }
  if (!gt_ico_literals_match()) return hipErrorInvalidValue; /* literal tables out of date */
  e = hipMemcpyToSymbol(HIP_SYMBOL(c_faces), f, sizeof(f), 0, hipMemcpyHostToDevice);
  if (e == hipSuccess && dev >= 0 && dev < 64) g_faces_ready[dev] = true;
  return e;
}

}  // namespace

extern "C" int gt_forward_prepare(void) {
  const hipError_t e = ensure_faces();
  if (e != hipSuccess) {
    gt_set_error("face table upload failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  return GT_OK;
}

static void tuning_defaults(const gt_tuning* t, uint32_t* tile_w, uint32_t* lds_depth) {
  *tile_w = (t && t->tile_w) ? t->tile_w : 8u;
  *lds_depth = (t && t->stack_lds) ? t->stack_lds : (uint32_t)GT_LDS_STACK;
}

/* Antialiasing running average (GS_shader.ts:303-307): avg = mix(avg, img,
 * 1 / (frames + 1)) per RGB channel, A = 1.  A separate pass so that the
 * render kernel's register allocation is that of the plain frame. */
__global__ __launch_bounds__(256) void k_aa_mix(const float4* __restrict__ img, float4* __restrict__ avg, uint32_t n,
                                                uint32_t frames) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const float4 c = img[i], a = avg[i];
  avg[i] = make_float4(gt_aa_mix(a.x, c.x, frames), gt_aa_mix(a.y, c.y, frames), gt_aa_mix(a.z, c.z, frames), 1.0f);
}

extern "C" int gt_launch_aa_mix(const float* img, float* avg, uint32_t n_pixels, uint32_t frames, void* stream) {
  if (n_pixels == 0) return GT_OK;
  k_aa_mix<<<(n_pixels + 255u) / 256u, 256, 0, (hipStream_t)stream>>>(reinterpret_cast<const float4*>(img),
                                                                       reinterpret_cast<float4*>(avg), n_pixels, frames);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    gt_set_error("k_aa_mix launch failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  return GT_OK;
}

extern "C" int gt_launch_colors(const gt_dev_scene* S, const gt_dev_frame* F, float* out, void* stream) {
  if (S->n == 0) return GT_OK;
  hipLaunchKernelGGL(k_colors, dim3((S->n + 255u) / 256u), dim3(256), 0, (hipStream_t)stream, *S, F->to_world[12],
                     F->to_world[13], F->to_world[14], F->sh_degree, reinterpret_cast<float4*>(out));
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    gt_set_error("colour launch failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  return GT_OK;
}

extern "C" int gt_launch_forward(const gt_dev_scene* S, const gt_dev_frame* F, const gt_dev_frame* cams_dev,
                                 uint32_t n_cams, uint32_t total_tiles, const gt_tuning* tune, void* stream,
                                 uint32_t tile_offset, unsigned int* work, uint32_t* rect_buf, size_t rect_words) {
  if (tile_offset >= total_tiles) return GT_OK;
  hipError_t e = ensure_faces();
  if (e != hipSuccess) {
    gt_set_error("face table upload failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  if ((F->hit_ids == nullptr) != (F->accept == nullptr)) {
    gt_set_error("hit_ids and accept must be requested together");
    return GT_E_INVALID;
  }
  const bool dump = F->hit_ids != nullptr && F->max_rounds > 0;
  const bool stats = F->stats != nullptr;
  const bool batch = n_cams > 1;
  /* default (path 0): packet traversal in the persistent kernel when work
   * counters are given, else in the tile kernel; path 1: per-lane
   * traversal; 3: packet traversal, tile kernel; 4: packet traversal,
   * persistent kernel */
  const uint32_t path = (!tune || tune->path == 0) ? (work ? 4u : 3u) : tune->path;
  if (path == 4 && !work) {
    gt_set_error("the persistent packet kernel (tuning path 4) needs its work counters");
    return GT_E_INVALID;
  }
  /* 0: one resident block per wave slot (Waves<MAXK> per SIMD) */
  const uint32_t bpc = tune && tune->blocks_per_cu ? tune->blocks_per_cu : 0u;
  uint32_t tile_w, depth;
  tuning_defaults(tune, &tile_w, &depth);
  hipStream_t st = (hipStream_t)stream;
  /* one camera, packet traversal: the frame's proxy rectangles first (not
   * for batches: in a 64-camera batch of views from outside the scene one
   * table per camera cost more than it culled) */
  const uint32_t* rects = nullptr;
  if (rect_buf && !batch && path != 1 && gt_rects_wanted(tune, F->width, F->height, S->n_prox)) {
    const uint32_t n = (uint32_t)GT_RECT_ENTRIES(S->n_prox);
    hipLaunchKernelGGL(k_rects, dim3((n + 255u) / 256u), dim3(256), 0, st, S->prox, S->n_prox, *F,
                       reinterpret_cast<uint2*>(rect_buf));
    e = hipGetLastError();
    if (e != hipSuccess) {
      gt_set_error("proxy rectangle launch failed: %s", hipGetErrorString(e));
      return GT_E_HIP;
    }
    rects = rect_buf;
  }
  /* persistent launch of a whole frame: its tile queue in order of
   * decreasing cost, estimated by the rectangles' coverage of each tile */
  const uint32_t* order = nullptr;
  if (rects && path == 4 && tile_offset == 0) {
    const uint32_t tiles_y = total_tiles / F->tiles_x;
    const size_t rw = 2 * GT_RECT_ENTRIES(S->n_prox);
    if (F->tiles_x * tiles_y == total_tiles && rect_words >= rw + gt_order_words(F->tiles_x, tiles_y)) {
      e = order_tiles(S->n_prox, *F, tile_w, tiles_y, total_tiles, reinterpret_cast<const uint2*>(rect_buf),
                      rect_buf + rw, st, &order);
      if (e != hipSuccess) {
        gt_set_error("tile order launch failed: %s", hipGetErrorString(e));
        return GT_E_HIP;
      }
    }
  }
#define GT_TILES(MK, MODE)                                                                                       \
  (batch ? launch_tiles<MK, MODE, true>(*S, *F, cams_dev, n_cams, total_tiles, tile_w, depth, st, dump, stats, \
                                       tile_offset, path, work, bpc, nullptr, nullptr)                           \
         : launch_tiles<MK, MODE, false>(*S, *F, cams_dev, n_cams, total_tiles, tile_w, depth, st, dump, stats, \
                                        tile_offset, path, work, bpc, rects, order))
  if (F->debug_mode == 1) {
    e = launch_tiles<1, MODE_CLOSEST, false>(*S, *F, cams_dev, 1, total_tiles, tile_w, depth, st, dump, stats,
                                             tile_offset, path, work, bpc, rects, order);
  } else if (F->debug_mode == 2) {
    e = launch_tiles<1, MODE_COUNT, false>(*S, *F, cams_dev, 1, total_tiles, tile_w, depth, st, dump, stats,
                                           tile_offset, path, work, bpc, rects, order);
  } else {
    /* every k renders with the 32-slot buffer: one traversal serves
     * several rounds of k (1080p 1M-Gaussian frame: k 16 8.28 -> 7.96 ms,
     * k 8 10.78 -> 8.45 ms against a k-slot buffer per round) */
    e = (F->colors && !batch) ? launch_tiles<32, MODE_RENDER_PC, false>(*S, *F, cams_dev, 1, total_tiles, tile_w,
                                                                        depth, st, dump, stats, tile_offset, path,
                                                                        work, bpc, rects, order)
                              : GT_TILES(32, MODE_RENDER);
  }
#undef GT_TILES
  if (e != hipSuccess) {
    gt_set_error("forward launch failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  return GT_OK;
}

extern "C" int gt_launch_persistent(const gt_dev_scene* S, const gt_dev_frame* cam0, const gt_dev_frame* cams_dev,
                                    uint32_t n_cams, uint32_t total_tiles, unsigned int* work,
                                    unsigned long long* stats, const gt_tuning* tune, void* stream) {
  hipError_t e = ensure_faces();
  if (e != hipSuccess) {
    gt_set_error("face table upload failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  if ((cam0->hit_ids == nullptr) != (cam0->accept == nullptr)) {
    gt_set_error("hit_ids and accept must be requested together");
    return GT_E_INVALID;
  }
  if ((uint64_t)total_tiles * 64ull > 0xffffff00ull) {
    gt_set_error("%u tiles exceed the 2^32 ray ids of one launch", total_tiles);
    return GT_E_LIMIT;
  }
  const bool dump = cam0->hit_ids != nullptr && cam0->max_rounds > 0;
  const bool batch = n_cams > 1;
  const uint32_t rays = total_tiles * 64u;
  const uint32_t k = cam0->k;
  hipStream_t st = (hipStream_t)stream;
  const uint32_t bpc = tune && tune->blocks_per_cu ? tune->blocks_per_cu : GT_RENDER_WAVES;
  const int grid = device_cus() * (int)bpc;
  const uint32_t refill = tune && tune->refill_min ? tune->refill_min : 48u;
  const uint32_t comp = tune && tune->comp_min ? tune->comp_min : 48u;
  uint32_t tile_w, depth;
  tuning_defaults(tune, &tile_w, &depth);
#define GT_DISPATCH(MK)                                                                                          \
  (batch ? launch_persistent<MK, true>(*S, *cam0, cams_dev, n_cams, rays, work, stats, st, dump, grid, refill,   \
                                       comp, tile_w, depth)                                                       \
         : launch_persistent<MK, false>(*S, *cam0, cams_dev, n_cams, rays, work, stats, st, dump, grid, refill,  \
                                        comp, tile_w, depth))
  if (k <= 4) e = GT_DISPATCH(4);
  else if (k <= 8) e = GT_DISPATCH(8);
  else if (k <= 16) e = GT_DISPATCH(16);
  else e = GT_DISPATCH(32);
#undef GT_DISPATCH
  if (e != hipSuccess) {
    gt_set_error("render launch failed: %s", hipGetErrorString(e));
    return GT_E_HIP;
  }
  return GT_OK;
}
