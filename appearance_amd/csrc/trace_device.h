// Device-side wide-BVH traversal shared by the ray-query kernels (trace.hip)
// and the wavefront integrator's extend and shadow stages.
#pragma once

#include "apt_internal.h"

namespace apt {

// Traversal stack: the bottom S entries of each lane live in LDS (S is a
// template parameter, tuned per scene), deeper entries in a private overflow
// array.  An entry is (node ref, entry distance of its box).
constexpr int TRACE_OVF_STACK = 64;
constexpr int TRACE_MAX_LDS_STACK = 32;

// Work-group -> tile mapping that gives each of the 8 XCDs of an MI355X one
// contiguous range of work-groups.  The dispatcher deals work-groups to XCDs
// round robin (group i runs on XCD i % 8), so consecutive logical groups
// (neighbouring rays or pixels) share an XCD's L2.  The mapping is a bijection
// on [0, G) for any G.
__device__ __forceinline__ uint32_t xcd_block(uint32_t i, uint32_t G) {
  const uint32_t q = G >> 3, r = G & 7u;
  const uint32_t xcd = i & 7u, k = i >> 3;
  return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

struct TraceRay {
  RayPre pre;
  v3 inv;
  v3 d;  // the direction as given (read only by the two-level traversal)
};

__device__ __forceinline__ TraceRay make_trace_ray(v3 o, v3 d, float tmin, float tmax) {
  TraceRay r;
  r.pre = ray_pre(o, d, tmin, tmax);
  r.inv = mk(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
  r.d = d;
  return r;
}

// Per-lane state of the two-level traversal: the world-space ray, the prim
// base of the instance being traversed and the stack depth at which it was
// entered (-1 in the top level).  When the stack falls back to that depth the
// instance is finished and the world ray is restored.
struct InstLane {
  v3 wo, wd;
  uint32_t prim_base;
  int sp_inst;
};

// Two-level traversal widens every box's entry and exit distances by 2^-18 of
// their magnitude.  An object-space origin can lie far from a small BLAS (a
// small instance seen from afar), where the slab test's rounding, about
// 2^-22 of the distance, outgrows the BLAS's scene-relative padding; and a
// hit found in object space maps back onto the world ray only up to the
// rounding of the instance transform.  A relative margin covers both at any
// distance.
constexpr float INST_WIDEN = 3.814697265625e-06f;  // 2^-18

struct HitRec {
  float t, u, v;
  uint32_t prim;
};

// Per-lane view of the LDS stack: entry e of this lane lives at base[e*stride].
struct LdsStack {
  uint2* base;
  int stride;
};

// Entry distance of each child box, +inf for a missed or empty child.  A
// grid plane origin + q * step lies at distance q * (step * inv) + (origin -
// o) * inv along the ray: one FMA per plane after three multiplies per node
// (step * inv is exact, step being a power of two).  The rounding of
// (origin - o) * inv and of the FMA is about 2^-23 of the node's distance
// from the origin, far below the boxes' padding of 2^-14 max|coord| for
// origins within ~100x the scene extent, so the test stays conservative.
template <int W, bool ROBUST = false>
__device__ __forceinline__ void child_dists(const WNode<W>& nd, const TraceRay& ray, float tcap,
                                            float (&dist)[W]) {
  const v3 inv = ray.inv, o = ray.pre.o;
  const float ax = qstep(nd.ex) * inv.x, ay = qstep(nd.ey) * inv.y, az = qstep(nd.ez) * inv.z;
  const float bx = (nd.ox - o.x) * inv.x, by = (nd.oy - o.y) * inv.y, bz = (nd.oz - o.z) * inv.z;
  if constexpr (W == 4) {
    // A row of four children's bytes is one word.  The near plane of an axis
    // is the lo row when the direction is positive (a >= 0), else the hi
    // row: q -> fma(q, a, b) is monotonic, so picking the rows once per node
    // gives every child the same entry and exit distances as a min and a max
    // per axis (3 % faster on the bvh_stress rays; 2 and 8 children keep the
    // min / max, whose row selection costs registers).
    uint32_t q32[6];
    __builtin_memcpy(q32, nd.q, sizeof(q32));
    const bool px = ax >= 0.0f, py = ay >= 0.0f, pz = az >= 0.0f;
    const uint32_t nx = px ? q32[QLO_X] : q32[QHI_X], fx = px ? q32[QHI_X] : q32[QLO_X];
    const uint32_t ny = py ? q32[QLO_Y] : q32[QHI_Y], fy = py ? q32[QHI_Y] : q32[QLO_Y];
    const uint32_t nz = pz ? q32[QLO_Z] : q32[QHI_Z], fz = pz ? q32[QHI_Z] : q32[QLO_Z];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int sh = 8 * k;
      const float tnx = __builtin_fmaf((float)((nx >> sh) & 0xFFu), ax, bx);
      const float tfx = __builtin_fmaf((float)((fx >> sh) & 0xFFu), ax, bx);
      const float tny = __builtin_fmaf((float)((ny >> sh) & 0xFFu), ay, by);
      const float tfy = __builtin_fmaf((float)((fy >> sh) & 0xFFu), ay, by);
      const float tnz = __builtin_fmaf((float)((nz >> sh) & 0xFFu), az, bz);
      const float tfz = __builtin_fmaf((float)((fz >> sh) & 0xFFu), az, bz);
      float tn, tf;
      if (ROBUST) {
        const float bn = fmaxf(fmaxf(tnx, tny), tnz), bf = fminf(fminf(tfx, tfy), tfz);
        tn = fmaxf(__builtin_fmaf(-fabsf(bn), INST_WIDEN, bn), ray.pre.tmin);
        tf = fminf(__builtin_fmaf(fabsf(bf), INST_WIDEN, bf), tcap);
      } else {
        tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, ray.pre.tmin));
        tf = fminf(fminf(tfx, tfy), fminf(tfz, tcap));
      }
      dist[k] = (tn <= tf && nd.child[k] != REF_EMPTY) ? tn : INFINITY;
    }
  } else {
#pragma unroll
    for (int k = 0; k < W; ++k) {
      float t0x = __builtin_fmaf((float)nd.q[QLO_X][k], ax, bx), t1x = __builtin_fmaf((float)nd.q[QHI_X][k], ax, bx);
      float t0y = __builtin_fmaf((float)nd.q[QLO_Y][k], ay, by), t1y = __builtin_fmaf((float)nd.q[QHI_Y][k], ay, by);
      float t0z = __builtin_fmaf((float)nd.q[QLO_Z][k], az, bz), t1z = __builtin_fmaf((float)nd.q[QHI_Z][k], az, bz);
      float tn, tf;
      if (ROBUST) {
        const float bn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fminf(t0z, t1z));
        const float bf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
        tn = fmaxf(__builtin_fmaf(-fabsf(bn), INST_WIDEN, bn), ray.pre.tmin);
        tf = fminf(__builtin_fmaf(fabsf(bf), INST_WIDEN, bf), tcap);
      } else {
        tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), ray.pre.tmin));
        tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tcap));
      }
      dist[k] = (tn <= tf && nd.child[k] != REF_EMPTY) ? tn : INFINITY;
    }
  }
}

// Compare-exchange without a branch: two selects per array.
template <int W>
__device__ __forceinline__ void cswap(float (&d)[W], uint32_t (&r)[W], int a, int b) {
  const bool sw = d[b] < d[a];
  const float da = d[a], db = d[b];
  const uint32_t ra = r[a], rb = r[b];
  d[a] = sw ? db : da;
  d[b] = sw ? da : db;
  r[a] = sw ? rb : ra;
  r[b] = sw ? ra : rb;
}

// Ascending sort of (dist, ref) by dist: optimal networks for 2, 4 and 8.
template <int W>
__device__ __forceinline__ void sort_children(float (&d)[W], uint32_t (&r)[W]) {
  if constexpr (W == 2) {
    cswap<W>(d, r, 0, 1);
  } else if constexpr (W == 4) {
    cswap<W>(d, r, 0, 1); cswap<W>(d, r, 2, 3); cswap<W>(d, r, 0, 2);
    cswap<W>(d, r, 1, 3); cswap<W>(d, r, 1, 2);
  } else {
    cswap<W>(d, r, 0, 2); cswap<W>(d, r, 1, 3); cswap<W>(d, r, 4, 6); cswap<W>(d, r, 5, 7);
    cswap<W>(d, r, 0, 4); cswap<W>(d, r, 1, 5); cswap<W>(d, r, 2, 6); cswap<W>(d, r, 3, 7);
    cswap<W>(d, r, 0, 1); cswap<W>(d, r, 2, 3); cswap<W>(d, r, 4, 5); cswap<W>(d, r, 6, 7);
    cswap<W>(d, r, 2, 4); cswap<W>(d, r, 3, 5);
    cswap<W>(d, r, 1, 4); cswap<W>(d, r, 3, 6);
    cswap<W>(d, r, 1, 2); cswap<W>(d, r, 3, 4); cswap<W>(d, r, 5, 6);
  }
}

// Push/pop on a lane's stack: entries below S live in LDS at lds[e * stride],
// deeper ones in the private array `ovf`.  sp, lds and ovf are separate
// locals of the traversal loop so the stack pointer stays in a register and
// the LDS accesses compile to ds_* instructions.
#define APT_STACK_PUSH(ref_, t_)                                                 \
  do {                                                                           \
    const uint2 e_ = make_uint2((ref_), f2u(t_));                                \
    if (sp < S) lds[sp * stride] = e_;                                           \
    else if (sp < S + TRACE_OVF_STACK) ovf[sp - S] = e_;                         \
    if (sp < S + TRACE_OVF_STACK) ++sp;                                          \
  } while (0)
// The empty asm keeps the two branches distinct so the compiler does not fuse
// them into one generic (flat) load through a selected pointer.
__device__ __forceinline__ uint2 stack_pop(int& sp, const uint2* lds, int stride, const uint2* ovf,
                                           int S) {
  --sp;
  uint2 e;
  if (sp < S) {
    e = lds[sp * stride];
  } else {
    e = ovf[sp - S];
    asm volatile("" : "+v"(e.x), "+v"(e.y));
  }
  return e;
}
#define APT_STACK_POP() stack_pop(sp, lds, stride, ovf, S)

// Closest hit: the lexicographic minimum of (t, prim) over all triangles that
// pass the watertight test with tmin < t < tmax.  A node is entered while its
// padded box starts no later than the best t so far, so equal-t candidates are
// still visited and the tie goes to the lower prim.  One ray per lane, in the
// single-step schedule (step_fetch / step_item, defined below).
// With INST the BVH is a two-level scene: nodes and tris are its arenas,
// root the TLAS root and insts its instance records.
template <int W, int S, bool INST = false>
__device__ __forceinline__ HitRec trace_closest(const WNode<W>* __restrict__ nodes,
                                                const float4* __restrict__ tris, uint32_t root,
                                                const TraceRay& ray, LdsStack st,
                                                const InstRec* __restrict__ insts = nullptr);
// Any hit: true as soon as one triangle passes with tmin < t < tmax
// (the reference's terminate-on-first-hit shadow query, helpers/trace.wgsl:20-26).
template <int W, int S, bool INST = false>
__device__ __forceinline__ bool trace_any(const WNode<W>* __restrict__ nodes,
                                          const float4* __restrict__ tris, uint32_t root,
                                          const TraceRay& ray, LdsStack st,
                                          const InstRec* __restrict__ insts = nullptr);

// Host-side dispatch over the supported LDS stack depths.
#define APT_FOR_STACK_DEPTH(S, ...) \
  switch (S) {                      \
    case 8: { constexpr int S_ = 8; __VA_ARGS__; } break;   \
    case 12: { constexpr int S_ = 12; __VA_ARGS__; } break; \
    default: { constexpr int S_ = 16; __VA_ARGS__; } break; \
  }

// 256-thread blocks that fill the device at 8 waves per SIMD (8 blocks per
// CU), the grid of the persistent traversal kernels; blocks that do not fit
// start late, find the ray pool drained and exit.
inline uint32_t pool_blocks() {
  static uint32_t g = 0;
  if (!g) {
    int dev = 0;
    hipDeviceProp_t p;
    g = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
            ? 8u * (uint32_t)p.multiProcessorCount
            : 2048u;
  }
  return g;
}

// The ray pool of a persistent launch is split into POOL_PARTS contiguous
// ranges, two per XCD, each with its own counter (POOL_STRIDE words apart, so
// the counters sit on separate 64-byte lines).  A wave starts on range
// b % POOL_PARTS of its work-group b, which belongs to the wave's own XCD
// (work-groups are dealt to the 8 XCDs round robin, so block b runs on XCD
// b % 8), and moves on to the next range when that one runs dry: neighbouring
// rays are traced by one XCD and share its L2, no XCD idles before the whole
// pool is drained, and each counter takes the atomic adds of half an XCD's
// waves (one counter per XCD measured 7 % slower: the adds to one address
// queue).
__device__ __forceinline__ uint32_t pool_part_lo(uint32_t n, uint32_t part) {
  return (uint32_t)(((uint64_t)n * part) / POOL_PARTS);
}

// Zeroes the counters of a pool; call with tid = 0 .. POOL_PARTS-1.
__device__ __forceinline__ void pool_zero(uint32_t* ctr, uint32_t tid) {
  if (tid < POOL_PARTS) ctr[tid * POOL_STRIDE] = 0u;
}

struct PoolCursor {
  uint32_t part;  // range the wave takes rays from
  uint32_t left;  // ranges not yet found empty
};

__device__ __forceinline__ PoolCursor pool_start() {
  return PoolCursor{blockIdx.x % POOL_PARTS, POOL_PARTS};
}

// Called by every lane of the wave: once at least `refill` lanes are idle (or
// all are), the idle lanes take consecutive rays of the current range; a wave
// with no ray left moves through the ranges until it gets rays or all are
// empty.  Returns the lane's new ray index, or ~0u.
__device__ __forceinline__ uint32_t pool_take(uint32_t* ctr, uint32_t n, PoolCursor& pc, bool active,
                                              uint32_t refill) {
  const unsigned long long idle = __ballot(!active);
  if (idle == 0ull) return ~0u;
  const uint32_t nidle = (uint32_t)__popcll(idle);
  const bool all_idle = nidle == 64u;
  if (nidle < refill && !all_idle) return ~0u;
  const uint32_t lane = __lane_id();
  const int leader = __ffsll(idle) - 1;
  uint32_t got = ~0u;
  while (pc.left != 0u) {
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(ctr + pc.part * POOL_STRIDE, nidle);
    // the cursor and the range are the same in every lane: keep them scalar
    base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
    const uint32_t lo = pool_part_lo(n, pc.part), len = pool_part_lo(n, pc.part + 1u) - lo;
    if (!active) {
      const uint32_t r = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
      if (r < len) got = lo + r;
    }
    if (base + nidle >= len) {
      pc.part = (uint32_t)__builtin_amdgcn_readfirstlane((int)((pc.part + 1u) % POOL_PARTS));
      pc.left = (uint32_t)__builtin_amdgcn_readfirstlane((int)(pc.left - 1u));
    }
    if (base < len || !all_idle) break;  // rays taken, or lanes still busy
  }
  return got;
}

// The first nodes in breadth-first order -- the top levels, which every ray
// visits -- are staged in LDS by each work-group of a pool kernel: 2 KB, i.e.
// 42 / 32 / 21 nodes of width 2 / 4 / 8, which with the 8-entry stack keeps
// eight 256-thread work-groups per CU within the 160 KB of LDS.
constexpr uint32_t TOP_BYTES = 2048;
template <int W>
constexpr uint32_t top_nodes() {
  return TOP_BYTES / (uint32_t)sizeof(WNode<W>);
}

// Copies min(n_nodes, top_nodes<W>()) nodes into `top` (TOP_BYTES of LDS)
// and returns that count; every thread of the work-group must call it.
template <int W>
__device__ __forceinline__ uint32_t stage_top(const WNode<W>* __restrict__ nodes, uint32_t n_nodes,
                                              float4* top) {
  const uint32_t k = n_nodes < top_nodes<W>() ? n_nodes : top_nodes<W>();
  const float4* src = reinterpret_cast<const float4*>(nodes);
  const uint32_t n4 = k * (uint32_t)(sizeof(WNode<W>) / 16);
  for (uint32_t i = threadIdx.x; i < n4; i += blockDim.x) top[i] = src[i];
  __syncthreads();
  return k;
}

// LDS entries per lane of a pool kernel: the S-entry stack plus one entry
// that holds the (u, v) of the best hit so far (rarely written, read once per
// ray, and kept out of the registers the fetch needs).
constexpr int pool_lds_entries(int S) { return S + 1; }

// Single-step traversal.  Every iteration a lane fetches one item and
// processes it: a node (its quantized boxes and child refs, at most 96 B), or
// the next one or two triangles of the leaf it is in (48 or 96 B).  Both land
// in the same six float4 registers and every lane's fetch of an iteration is
// in flight at once: one memory round trip per iteration, whatever mix of
// nodes and leaves the wave holds.  (A while-while loop runs inner nodes
// until every lane of the wave has reached a leaf, so lanes that get there
// early wait a round trip per step of the others.)  A ray visits children
// nearest first, pushes the others far to near and drops popped entries whose
// box starts beyond its best hit; `cur` is the node to fetch next or the leaf
// ref of the triangles still to test (first advanced and count lowered as the
// leaf is consumed).
constexpr int STEP_NQ = 6;  // float4 per fetch: a node or two triangles

// With INST (a two-level scene) a leaf ref met in the top level (in_top) is an
// instance: its 64-byte record is the item fetched.
template <int W, bool INST = false>
__device__ __forceinline__ void step_fetch(const WNode<W>* __restrict__ nodes, const float4* __restrict__ tris,
                                           const float4* top, uint32_t ktop, uint32_t cur,
                                           float4 (&q)[STEP_NQ], const InstRec* __restrict__ insts = nullptr,
                                           bool in_top = false) {
  constexpr int N4 = (int)(sizeof(WNode<W>) / 16);  // float4 per node: 3, 4, 6
  static_assert(N4 <= STEP_NQ, "a node must fit the fetch registers");
  const bool in_leaf = ref_is_leaf(cur);
  if (INST && in_leaf && in_top) {
    const float4* p = reinterpret_cast<const float4*>(insts + ref_leaf_first(cur));
#pragma unroll
    for (int k = 0; k < 4; ++k) q[k] = p[k];
    asm volatile("" ::: "memory");
    return;
  }
  const bool two = in_leaf && ref_leaf_count(cur) >= 2u;
  if (!in_leaf && cur < ktop) {
#pragma unroll
    for (int k = 0; k < N4; ++k) q[k] = top[cur * N4 + k];
  } else {
    const float4* p = in_leaf ? tris + 3 * (size_t)ref_leaf_first(cur)
                              : reinterpret_cast<const float4*>(nodes + cur);
#pragma unroll
    for (int k = 0; k < 3; ++k) q[k] = p[k];
    if ((!in_leaf && N4 > 3) || two) {
#pragma unroll
      for (int k = 3; k < STEP_NQ; ++k) q[k] = (k < N4 || two) ? p[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // keep all of the fetch's loads ahead of the processing, so they are in
    // flight together
    asm volatile("" ::: "memory");
  }
}

// Processes the item step_fetch loaded for `cur` and moves `cur` on; returns
// true when the ray is finished (stack empty, or an any-hit occluder found).
// The (u, v) of an improved closest hit go to *uv when UV_LDS, else to hu, hv.
// With INST: an instance record maps the ray into the instance's object space
// and continues at its BLAS root; triangle prims are offset by the instance's
// prim base; when the stack falls back to the depth at which the instance was
// entered, the world ray is restored.
template <int W, int S, bool ANY, bool UV_LDS, bool INST = false>
__device__ __forceinline__ bool step_item(const float4 (&q)[STEP_NQ], TraceRay& ray, uint32_t& cur,
                                          float& best_t, uint32_t& best_prim, float& hu, float& hv, uint2* uv,
                                          bool& occluded, uint2* const lds, const int stride,
                                          uint2 (&ovf)[TRACE_OVF_STACK], int& sp, InstLane& il) {
  constexpr int N4 = (int)(sizeof(WNode<W>) / 16);
  const bool in_leaf = ref_is_leaf(cur);
  if (INST && in_leaf && il.sp_inst < 0) {
    const uint32_t flags = f2u(q[3].z);
    v3 o = il.wo, d = il.wd;
    if (!(flags & INST_IDENTITY)) {
      o = xform_point(q[0], q[1], q[2], o);
      d = xform_dir(q[0], q[1], q[2], d);
    }
    ray = make_trace_ray(o, d, ray.pre.tmin, ray.pre.tmax);
    il.prim_base = f2u(q[3].y);
    il.sp_inst = sp;
    cur = f2u(q[3].x);
    return false;
  }
  const uint32_t left = ref_leaf_count(cur);
  const bool two = in_leaf && left >= 2u;
  uint32_t next = REF_EMPTY;  // ref to enter after this item; REF_EMPTY: pop
  bool stay = false;
  if (!in_leaf) {
    WNode<W> nd;
    __builtin_memcpy(&nd, q, sizeof(float4) * N4);
    float dist[W];
    child_dists<W, INST>(nd, ray, ANY ? ray.pre.tmax : best_t, dist);
    if (ANY) {
#pragma unroll
      for (int k = 0; k < W; ++k) {
        if (dist[k] == INFINITY) continue;
        if (next == REF_EMPTY) next = nd.child[k];
        else APT_STACK_PUSH(nd.child[k], dist[k]);
      }
    } else {
      uint32_t refs[W];
#pragma unroll
      for (int k = 0; k < W; ++k) refs[k] = nd.child[k];
      sort_children<W>(dist, refs);
      if (dist[0] != INFINITY) {
#pragma unroll
        for (int k = W - 1; k >= 1; --k)
          if (dist[k] != INFINITY) APT_STACK_PUSH(refs[k], dist[k]);
        next = refs[0];
      }
    }
  } else {
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2) {
      if (h2 == 1 && !two) break;
      const float4 a = q[3 * h2], b = q[3 * h2 + 1], c = q[3 * h2 + 2];
      TriHit h;
      if (intersect_tri(ray.pre, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), mk(c.x, c.y, c.z), h)) {
        if (ANY) {
          occluded = true;
        } else {
          const uint32_t prim = INST ? il.prim_base + f2u(a.w) : f2u(a.w);
          if (hit_less(h.t, prim, best_t, best_prim)) {
            best_t = h.t;
            best_prim = prim;
            if (UV_LDS) {
              *uv = make_uint2(f2u(h.u), f2u(h.v));
            } else {
              hu = h.u;
              hv = h.v;
            }
          }
        }
      }
    }
    stay = left > 2u;
    if (stay) cur += 2u - (2u << 28);  // first + 2, count - 2
  }
  if (occluded) return true;
  if (stay) return false;
  if (next == REF_EMPTY) {
    while (true) {
      if (INST && sp == il.sp_inst) {
        // the instance is done: back to the world ray and the top level
        ray = make_trace_ray(il.wo, il.wd, ray.pre.tmin, ray.pre.tmax);
        il.sp_inst = -1;
      }
      if (sp <= 0) break;
      const uint2 e = APT_STACK_POP();
      if (ANY || u2f(e.y) <= best_t) {
        next = e.x;
        break;
      }
    }
  }
  if (next == REF_EMPTY) return true;
  cur = next;
  return false;
}

// Persistent traversal with dynamic ray fetch (Aila and Laine, HPG 2009),
// shared by the ray-query kernels and the integrators' extend and shadow
// stages.  A grid that just fills the device pulls ray indices [0, n) from
// the pool counters at `counter` (POOL_WORDS words, zeroed by an earlier
// kernel on the stream, see pool_take); a wave refetches once `refill` of
// its 64 lanes are idle, so lanes do not wait for the slowest ray of their
// wave.
//
// Every iteration each active lane runs one step (step_fetch, step_item).
//
// Src supplies the rays: with kOneScene, members nodes/tris/root and
// `void fetch(uint32_t i, TraceRay& ray)`; otherwise `void fetch(uint32_t i,
// TraceRay& ray, const WNode<W>*& nodes, const float4*& tris, uint32_t& root)`
// (rays of different scenes in one launch); and `void closest(uint32_t i,
// const HitRec&)` or `void any(uint32_t i, bool hit)`.  Nodes below `ktop`
// are read from the LDS copy `top` (stage_top; ktop = 0 with several scenes).
// With INST (one two-level scene, kOneScene) nodes/tris are the scene's
// arenas, root its TLAS root and insts its instance records.
template <int W, int S, bool ANY, class Src, bool INST = false>
__device__ __forceinline__ void trace_pool(Src& src, uint32_t n, uint32_t* counter, uint32_t refill,
                                           uint2* const lds, const int stride, const float4* top,
                                           uint32_t ktop, const InstRec* __restrict__ insts = nullptr) {
  static_assert(!INST || Src::kOneScene, "a two-level launch traces one scene");
  // The first ceil(n / 64) waves take every ray on their first visits (64
  // each); the rest of the grid would only add to the queue on the range
  // counters and find them drained, which over a few thousand rays costs
  // more than tracing them: those waves leave at once.
  const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if ((uint64_t)(blockIdx.x * (blockDim.x >> 6) + wave) * 64u >= n) return;
  uint2 ovf[TRACE_OVF_STACK];
  int sp = 0;
  PoolCursor pc = pool_start();
  bool active = false;
  // cur: the node to fetch, or the leaf ref of the triangles still to test
  // (first advanced and count lowered as the leaf is consumed)
  uint32_t rid = 0, cur = 0, root = 0;
  const WNode<W>* nodes = nullptr;
  const float4* tris = nullptr;
  if constexpr (Src::kOneScene) {
    nodes = src.nodes;
    tris = src.tris;
    root = src.root;
  }
  TraceRay ray;
  InstLane il{};
  float best_t = 0.0f;
  uint32_t best_prim = INVALID;
  uint2* const uv = lds + S * stride;  // (u, v) of the best hit
  while (true) {
    const uint32_t r = pool_take(counter, n, pc, active, refill);
    if (r != ~0u) {
      rid = r;
      if constexpr (Src::kOneScene) src.fetch(r, ray);
      else src.fetch(r, ray, nodes, tris, root);
      best_t = ray.pre.tmax;
      best_prim = INVALID;
      sp = 0;
      cur = root;
      active = true;
      if (INST) {
        il.wo = ray.pre.o;
        il.wd = ray.d;
        il.sp_inst = -1;
      }
    }
    if (__ballot(active) == 0ull) break;
    if (!active) continue;
    float4 q[STEP_NQ];
    step_fetch<W, INST>(nodes, tris, top, ktop, cur, q, insts, il.sp_inst < 0);
    bool occluded = false;
    float hu, hv;
    const bool done = step_item<W, S, ANY, true, INST>(q, ray, cur, best_t, best_prim, hu, hv, uv, occluded, lds,
                                                       stride, ovf, sp, il);
    if (done) {
      if constexpr (ANY) {
        src.any(rid, occluded);
      } else {
        const uint2 e = *uv;
        src.closest(rid, HitRec{best_t, best_prim == INVALID ? 0.0f : u2f(e.x),
                                best_prim == INVALID ? 0.0f : u2f(e.y), best_prim});
      }
      active = false;
    }
  }
}

template <int W, int S, bool INST>
__device__ __forceinline__ HitRec trace_closest(const WNode<W>* __restrict__ nodes,
                                                const float4* __restrict__ tris, uint32_t root,
                                                const TraceRay& ray0, LdsStack st,
                                                const InstRec* __restrict__ insts) {
  uint2* const lds = st.base;
  const int stride = st.stride;
  uint2 ovf[TRACE_OVF_STACK];
  int sp = 0;
  uint32_t cur = root;
  TraceRay ray = ray0;
  InstLane il{};
  if (INST) {
    il.wo = ray0.pre.o;
    il.wd = ray0.d;
    il.sp_inst = -1;
  }
  float best_t = ray.pre.tmax, hu = 0.0f, hv = 0.0f;
  uint32_t best_prim = INVALID;
  bool occluded = false;
  while (true) {
    float4 q[STEP_NQ];
    step_fetch<W, INST>(nodes, tris, nullptr, 0u, cur, q, insts, il.sp_inst < 0);
    if (step_item<W, S, false, false, INST>(q, ray, cur, best_t, best_prim, hu, hv, nullptr, occluded, lds,
                                            stride, ovf, sp, il))
      break;
  }
  return HitRec{best_t, hu, hv, best_prim};
}

template <int W, int S, bool INST>
__device__ __forceinline__ bool trace_any(const WNode<W>* __restrict__ nodes,
                                          const float4* __restrict__ tris, uint32_t root,
                                          const TraceRay& ray0, LdsStack st, const InstRec* __restrict__ insts) {
  uint2* const lds = st.base;
  const int stride = st.stride;
  uint2 ovf[TRACE_OVF_STACK];
  int sp = 0;
  uint32_t cur = root;
  TraceRay ray = ray0;
  InstLane il{};
  if (INST) {
    il.wo = ray0.pre.o;
    il.wd = ray0.d;
    il.sp_inst = -1;
  }
  float best_t = ray.pre.tmax, hu, hv;
  uint32_t best_prim = INVALID;
  bool occluded = false;
  while (true) {
    float4 q[STEP_NQ];
    step_fetch<W, INST>(nodes, tris, nullptr, 0u, cur, q, insts, il.sp_inst < 0);
    if (step_item<W, S, true, false, INST>(q, ray, cur, best_t, best_prim, hu, hv, nullptr, occluded, lds, stride,
                                           ovf, sp, il))
      break;
  }
  return occluded;
}

}  // namespace apt
