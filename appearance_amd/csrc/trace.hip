// Wide-BVH traversal kernels for gfx950: closest hit and any hit.
//
// These replace the reference's hardware ray query (rayQueryInitialize /
// rayQueryProceed / rayQueryGetCommittedIntersection, trace.wgsl:98-102 and
// helpers/trace.wgsl:20-26).  The traversal stack keeps its bottom S entries
// in LDS (laid out [entry][lane] so lanes at the same depth hit distinct
// banks; S = 8, 12 or 16 per BVH, trading LDS for occupancy) and spills
// deeper entries to a small private array; the host checks that the BVH's
// worst-case stack fits.
//
// Two schedules.  k_trace gives each lane one ray.  k_trace_pool is
// persistent: a grid that just fills the device pulls rays from a shared
// counter, and a wave refetches as soon as `refill` of its 64 lanes have
// finished their rays, so lanes no longer idle behind the slowest ray of
// their wave (persistent threads with dynamic ray fetch, Aila and Laine,
// HPG 2009).  For 2- and 4-wide BVHs each iteration of the pool kernel
// fetches one item per lane, a node or two triangles, so the wave pays one
// memory round trip per iteration whatever its lanes are doing.  Both
// schedules return the same hits: every ray visits children nearest first
// and the closest hit is the lexicographic (t, prim) minimum.
#include <algorithm>
#include <mutex>

#include "apt_internal.h"
#include "tlas.h"
#include "trace_device.h"

namespace apt {

namespace {

constexpr int BLOCK = 256;

template <int W, int S, bool ANY, bool INST = false>
__global__ __launch_bounds__(BLOCK) void k_trace(const WNode<W>* __restrict__ nodes,
                                                 const float4* __restrict__ tris, uint32_t root,
                                                 const apt_ray* __restrict__ rays, uint32_t n,
                                                 apt_hit* __restrict__ hits,
                                                 uint32_t* __restrict__ occluded,
                                                 const InstRec* __restrict__ insts = nullptr) {
  __shared__ uint2 s_stack[S * BLOCK];
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t i = blk * BLOCK + threadIdx.x;
  if (i >= n) return;
  const float4* rp = reinterpret_cast<const float4*>(rays) + 2 * (size_t)i;
  const float4 r0 = rp[0];
  const float4 r1 = rp[1];
  TraceRay ray = make_trace_ray(mk(r0.x, r0.y, r0.z), mk(r1.x, r1.y, r1.z), r0.w, r1.w);
  LdsStack st{s_stack + threadIdx.x, BLOCK};
  if (ANY) {
    occluded[i] = trace_any<W, S, INST>(nodes, tris, root, ray, st, insts) ? 1u : 0u;
  } else {
    HitRec h = trace_closest<W, S, INST>(nodes, tris, root, ray, st, insts);
    float4 out = make_float4(h.prim == INVALID ? INFINITY : h.t, h.u, h.v, u2f(h.prim));
    reinterpret_cast<float4*>(hits)[i] = out;
  }
}

__device__ __forceinline__ TraceRay load_ray(const apt_ray* __restrict__ rays, uint32_t i) {
  const float4* rp = reinterpret_cast<const float4*>(rays) + 2 * (size_t)i;
  const float4 r0 = rp[0];
  const float4 r1 = rp[1];
  return make_trace_ray(mk(r0.x, r0.y, r0.z), mk(r1.x, r1.y, r1.z), r0.w, r1.w);
}

__global__ void k_zero_pool(uint32_t* c) { pool_zero(c, threadIdx.x); }

// Rays and results of the ray-query entry points for trace_pool.
template <int W>
struct RaySrc {
  static constexpr bool kOneScene = true;
  const WNode<W>* nodes;
  const float4* tris;
  uint32_t root;
  const apt_ray* rays;
  apt_hit* hits;
  uint32_t* occluded;
  __device__ __forceinline__ void fetch(uint32_t i, TraceRay& ray) const { ray = load_ray(rays, i); }
  __device__ __forceinline__ void closest(uint32_t i, const HitRec& h) const {
    reinterpret_cast<float4*>(hits)[i] =
        make_float4(h.prim == INVALID ? INFINITY : h.t, h.u, h.v, u2f(h.prim));
  }
  __device__ __forceinline__ void any(uint32_t i, bool hit) const { occluded[i] = hit ? 1u : 0u; }
};

// INST: a two-level scene (nodes/tris its arenas, root its TLAS root,
// n_nodes its TLAS node count for the LDS copy).
template <int W, int S, bool ANY, bool INST = false>
__global__ __launch_bounds__(BLOCK, INST ? 6 : 8) void k_trace_pool(const WNode<W>* __restrict__ nodes,
                                                         const float4* __restrict__ tris, uint32_t root,
                                                         const apt_ray* __restrict__ rays, uint32_t n,
                                                         apt_hit* __restrict__ hits,
                                                         uint32_t* __restrict__ occluded,
                                                         uint32_t* __restrict__ counter, uint32_t refill,
                                                         uint32_t n_nodes,
                                                         const InstRec* __restrict__ insts = nullptr) {
  __shared__ uint2 s_stack[pool_lds_entries(S) * BLOCK];
  __shared__ float4 s_top[TOP_BYTES / 16];
  const uint32_t ktop = stage_top<W>(nodes, n_nodes, s_top);
  RaySrc<W> src{nodes, tris, root, rays, hits, occluded};
  trace_pool<W, S, ANY, RaySrc<W>, INST>(src, n, counter, refill, s_stack + threadIdx.x, BLOCK, s_top, ktop,
                                         insts);
}

// A closest-hit miss or an unoccluded ray for every entry (a two-level scene
// without instances).
__global__ void k_fill_miss(uint32_t n, apt_hit* __restrict__ hits, uint32_t* __restrict__ occluded) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (hits) reinterpret_cast<float4*>(hits)[i] = make_float4(INFINITY, 0.0f, 0.0f, u2f(INVALID));
  if (occluded) occluded[i] = 0u;
}

// Blocks of k_trace_pool that fit on the device at once.
template <int W, int S, bool ANY, bool INST = false>
uint32_t pool_grid() {
  static uint32_t g = 0;
  if (!g) {
    int per_cu = 0, dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_pool<W, S, ANY, INST>, BLOCK, 0) ==
            hipSuccess &&
        per_cu > 0)
      g = (uint32_t)per_cu * (uint32_t)p.multiProcessorCount;
    else
      g = 2048;
  }
  return g;
}

// counter_override: the ray-pool counters to use instead of one of the BVH's
// ring (NULL: the ring).
template <int W>
apt_status launch(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n, apt_hit* d_hits,
                  uint32_t* d_occ, bool any, hipStream_t s, uint32_t* counter_override) {
  if (n == 0) return APT_OK;
  const WNode<W>* nodes = static_cast<const WNode<W>*>(bvh->d_nodes);
  if (bvh->refill > 64) {
    dim3 grid((n + BLOCK - 1) / BLOCK);
    APT_FOR_STACK_DEPTH(bvh->lds_stack, {
      if (any)
        hipLaunchKernelGGL((k_trace<W, S_, true>), grid, dim3(BLOCK), 0, s, nodes, bvh->d_tris,
                           bvh->root, d_rays, n, d_hits, d_occ);
      else
        hipLaunchKernelGGL((k_trace<W, S_, false>), grid, dim3(BLOCK), 0, s, nodes, bvh->d_tris,
                           bvh->root, d_rays, n, d_hits, d_occ);
    })
  } else {
    // one pool of the BVH's ring per call: calls in flight on different
    // streams (up to the ring size) do not share counters
    uint32_t* counter = counter_override;
    if (!counter) {
      const uint32_t slot = __atomic_fetch_add(&bvh->counter_next, 1u, __ATOMIC_RELAXED) % APT_BVH_COUNTERS;
      counter = bvh->d_counters + (size_t)slot * POOL_WORDS;
    }
    hipLaunchKernelGGL(k_zero_pool, dim3(1), dim3(POOL_PARTS), 0, s, counter);
    APT_FOR_STACK_DEPTH(bvh->lds_stack, {
      if (any)
        hipLaunchKernelGGL((k_trace_pool<W, S_, true>), dim3(std::min(pool_grid<W, S_, true>(), (n + 63) / 64)),
                           dim3(BLOCK), 0, s, nodes, bvh->d_tris, bvh->root, d_rays, n, d_hits, d_occ,
                           counter, bvh->refill, bvh->n_nodes);
      else
        hipLaunchKernelGGL((k_trace_pool<W, S_, false>), dim3(std::min(pool_grid<W, S_, false>(), (n + 63) / 64)),
                           dim3(BLOCK), 0, s, nodes, bvh->d_tris, bvh->root, d_rays, n, d_hits, d_occ,
                           counter, bvh->refill, bvh->n_nodes);
    })
  }
  APT_HIP(hipGetLastError());
  return APT_OK;
}

apt_status dispatch(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n, apt_hit* d_hits,
                    uint32_t* d_occ, bool any, hipStream_t s, uint32_t* counter_override = nullptr) {
  APT_REQUIRE(bvh && d_rays && (any ? d_occ != nullptr : d_hits != nullptr),
              APT_ERR_INVALID_ARG, "trace: null argument");
  apt_status st = check_trace_stack(bvh->stats, "trace");
  if (st != APT_OK) return st;
  switch (bvh->width) {
    case 2: return launch<2>(bvh, d_rays, n, d_hits, d_occ, any, s, counter_override);
    case 4: return launch<4>(bvh, d_rays, n, d_hits, d_occ, any, s, counter_override);
    case 8: return launch<8>(bvh, d_rays, n, d_hits, d_occ, any, s, counter_override);
    default: set_error("trace: unsupported width %u", bvh->width); return APT_ERR_INVALID_ARG;
  }
}

template <int W>
apt_status launch_tlas(const apt_tlas* t, const apt_ray* d_rays, uint32_t n, apt_hit* d_hits, uint32_t* d_occ,
                       bool any, hipStream_t s) {
  const WNode<W>* nodes = static_cast<const WNode<W>*>(t->d_nodes);
  const InstRec* insts = t->d_inst;
  if (t->refill > 64) {
    dim3 grid((n + BLOCK - 1) / BLOCK);
    APT_FOR_STACK_DEPTH(t->lds_stack, {
      if (any)
        hipLaunchKernelGGL((k_trace<W, S_, true, true>), grid, dim3(BLOCK), 0, s, nodes, t->d_tris, t->root,
                           d_rays, n, d_hits, d_occ, insts);
      else
        hipLaunchKernelGGL((k_trace<W, S_, false, true>), grid, dim3(BLOCK), 0, s, nodes, t->d_tris, t->root,
                           d_rays, n, d_hits, d_occ, insts);
    })
  } else {
    const uint32_t slot = __atomic_fetch_add(&t->counter_next, 1u, __ATOMIC_RELAXED) % APT_BVH_COUNTERS;
    uint32_t* counter = t->d_counters + (size_t)slot * POOL_WORDS;
    hipLaunchKernelGGL(k_zero_pool, dim3(1), dim3(POOL_PARTS), 0, s, counter);
    APT_FOR_STACK_DEPTH(t->lds_stack, {
      if (any)
        hipLaunchKernelGGL((k_trace_pool<W, S_, true, true>),
                           dim3(std::min(pool_grid<W, S_, true, true>(), (n + 63) / 64)), dim3(BLOCK), 0, s, nodes,
                           t->d_tris, t->root, d_rays, n, d_hits, d_occ, counter, t->refill, t->n_tlas_nodes,
                           insts);
      else
        hipLaunchKernelGGL((k_trace_pool<W, S_, false, true>),
                           dim3(std::min(pool_grid<W, S_, false, true>(), (n + 63) / 64)), dim3(BLOCK), 0, s,
                           nodes, t->d_tris, t->root, d_rays, n, d_hits, d_occ, counter, t->refill,
                           t->n_tlas_nodes, insts);
    })
  }
  APT_HIP(hipGetLastError());
  return APT_OK;
}

}  // namespace

apt_status launch_trace_tlas(const apt_tlas* t, const apt_ray* d_rays, uint32_t n, apt_hit* d_hits,
                             uint32_t* d_occ, bool any, hipStream_t s) {
  APT_REQUIRE(t && d_rays && (any ? d_occ != nullptr : d_hits != nullptr), APT_ERR_INVALID_ARG,
              "trace (two-level): null argument");
  if (n == 0) return APT_OK;
  if (t->n_inst == 0) {
    hipLaunchKernelGGL(k_fill_miss, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, n, any ? nullptr : d_hits,
                       any ? d_occ : nullptr);
    APT_HIP(hipGetLastError());
    return APT_OK;
  }
  APT_REQUIRE(t->max_stack <= trace_stack_capacity(), APT_ERR_UNSUPPORTED,
              "trace (two-level): the scene needs a %u-entry stack", t->max_stack);
  switch (t->width) {
    case 2: return launch_tlas<2>(t, d_rays, n, d_hits, d_occ, any, s);
    case 4: return launch_tlas<4>(t, d_rays, n, d_hits, d_occ, any, s);
    case 8: return launch_tlas<8>(t, d_rays, n, d_hits, d_occ, any, s);
    default: set_error("trace (two-level): unsupported width %u", t->width); return APT_ERR_INVALID_ARG;
  }
}

uint32_t trace_stack_capacity() { return 8u + (uint32_t)TRACE_OVF_STACK; }

apt_status check_trace_stack(const apt_bvh_stats& st, const char* who) {
  APT_REQUIRE(st.max_stack <= trace_stack_capacity(), APT_ERR_UNSUPPORTED,
              "%s: the BVH needs a %u-entry traversal stack; the kernels hold %u (8 in LDS at the least, "
              "%d private)", who, st.max_stack, trace_stack_capacity(), TRACE_OVF_STACK);
  return APT_OK;
}

apt_status launch_trace_closest(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n,
                                apt_hit* d_hits, hipStream_t s) {
  return dispatch(bvh, d_rays, n, d_hits, nullptr, false, s);
}

apt_status launch_trace_any(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n,
                            uint32_t* d_occ, hipStream_t s) {
  return dispatch(bvh, d_rays, n, nullptr, d_occ, true, s);
}

}  // namespace apt

// ---------------------------------------------------------------------------
// apt_trace_closest_host_noalloc: apt_trace_closest_host without allocation.
// The rays, the hits and the ray-pool counters are staged in `scratch`, a
// fixed device array of this library (each GPU the library runs on holds its
// own copy until the process ends): APT_TRACE_NOALLOC_MAX_RAYS rays of 32 B,
// as many 16 B hits, then the counters, 3 MiB + 1 KiB.  Host code never
// passes `scratch` itself (its name is a host-side stand-in, not a device
// address): its address on each device comes from hipGetSymbolAddress on the
// first call with that device current, and is kept per device.
// ---------------------------------------------------------------------------
namespace apt {
namespace {

constexpr size_t NOALLOC_RAY_BYTES = (size_t)APT_TRACE_NOALLOC_MAX_RAYS * sizeof(apt_ray);
constexpr size_t NOALLOC_HIT_BYTES = (size_t)APT_TRACE_NOALLOC_MAX_RAYS * sizeof(apt_hit);
constexpr size_t NOALLOC_CTR_OFFSET = NOALLOC_RAY_BYTES + NOALLOC_HIT_BYTES;
constexpr size_t NOALLOC_BYTES = NOALLOC_CTR_OFFSET + (size_t)POOL_WORDS * sizeof(uint32_t);
static_assert(NOALLOC_CTR_OFFSET % 256 == 0, "counters stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[MAX_DEVICES] = {};

apt_status scratch_for_current_device(unsigned char** out) {
  int dev = 0;
  APT_HIP(hipGetDevice(&dev));
  APT_REQUIRE(dev >= 0 && dev < MAX_DEVICES, APT_ERR_DEVICE, "noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("noalloc: hipGetSymbolAddress of the scratch array failed on device %d: %s", dev,
                hipGetErrorString(e));
      return APT_ERR_DEVICE;
    }
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return APT_OK;
}

}  // namespace
}  // namespace apt

extern "C" apt_status apt_trace_closest_host_noalloc(const apt_bvh* bvh, const apt_ray* rays, uint32_t n,
                                                     apt_hit* hits) {
  using namespace apt;
  clear_error();
  APT_REQUIRE(bvh && rays && hits, APT_ERR_INVALID_ARG, "apt_trace_closest_host_noalloc: null argument");
  APT_REQUIRE(n <= APT_TRACE_NOALLOC_MAX_RAYS, APT_ERR_INVALID_ARG,
              "apt_trace_closest_host_noalloc: %u rays exceed the %u of its fixed device array", n,
              APT_TRACE_NOALLOC_MAX_RAYS);
  if (n == 0) return APT_OK;
  unsigned char* base = nullptr;
  apt_status st = scratch_for_current_device(&base);
  if (st != APT_OK) return st;
  apt_ray* d_rays = reinterpret_cast<apt_ray*>(base);
  apt_hit* d_hits = reinterpret_cast<apt_hit*>(base + NOALLOC_RAY_BYTES);
  uint32_t* d_ctr = reinterpret_cast<uint32_t*>(base + NOALLOC_CTR_OFFSET);
  hipStream_t s = lib_stream();
  APT_REQUIRE(s != nullptr, APT_ERR_DEVICE, "apt_trace_closest_host_noalloc: no library stream");
  // The pool counters are cleared by the launch's own k_zero_pool (a kernel,
  // not a memset); rays and hits are written before they are read.
  APT_HIP(hipMemcpyAsync(d_rays, rays, (size_t)n * sizeof(apt_ray), hipMemcpyHostToDevice, s));
  st = dispatch(bvh, d_rays, n, d_hits, nullptr, false, s, d_ctr);
  if (st != APT_OK) return st;
  APT_HIP(hipMemcpyAsync(hits, d_hits, (size_t)n * sizeof(apt_hit), hipMemcpyDeviceToHost, s));
  APT_HIP(hipStreamSynchronize(s));
  return APT_OK;
}
