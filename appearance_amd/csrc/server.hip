// Serving-loop kernels: one wavefront batch over path slots drawn from the
// tiles of many admitted requests (different scenes, cameras, resolutions,
// sample counts and depths in the same launch).
//
// A batch is raygen -> (extend, shade, shadow) x max_bounces -> film, all
// over the same slot pool, captured once into a hipGraph and replayed; the
// per-batch inputs (tile list, path count) live in device memory so the
// captured kernel arguments never change.  Film adds each finished path to its
// request's accumulator with 64-bit fixed-point atomics: integer addition is
// associative, so a request's sum does not depend on the order its paths
// finish in or on what else shared the batch.
#include <algorithm>

#include "pathtrace_device.h"
#include "server.h"

namespace apt {

namespace {

constexpr int BLOCK = 256;
constexpr float FIXED_SCALE = 16777216.0f;  // 2^24

// Tile of the batch containing batch path p (tiles are sorted by offset).
__device__ __forceinline__ uint32_t find_tile(const SrvTile* __restrict__ tiles, uint32_t n,
                                              uint32_t p) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (tiles[mid].offset <= p) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(BLOCK) void k_srv_raygen(const SrvBatch* __restrict__ batch,
                                                      const SrvReq* __restrict__ reqs, PathView ps,
                                                      uint32_t* __restrict__ path_req,
                                                      uint32_t* __restrict__ path_pix) {
  const uint32_t n_paths = batch->n_paths;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ps.counters[CNT_QUEUE0] = n_paths;
    ps.counters[CNT_QUEUE1] = 0;
    ps.counters[CNT_SHADOW] = 0;
  }
  if (blockIdx.x == 0) pool_zero(&ps.counters[CNT_POOL_EXT], threadIdx.x);
  // a device-filling grid strides over the batch, whose size the graph
  // replay does not know
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < n_paths; p += gridDim.x * BLOCK) {
    const SrvTile& t = batch->tiles[find_tile(batch->tiles, batch->n_tiles, p)];
    const SrvReq& r = reqs[t.req_slot];
    const uint32_t idx = t.first + (p - t.offset);  // request path index, sample-major
    const uint32_t npix = r.width * r.height;
    const uint32_t sample = idx / npix, pixel = idx - sample * npix;
    const uint32_t x = pixel % r.width, y = pixel / r.width;
    start_path(r.sc, ps, p, x, y, r.width, r.height, sample + r.seed);
    path_req[p] = t.req_slot;
    path_pix[p] = pixel;
    ps.queue[0][p] = p;
  }
}

constexpr uint32_t SRV_REFILL = 32;  // persistent traversal refill threshold

// Waves per SIMD of the persistent batch kernel (shading and both traversals
// in one kernel), and its grid: every CU's share at that occupancy.
constexpr int SRV_PERSIST_WAVES = 4;
inline uint32_t srv_persist_blocks() {
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  return (uint32_t)std::max(1, cus) * (uint32_t)SRV_PERSIST_WAVES;  // 4 SIMDs x waves / 4 waves per block
}

// Rays of paths from many requests (and scenes) in one launch.
template <int W>
struct SrvExtendSrc {
  static constexpr bool kOneScene = false;
  const SrvReq* reqs;
  PathView ps;
  const uint32_t* path_req;
  const uint32_t* queue;
  __device__ __forceinline__ void fetch(uint32_t i, TraceRay& ray, const WNode<W>*& nodes, const float4*& tris,
                                        uint32_t& root) const {
    const uint32_t p = queue[i];
    const SceneView& sc = reqs[path_req[p]].sc;
    nodes = static_cast<const WNode<W>*>(sc.nodes);
    tris = sc.bvh_tris;
    root = sc.root;
    const float4 o = ps.o_pdf[p], d = ps.d_flags[p];
    ray = make_trace_ray(ld3(o), ld3(d), 0.0f, 1e30f);
  }
  __device__ __forceinline__ void closest(uint32_t i, const HitRec& h) const {
    ps.hit[queue[i]] = make_float4(h.t, h.u, h.v, u2f(h.prim));
  }
};

template <int W, int S>
__global__ __launch_bounds__(BLOCK, 8) void k_srv_extend(const SrvReq* __restrict__ reqs, PathView ps,
                                                         const uint32_t* __restrict__ path_req, int qin) {
  __shared__ uint2 s_stack[pool_lds_entries(S) * BLOCK];
  const uint32_t count = ps.counters[CNT_QUEUE0 + qin];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ps.counters[CNT_QUEUE0 + (qin ^ 1)] = 0;
    ps.counters[CNT_SHADOW] = 0;
    ps.counters[CNT_STAT_EXT] += count;
  }
  if (blockIdx.x == 0) pool_zero(&ps.counters[CNT_POOL_SHADOW], threadIdx.x);
  SrvExtendSrc<W> src{reqs, ps, path_req, ps.queue[qin]};
  trace_pool<W, S, false>(src, count, &ps.counters[CNT_POOL_EXT], SRV_REFILL, s_stack + threadIdx.x, BLOCK,
                          nullptr, 0u);  // rays of several scenes: no shared top levels
}

__global__ __launch_bounds__(BLOCK) void k_srv_shade(const SrvReq* __restrict__ reqs, PathView ps,
                                                     const uint32_t* __restrict__ path_req,
                                                     int qin, uint32_t bounce) {
  const uint32_t count = ps.counters[CNT_QUEUE0 + qin];
  // grid-stride over the queue, as the plain integrator's shade stage
  for (uint32_t tile = blockIdx.x; tile * BLOCK < count; tile += gridDim.x) {
    const uint32_t i = tile * BLOCK + threadIdx.x;
    bool alive = false, want_shadow = false;
    uint32_t p = 0;
    if (i < count) {
      p = ps.queue[qin][i];
      const SrvReq& r = reqs[path_req[p]];
      // resident scenes carry no textures
      shade_path(scene_view<false>(r.sc), ps, p, bounce, r.max_bounces, r.rr_start, alive, want_shadow);
    }
    int slot, sslot;
    block_append2<BLOCK>(alive, &ps.counters[CNT_QUEUE0 + (qin ^ 1)], slot, want_shadow,
                         &ps.counters[CNT_SHADOW], sslot);
    if (slot >= 0) ps.queue[qin ^ 1][slot] = p;
    if (sslot >= 0) ps.shadow_queue[sslot] = p;
  }
}

template <int W>
struct SrvShadowSrc {
  static constexpr bool kOneScene = false;
  const SrvReq* reqs;
  PathView ps;
  const uint32_t* path_req;
  __device__ __forceinline__ void fetch(uint32_t i, TraceRay& ray, const WNode<W>*& nodes, const float4*& tris,
                                        uint32_t& root) const {
    const uint32_t p = ps.shadow_queue[i];
    const SceneView& sc = reqs[path_req[p]].sc;
    nodes = static_cast<const WNode<W>*>(sc.nodes);
    tris = sc.bvh_tris;
    root = sc.root;
    const float4 o = ps.sh_o[p], d = ps.sh_d[p];
    ray = make_trace_ray(ld3(o), ld3(d), 0.0f, o.w);
  }
  __device__ __forceinline__ void any(uint32_t i, bool hit) const {
    if (hit) return;
    const uint32_t p = ps.shadow_queue[i];
    const float4 c = ps.sh_c[p];
    const float4 l = ps.rad[p];
    ps.rad[p] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
  }
};

template <int W, int S>
__global__ __launch_bounds__(BLOCK, 8) void k_srv_shadow(const SrvReq* __restrict__ reqs, PathView ps,
                                                         const uint32_t* __restrict__ path_req) {
  __shared__ uint2 s_stack[pool_lds_entries(S) * BLOCK];
  const uint32_t count = ps.counters[CNT_SHADOW];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ps.counters[CNT_STAT_SHADOW] += count;
  }
  if (blockIdx.x == 0) pool_zero(&ps.counters[CNT_POOL_EXT], threadIdx.x);  // next bounce's extend
  SrvShadowSrc<W> src{reqs, ps, path_req};
  trace_pool<W, S, true>(src, count, &ps.counters[CNT_POOL_SHADOW], SRV_REFILL, s_stack + threadIdx.x, BLOCK,
                         nullptr, 0u);  // rays of several scenes: no shared top levels
}

__device__ __forceinline__ unsigned long long to_fixed(float v) {
  // radiance is >= 0; non-finite or absurd values are dropped rather than
  // allowed to wrap the accumulator
  if (!(v > 0.0f)) return 0ull;
  const float s = v * FIXED_SCALE;
  return s < 1.0e18f ? (unsigned long long)rintf(s) : 0ull;
}

__global__ __launch_bounds__(BLOCK) void k_srv_film(const SrvBatch* __restrict__ batch,
                                                    const SrvReq* __restrict__ reqs, PathView ps,
                                                    const uint32_t* __restrict__ path_req,
                                                    const uint32_t* __restrict__ path_pix) {
  const uint32_t n_paths = batch->n_paths;
  for (uint32_t p = blockIdx.x * BLOCK + threadIdx.x; p < n_paths; p += gridDim.x * BLOCK) {
    const float4 l = ps.rad[p];
    unsigned long long* a = reqs[path_req[p]].accum + 3 * (size_t)path_pix[p];
    atomicAdd(a + 0, to_fixed(l.x));
    atomicAdd(a + 1, to_fixed(l.y));
    atomicAdd(a + 2, to_fixed(l.z));
  }
}

__global__ __launch_bounds__(BLOCK) void k_srv_resolve(const unsigned long long* __restrict__ accum,
                                                       uint32_t npix, uint32_t spp,
                                                       float* __restrict__ out) {
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= npix * 3) return;
  const double scale = 1.0 / ((double)FIXED_SCALE * (double)spp);
  out[i] = (float)((double)accum[i] * scale);
}

__global__ void k_srv_zero(unsigned long long* __restrict__ a, size_t n) {
  const size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) a[i] = 0ull;
}

// ---- one persistent kernel per batch ------------------------------------------
// Every lane owns a path slot and pulls the batch's paths from a counter.  In
// each round the lanes of a wave that hold no path (at the start, or whose
// path just ended) take the next batch paths with one atomic add per wave;
// then every lane with a path runs its next segment -- the extension ray,
// the shading of its hit, the shadow ray -- so the paths of a wave advance
// together whatever request, scene, resolution or depth they belong to.  A
// finished path is added to its request's accumulator (64-bit fixed point,
// as k_srv_film) before its lane takes the next.  A wave leaves once the
// counter is past the batch and none of its lanes holds a path.  The
// per-path arithmetic is the wavefront stages' (start_path, shade_path, the
// bit-exact traversals, the same additions), so a request's image is the same
// bits in either mode.
__global__ void k_srv_persist_reset(PathView ps) {
  if (threadIdx.x == 0) ps.counters[CNT_QUEUE0] = 0;
}

template <int W, int S>
__global__ __launch_bounds__(BLOCK, SRV_PERSIST_WAVES) void k_srv_persistent(const SrvBatch* __restrict__ batch,
                                                                      const SrvReq* __restrict__ reqs, PathView ps) {
  __shared__ uint2 s_stack[S * BLOCK];
  const LdsStack st{s_stack + threadIdx.x, BLOCK};
  const uint32_t slot = blockIdx.x * BLOCK + threadIdx.x;
  const uint32_t n_paths = batch->n_paths;
  const uint32_t lane = threadIdx.x & 63u;
  uint32_t* next = &ps.counters[CNT_QUEUE0];
  bool active = false;
  uint32_t bounce = 0, req = 0, pixel = 0;
  while (true) {
    const unsigned long long idle = __ballot(!active);
    if (idle) {
      const int leader = __ffsll((long long)idle) - 1;
      uint32_t base = 0;
      if ((int)lane == leader) {
        // once the batch is handed out no further adds: the counter cannot run
        // past n_paths by more than the waves' idle lanes
        base = __hip_atomic_load(next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (base < n_paths) base = atomicAdd(next, (uint32_t)__popcll(idle));
      }
      base = __shfl(base, leader);
      if (!active) {
        const uint32_t path = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
        if (base < n_paths && path < n_paths) {
          const SrvTile& t = batch->tiles[find_tile(batch->tiles, batch->n_tiles, path)];
          const SrvReq& r = reqs[t.req_slot];
          const uint32_t idx = t.first + (path - t.offset);  // request path index, sample-major
          const uint32_t npix = r.width * r.height;
          const uint32_t sample = idx / npix;
          pixel = idx - sample * npix;
          start_path(r.sc, ps, slot, pixel % r.width, pixel / r.width, r.width, r.height, sample + r.seed);
          req = t.req_slot;
          bounce = 0;
          active = true;
        }
      }
    }
    if (!__any(active)) break;
    if (active) {
      const SrvReq& r = reqs[req];
      const WNode<W>* nodes = static_cast<const WNode<W>*>(r.sc.nodes);
      const float4 o = ps.o_pdf[slot], d = ps.d_flags[slot];
      const HitRec h = trace_closest<W, S, false>(nodes, r.sc.bvh_tris, r.sc.root,
                                                   make_trace_ray(ld3(o), ld3(d), 0.0f, 1e30f), st, nullptr);
      ps.hit[slot] = make_float4(h.t, h.u, h.v, u2f(h.prim));
      bool alive = false, want_shadow = false;
      // resident scenes carry no textures
      shade_path(scene_view<false>(r.sc), ps, slot, bounce, r.max_bounces, r.rr_start, alive, want_shadow);
      if (want_shadow) {
        const float4 so = ps.sh_o[slot], sd = ps.sh_d[slot];
        if (!trace_any<W, S, false>(nodes, r.sc.bvh_tris, r.sc.root, make_trace_ray(ld3(so), ld3(sd), 0.0f, so.w),
                                    st, nullptr)) {
          const float4 c = ps.sh_c[slot], l = ps.rad[slot];
          ps.rad[slot] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
        }
      }
      ++bounce;
      if (!alive) {
        const float4 l = ps.rad[slot];
        unsigned long long* a = r.accum + 3 * (size_t)pixel;
        atomicAdd(a + 0, to_fixed(l.x));
        atomicAdd(a + 1, to_fixed(l.y));
        atomicAdd(a + 2, to_fixed(l.z));
        active = false;
      }
    }
  }
}

template <int W, int S>
void enqueue_batch_t(const SrvDevice& d, hipStream_t s) {
  if (d.persistent) {
    // a grid that fills the device at the kernel's occupancy, within the path slots
    const uint32_t blocks = std::min<uint32_t>(srv_persist_blocks(), d.path_slots / BLOCK);
    hipLaunchKernelGGL(k_srv_persist_reset, dim3(1), dim3(64), 0, s, d.ps);
    hipLaunchKernelGGL((k_srv_persistent<W, S>), dim3(std::max(1u, blocks)), dim3(BLOCK), 0, s, d.batch, d.reqs,
                       d.ps);
    return;
  }
  // every stage strides over (or pulls from) the batch with a grid that fills the device
  const dim3 grid(std::min<uint32_t>(pool_blocks(), (d.path_slots + BLOCK - 1) / BLOCK)), block(BLOCK);
  const dim3 pool(std::min<uint32_t>(pool_blocks(), (d.path_slots + 63) / 64));
  hipLaunchKernelGGL(k_srv_raygen, grid, block, 0, s, d.batch, d.reqs, d.ps, d.path_req, d.path_pix);
  for (uint32_t b = 0; b < d.max_bounces; ++b) {
    const int qin = (int)(b & 1u);
    hipLaunchKernelGGL((k_srv_extend<W, S>), pool, block, 0, s, d.reqs, d.ps, d.path_req, qin);
    hipLaunchKernelGGL(k_srv_shade, grid, block, 0, s, d.reqs, d.ps, d.path_req, qin, b);
    hipLaunchKernelGGL((k_srv_shadow<W, S>), pool, block, 0, s, d.reqs, d.ps, d.path_req);
  }
  hipLaunchKernelGGL(k_srv_film, grid, block, 0, s, d.batch, d.reqs, d.ps, d.path_req, d.path_pix);
}

}  // namespace

apt_status srv_enqueue_batch(const SrvDevice& d, hipStream_t s) {
  switch (d.width) {
    case 2: APT_FOR_STACK_DEPTH(d.lds_stack, enqueue_batch_t<2, S_>(d, s)); break;
    case 4: APT_FOR_STACK_DEPTH(d.lds_stack, enqueue_batch_t<4, S_>(d, s)); break;
    case 8: APT_FOR_STACK_DEPTH(d.lds_stack, enqueue_batch_t<8, S_>(d, s)); break;
    default: set_error("server: unsupported BVH width %u", d.width); return APT_ERR_INVALID_ARG;
  }
  APT_HIP(hipGetLastError());
  return APT_OK;
}

apt_status srv_resolve(const unsigned long long* accum, uint32_t npix, uint32_t spp, float* out,
                       hipStream_t s) {
  const uint32_t n = npix * 3;
  hipLaunchKernelGGL(k_srv_resolve, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, accum, npix,
                     spp, out);
  APT_HIP(hipGetLastError());
  return APT_OK;
}

apt_status srv_zero(unsigned long long* a, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_srv_zero, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, s, a, n);
  APT_HIP(hipGetLastError());
  return APT_OK;
}

}  // namespace apt
