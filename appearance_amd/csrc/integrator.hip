// Wavefront path integrator for gfx950.
//
// Replaces the reference's per-bounce megakernel (trace.wgsl:main, which
// launches every pixel each bounce and early-outs dead lanes, :140) and its
// apply_di/apply_gi passes by compacted stages:
//   raygen  -> extend (closest hit) -> shade (emission, NEE, BSDF sampling,
//   Russian roulette) -> shadow (any hit) -> extend ... -> film -> resolve.
// Live paths and pending shadow rays are compacted into queues with a wave
// ballot, a popcount prefix and one atomic per wave.  Each path owns its RNG
// state and its radiance sum, so queue order never changes a result; the film
// adds samples to a pixel in sample order.
#include <algorithm>

#include "pathtrace_device.h"

namespace apt {

namespace {

constexpr int BLOCK = 256;
// Pixel-sample -> primary ray (raygen.wgsl:21-40 with a jittered pixel
// position; the reference samples the pixel centre).
__global__ __launch_bounds__(BLOCK) void k_raygen(FrameView fv, SceneView sc, PathView ps,
                                                  uint32_t n_paths, uint32_t sample_base) {
  const uint32_t p = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ps.counters[CNT_QUEUE0] = n_paths;
    ps.counters[CNT_QUEUE1] = 0;
    ps.counters[CNT_SHADOW] = 0;
  }
  if (blockIdx.x == 0) pool_zero(&ps.counters[CNT_POOL_EXT], threadIdx.x);
  if (p >= n_paths) return;
  const uint32_t npix = fv.width * fv.rows;
  const uint32_t s_local = p / npix, pix_local = p - s_local * npix;
  const uint32_t x = pix_local % fv.width, y = fv.row_start + pix_local / fv.width;
  const uint32_t sample_seed = fv.frame_idx * fv.spp + sample_base + s_local + fv.seed;
  start_path(sc, ps, p, x, y, fv.width, fv.height, sample_seed);
  ps.queue[0][p] = p;
}

// Extension rays of the queued paths (persistent ray pool over the queue).
template <int W>
struct ExtendSrc {
  static constexpr bool kOneScene = true;
  const WNode<W>* nodes;
  const float4* tris;
  uint32_t root;
  PathView ps;
  const uint32_t* queue;
  __device__ __forceinline__ void fetch(uint32_t i, TraceRay& ray) const {
    const uint32_t p = queue[i];
    const float4 o = ps.o_pdf[p], d = ps.d_flags[p];
    ray = make_trace_ray(ld3(o), ld3(d), 0.0f, 1e30f);
  }
  __device__ __forceinline__ void closest(uint32_t i, const HitRec& h) const {
    ps.hit[queue[i]] = make_float4(h.t, h.u, h.v, u2f(h.prim));
  }
};

// INST: a two-level scene (sc.insts); its kernels trade occupancy for the
// instance state's registers.
template <int W, int S, bool INST>
__global__ __launch_bounds__(BLOCK, INST ? 6 : 8) void k_extend(SceneView sc, PathView ps, int qin,
                                                                const WNode<W>* __restrict__ nodes) {
  __shared__ uint2 s_stack[pool_lds_entries(S) * BLOCK];
  __shared__ float4 s_top[TOP_BYTES / 16];
  const uint32_t ktop = stage_top<W>(nodes, sc.n_nodes, s_top);
  const uint32_t count = ps.counters[CNT_QUEUE0 + qin];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ps.counters[CNT_QUEUE0 + (qin ^ 1)] = 0;
    ps.counters[CNT_SHADOW] = 0;
    ps.counters[CNT_STAT_EXT] += count;
  }
  if (blockIdx.x == 0) pool_zero(&ps.counters[CNT_POOL_SHADOW], threadIdx.x);
  ExtendSrc<W> src{nodes, sc.bvh_tris, sc.root, ps, ps.queue[qin]};
  trace_pool<W, S, false, ExtendSrc<W>, INST>(src, count, &ps.counters[CNT_POOL_EXT], min(sc.refill, 64u),
                                              s_stack + threadIdx.x, BLOCK, s_top, ktop, sc.insts);
}

// Four waves per SIMD: the compiler's own choice (136 VGPRs) allows three.
// TEX: the scene has material textures or a sky panorama (the textured code is compiled in, at
// three waves per SIMD: at four it spills).
template <bool TEX>
__global__ __launch_bounds__(BLOCK, TEX ? 3 : 4) void k_shade(FrameView fv, SceneView sc_, PathView ps, int qin,
                                                 uint32_t bounce) {
  const SceneView sc = scene_view<TEX>(sc_);
  const uint32_t count = ps.counters[CNT_QUEUE0 + qin];
  // A grid of a few work-groups per CU strides over the queue: work-group b
  // takes tiles b, b + G, ..., so a short queue (a late bounce) still spreads
  // over every XCD, and the work-groups past it leave at once.
  for (uint32_t tile = blockIdx.x; tile * BLOCK < count; tile += gridDim.x) {
    const uint32_t i = tile * BLOCK + threadIdx.x;
    bool alive = false, want_shadow = false;
    uint32_t p = 0;
    if (i < count) {
      p = ps.queue[qin][i];
      shade_path(sc, ps, p, bounce, fv.max_bounces, fv.rr_start, alive, want_shadow);
    }
    int slot, sslot;
    block_append2<BLOCK>(alive, &ps.counters[CNT_QUEUE0 + (qin ^ 1)], slot, want_shadow,
                         &ps.counters[CNT_SHADOW], sslot);
    if (slot >= 0) ps.queue[qin ^ 1][slot] = p;
    if (sslot >= 0) ps.shadow_queue[sslot] = p;
  }
}

// Shadow rays of the paths that asked for one; unoccluded ones add their
// contribution.
template <int W>
struct ShadowSrc {
  static constexpr bool kOneScene = true;
  const WNode<W>* nodes;
  const float4* tris;
  uint32_t root;
  PathView ps;
  __device__ __forceinline__ void fetch(uint32_t i, TraceRay& ray) const {
    const uint32_t p = ps.shadow_queue[i];
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

template <int W, int S, bool INST>
__global__ __launch_bounds__(BLOCK, INST ? 6 : 8) void k_shadow(SceneView sc, PathView ps,
                                                                const WNode<W>* __restrict__ nodes) {
  __shared__ uint2 s_stack[pool_lds_entries(S) * BLOCK];
  __shared__ float4 s_top[TOP_BYTES / 16];
  const uint32_t ktop = stage_top<W>(nodes, sc.n_nodes, s_top);
  const uint32_t count = ps.counters[CNT_SHADOW];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ps.counters[CNT_STAT_SHADOW] += count;
  }
  if (blockIdx.x == 0) pool_zero(&ps.counters[CNT_POOL_EXT], threadIdx.x);  // next bounce's extend
  ShadowSrc<W> src{nodes, sc.bvh_tris, sc.root, ps};
  trace_pool<W, S, true, ShadowSrc<W>, INST>(src, count, &ps.counters[CNT_POOL_SHADOW], min(sc.refill, 64u),
                                             s_stack + threadIdx.x, BLOCK, s_top, ktop, sc.insts);
}

// Adds this pass's samples of every pixel to its accumulator, one sample at a
// time in sample order, so the sum does not depend on how samples are grouped
// into passes.
__global__ __launch_bounds__(BLOCK) void k_film(FrameView fv, PathView ps, uint32_t s_count,
                                                float4* __restrict__ frame_sum) {
  const uint32_t pix = blockIdx.x * BLOCK + threadIdx.x;
  const uint32_t npix = fv.width * fv.rows;
  if (pix >= npix) return;
  float4 a = frame_sum[pix];
  for (uint32_t s = 0; s < s_count; ++s) {
    const float4 l = ps.rad[(size_t)s * npix + pix];
    a.x += l.x;
    a.y += l.y;
    a.z += l.z;
  }
  frame_sum[pix] = a;
}

__device__ __forceinline__ float aces(float x) {  // resolve.wgsl:29-38
  const float s = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
  return clampf(s, 0.0f, 1.0f);
}

__device__ __forceinline__ uint32_t unorm8(float v) {
  return (uint32_t)rintf(clampf(v, 0.0f, 1.0f) * 255.0f);
}

// resolve.wgsl:40-59: frame mean -> ACES -> progressive mean of tonemapped
// frames -> RGBA8 at the 64x64 block-major position (block.wgsl:3-15).  Also
// keeps the linear sum of all samples for the FP32 export.
__global__ __launch_bounds__(BLOCK) void k_resolve(FrameView fv, const float4* __restrict__ frame_sum,
                                                   float4* __restrict__ accum_tm,
                                                   float4* __restrict__ accum_lin,
                                                   uint32_t accum_frames, uint32_t* __restrict__ out_rgba) {
  const uint32_t pix = blockIdx.x * BLOCK + threadIdx.x;
  const uint32_t npix = fv.width * fv.rows;
  if (pix >= npix) return;
  const float4 s = frame_sum[pix];
  const float inv = 1.0f / (float)fv.spp;
  float4 tm = accum_tm[pix];
  tm.x += aces(s.x * inv);
  tm.y += aces(s.y * inv);
  tm.z += aces(s.z * inv);
  accum_tm[pix] = tm;
  float4 lin = accum_lin[pix];
  lin.x += s.x;
  lin.y += s.y;
  lin.z += s.z;
  lin.w += (float)fv.spp;
  accum_lin[pix] = lin;
  const float k = 1.0f / (float)(accum_frames + 1);
  const uint32_t x = pix % fv.width, y = pix / fv.width;
  const uint32_t B = 64;
  const uint32_t bid = (y / B) * ((fv.width + B - 1) / B) + x / B;
  const uint32_t dst = bid * (B * B) + (y % B) * B + (x % B);
  if (out_rgba) out_rgba[dst] = unorm8(tm.x * k) | (unorm8(tm.y * k) << 8) | (unorm8(tm.z * k) << 16) | (255u << 24);
}

template <int W, bool INST>
void launch_bounce_kernels(const SceneView& sc, const PathView& ps, uint32_t n_paths, int qin,
                           const FrameView& fv, uint32_t bounce, hipStream_t s, hipEvent_t* ev) {
  // the traversal stages pull from their queue and the shade stage strides
  // over it: grids that fill the device
  const dim3 grid(std::min<uint32_t>(pool_blocks(), (n_paths + BLOCK - 1) / BLOCK));
  const dim3 pool(std::min<uint32_t>(pool_blocks(), (n_paths + 63) / 64));
  const WNode<W>* nodes = static_cast<const WNode<W>*>(sc.nodes);
  const auto mark = [&](int k) {
    if (ev) (void)hipEventRecord(ev[k], s);
  };
  APT_FOR_STACK_DEPTH(sc.lds_stack, {
    mark(0);
    hipLaunchKernelGGL((k_extend<W, S_, INST>), pool, dim3(BLOCK), 0, s, sc, ps, qin, nodes);
    mark(1);
    mark(2);
    if (scene_textured(sc)) hipLaunchKernelGGL(k_shade<true>, grid, dim3(BLOCK), 0, s, fv, sc, ps, qin, bounce);
    else hipLaunchKernelGGL(k_shade<false>, grid, dim3(BLOCK), 0, s, fv, sc, ps, qin, bounce);
    mark(3);
    mark(4);
    hipLaunchKernelGGL((k_shadow<W, S_, INST>), pool, dim3(BLOCK), 0, s, sc, ps, nodes);
    mark(5);
  })
}

}  // namespace

apt_status integrate_pass(const FrameView& fv, const SceneView& sc, const PathView& ps,
                          uint32_t n_paths, uint32_t sample_base, uint32_t s_count,
                          float4* frame_sum, hipStream_t s, hipEvent_t* stage_ev) {
  const dim3 grid((n_paths + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_raygen, grid, dim3(BLOCK), 0, s, fv, sc, ps, n_paths, sample_base);
  for (uint32_t b = 0; b < fv.max_bounces; ++b) {
    const int qin = (int)(b & 1u);
    hipEvent_t* ev = stage_ev ? stage_ev + 6 * b : nullptr;
    const bool inst = sc.insts != nullptr;
    switch (sc.width) {
      case 2:
        if (inst) launch_bounce_kernels<2, true>(sc, ps, n_paths, qin, fv, b, s, ev);
        else launch_bounce_kernels<2, false>(sc, ps, n_paths, qin, fv, b, s, ev);
        break;
      case 4:
        if (inst) launch_bounce_kernels<4, true>(sc, ps, n_paths, qin, fv, b, s, ev);
        else launch_bounce_kernels<4, false>(sc, ps, n_paths, qin, fv, b, s, ev);
        break;
      case 8:
        if (inst) launch_bounce_kernels<8, true>(sc, ps, n_paths, qin, fv, b, s, ev);
        else launch_bounce_kernels<8, false>(sc, ps, n_paths, qin, fv, b, s, ev);
        break;
      default: set_error("integrator: unsupported BVH width %u", sc.width); return APT_ERR_INVALID_ARG;
    }
  }
  const uint32_t npix = fv.width * fv.rows;
  hipLaunchKernelGGL(k_film, dim3((npix + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, fv, ps, s_count,
                     frame_sum);
  APT_HIP(hipGetLastError());
  return APT_OK;
}

apt_status resolve_frame(const FrameView& fv, const float4* frame_sum, float4* accum_tm,
                         float4* accum_lin, uint32_t accum_frames, uint32_t* out_rgba,
                         hipStream_t s) {
  const uint32_t npix = fv.width * fv.rows;
  hipLaunchKernelGGL(k_resolve, dim3((npix + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, fv, frame_sum,
                     accum_tm, accum_lin, accum_frames, out_rgba);
  APT_HIP(hipGetLastError());
  return APT_OK;
}

}  // namespace apt
