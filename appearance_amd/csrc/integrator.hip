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
#include "integrator.h"
#include "shading.h"
#include "trace_device.h"

namespace apt {

namespace {

constexpr int BLOCK = 256;
constexpr float TRACE_EPSILON = 1e-4f;  // helpers/trace.wgsl:10
constexpr float SUN_DISTANCE = 1e6f;    // shared/sky_bindings.wgsl:4

constexpr uint32_t FLAG_NEE_SIDE = 1u;  // previous vertex: light sampling could reach wi

// Wave-level compaction: returns this lane's slot in `queue`, or -1.
__device__ __forceinline__ int wave_append(bool want, uint32_t* counter) {
  const uint64_t mask = __ballot(want);
  if (mask == 0) return -1;
  const uint32_t lane = __lane_id();
  const uint32_t below = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  const int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t base = 0;
  if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  return want ? (int)(base + below) : -1;
}

__device__ __forceinline__ v3 ld3(const float4& a) { return mk(a.x, a.y, a.z); }

struct TriGeom {
  v3 p0, p1, p2;
};

__device__ __forceinline__ TriGeom load_tri(const SceneView& sc, uint32_t prim) {
  const float4* t = sc.tri_pos + 3 * (size_t)prim;
  return {ld3(t[0]), ld3(t[1]), ld3(t[2])};
}

__device__ __forceinline__ float tri_area(const TriGeom& g) {
  return 0.5f * length(cross(g.p1 - g.p0, g.p2 - g.p0));
}

__device__ __forceinline__ bool has_emission(const apt_material& m) {
  return m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f;
}

__device__ __forceinline__ float sun_pick(const SceneView& sc) {
  if (!(sc.sky.sun_enabled > 0.0f)) return 0.0f;
  return sc.n_emissive > 0 ? 0.5f : 1.0f;  // Nee::sample_uniform (nee.wgsl:102-115)
}

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
  if (p >= n_paths) return;
  const uint32_t npix = fv.width * fv.rows;
  const uint32_t s_local = p / npix, pix_local = p - s_local * npix;
  const uint32_t x = pix_local % fv.width, y = fv.row_start + pix_local / fv.width;
  const uint32_t pixel = y * fv.width + x;
  const uint32_t sample_seed = fv.frame_idx * fv.spp + sample_base + s_local + fv.seed;
  uint32_t rng = path_seed(pixel, sample_seed);
  const float jx = rng_next(rng), jy = rng_next(rng);
  const float ux = ((float)x + jx) / (float)fv.width * 2.0f - 1.0f;
  const float uy = -(((float)y + jy) / (float)fv.height * 2.0f - 1.0f);
  v3 dl = normalize(mk(ux * sc.cam_tan_half * sc.cam_aspect, uy * sc.cam_tan_half, -1.0f));
  v3 d = sc.cam_right * dl.x + sc.cam_up * dl.y + sc.cam_back * dl.z;
  ps.o_pdf[p] = make_float4(sc.cam_pos.x, sc.cam_pos.y, sc.cam_pos.z, 0.0f);
  ps.d_flags[p] = make_float4(d.x, d.y, d.z, u2f(0u));
  ps.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  ps.rad[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ps.rng[p] = rng;
  ps.queue[0][p] = p;
}

template <int W, int S>
__global__ __launch_bounds__(BLOCK) void k_extend(SceneView sc, PathView ps, int qin,
                                                  const WNode<W>* __restrict__ nodes) {
  __shared__ uint2 s_stack[S * BLOCK];
  const uint32_t count = ps.counters[CNT_QUEUE0 + qin];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ps.counters[CNT_QUEUE0 + (qin ^ 1)] = 0;
    ps.counters[CNT_SHADOW] = 0;
    ps.counters[CNT_STAT_EXT] += count;
  }
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (i >= count) return;
  const uint32_t p = ps.queue[qin][i];
  const float4 o = ps.o_pdf[p], d = ps.d_flags[p];
  TraceRay ray = make_trace_ray(ld3(o), ld3(d), 0.0f, 1e30f);
  HitRec h = trace_closest<W, S>(nodes, sc.bvh_tris, sc.root, ray, LdsStack{s_stack + threadIdx.x, BLOCK});
  ps.hit[p] = make_float4(h.t, h.u, h.v, u2f(h.prim));
}

__global__ __launch_bounds__(BLOCK) void k_shade(FrameView fv, SceneView sc, PathView ps, int qin,
                                                 uint32_t bounce) {
  const uint32_t count = ps.counters[CNT_QUEUE0 + qin];
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  const bool active = i < count;
  bool alive = false, want_shadow = false;
  uint32_t p = 0;
  if (active) {
    p = ps.queue[qin][i];
    float4 o4 = ps.o_pdf[p], d4 = ps.d_flags[p], t4 = ps.thr[p], l4 = ps.rad[p], h4 = ps.hit[p];
    const v3 ro = ld3(o4), rd = ld3(d4);
    const float pdf_prev = o4.w;
    const uint32_t flags = f2u(d4.w);
    v3 T = ld3(t4), L = ld3(l4);
    uint32_t rng = ps.rng[p];
    const uint32_t prim = f2u(h4.w);
    const float p_sun = sun_pick(sc);
    if (prim == INVALID) {
      // Miss: gradient sky (not light-sampled) plus the sun disk, MIS-weighted
      // against sun sampling at the previous vertex.
      v3 Le = sky_gradient(sc.sky, rd);
      float si = sun_intensity(sc.sky, rd);
      if (si > 0.0f) {
        float w = 1.0f;
        if (pdf_prev > 0.0f && (flags & FLAG_NEE_SIDE))
          w = power_heuristic(pdf_prev, p_sun / sun_solid_angle(sc.sky));
        Le += mk(sc.sky.sun_color[0], sc.sky.sun_color[1], sc.sky.sun_color[2]) * (si * w);
      }
      L += T * Le;
    } else {
      const TriGeom g = load_tri(sc, prim);
      const float t = h4.x, bu = h4.y, bv = h4.z;
      const apt_material& mat = sc.mats[sc.tri_mat[prim]];
      const v3 hp = ro + rd * t;
      const v3 ng = normalize(cross(g.p1 - g.p0, g.p2 - g.p0));
      v3 ns = ng;
      if (sc.tri_nrm) {
        const float4* tn = sc.tri_nrm + 3 * (size_t)prim;
        ns = normalize(ld3(tn[0]) * (1.0f - bu - bv) + ld3(tn[1]) * bu + ld3(tn[2]) * bv);
      }
      const v3 wo = -rd;
      const bool back_face = dot(wo, ns) < 0.0f;
      if (back_face) ns = -ns;
      // Emission (one-sided: the emitter's front is its winding normal).
      if (has_emission(mat) && dot(ng, rd) < 0.0f) {
        v3 Le = mk(mat.emission[0], mat.emission[1], mat.emission[2]);
        float w = 1.0f;
        if (pdf_prev > 0.0f && (flags & FLAG_NEE_SIDE) && sc.n_emissive > 0) {
          float cos_l = -dot(ng, rd);
          float p_l = (1.0f - p_sun) / (float)sc.n_emissive / tri_area(g) * (t * t) / cos_l;
          w = power_heuristic(pdf_prev, p_l);
        }
        L += T * Le * w;
      }
      const bool lambert = mat.bsdf_model == APT_BSDF_LAMBERT;
      const v3 albedo = mk(mat.color[0], mat.color[1], mat.color[2]);
      const Disney bd = disney_from_material(mat);
      ShadeFrame sf;
      sf.f = onb(ns);
      sf.cc = sf.f;

      // Next-event estimation: one light sample (sun with probability p_sun,
      // else a uniformly chosen emissive triangle, uniform in area), pdf in
      // solid angle, power-heuristic MIS against BSDF sampling.
      const float u_sel = rng_next(rng), u_a = rng_next(rng), u_b = rng_next(rng);
      if (p_sun > 0.0f || sc.n_emissive > 0) {
        v3 wi, Le;
        float p_l = 0.0f, dist = 0.0f;
        // rng_next can return exactly 1.0, so the sun branch is also taken
        // whenever there is no emissive triangle to index.
        if (u_sel < p_sun || sc.n_emissive == 0) {
          wi = normalize(perturb_direction(u_a, u_b, sun_to(sc.sky), sun_half_angle(sc.sky)));
          float si = sun_intensity(sc.sky, wi);
          Le = mk(sc.sky.sun_color[0], sc.sky.sun_color[1], sc.sky.sun_color[2]) * si;
          p_l = p_sun / sun_solid_angle(sc.sky);
          dist = SUN_DISTANCE;
        } else {
          const float us = (u_sel - p_sun) / (1.0f - p_sun);
          uint32_t k = (uint32_t)fmaxf(0.0f, us * (float)sc.n_emissive);
          k = k < sc.n_emissive ? k : sc.n_emissive - 1;
          const uint32_t lp = sc.emissive[k];
          const TriGeom lg = load_tri(sc, lp);
          const float su = sqrtf(u_a);
          const v3 x = lg.p0 * (1.0f - su) + lg.p1 * (su * (1.0f - u_b)) + lg.p2 * (su * u_b);
          const v3 nl = normalize(cross(lg.p1 - lg.p0, lg.p2 - lg.p0));
          const v3 dv = x - hp;
          dist = length(dv);
          wi = dv / dist;
          const float cos_l = -dot(nl, wi);
          const apt_material& lm = sc.mats[sc.tri_mat[lp]];
          Le = mk(lm.emission[0], lm.emission[1], lm.emission[2]);
          p_l = cos_l > 0.0f ? (1.0f - p_sun) / (float)sc.n_emissive / tri_area(lg) * (dist * dist) / cos_l : 0.0f;
        }
        const float cos_i = dot(wi, ns);
        if (p_l > 0.0f && cos_i > 0.0f && maxc(Le) > 0.0f) {
          float p_b = 0.0f;
          v3 f = lambert ? lambert_eval(albedo, sf.f, wo, wi, p_b)
                         : disney_eval(bd, sf, wo, wi, back_face, p_b);
          float w = power_heuristic(p_l, p_b);
          v3 c = T * f * Le * (cos_i * w / p_l);
          if (maxc(c) > 0.0f && c.x == c.x && c.y == c.y && c.z == c.z) {
            v3 so = hp + ns * TRACE_EPSILON;
            ps.sh_o[p] = make_float4(so.x, so.y, so.z, dist - 2.0f * TRACE_EPSILON);
            ps.sh_d[p] = make_float4(wi.x, wi.y, wi.z, 0.0f);
            ps.sh_c[p] = make_float4(c.x, c.y, c.z, 0.0f);
            want_shadow = true;
          }
        }
      }

      // Continuation (BSDF sampling) while segments remain.
      if (bounce + 1 < fv.max_bounces) {
        const float r0 = rng_next(rng), r1 = rng_next(rng), r2 = rng_next(rng);
        BsdfSample bs = lambert ? lambert_sample(albedo, sf.f, wo, r0, r1)
                                : disney_sample(bd, sf, wo, t, back_face, r0, r1, r2);
        if (bs.pdf > 0.0f) {
          const float cosw = fabsf(dot(bs.wi, ns));
          v3 Tn = T * bs.f * (cosw / bs.pdf);
          bool ok = Tn.x == Tn.x && Tn.y == Tn.y && Tn.z == Tn.z && maxc(Tn) > 0.0f;
          if (ok && bounce + 1 >= fv.rr_start) {
            const float q = fminf(1.0f, maxc(Tn));
            const float u = rng_next(rng);
            if (u > q) ok = false;
            else Tn = Tn * (1.0f / q);
          }
          if (ok) {
            const float side = dot(bs.wi, ns) > 0.0f ? 1.0f : -1.0f;
            const v3 no = hp + ns * (side * TRACE_EPSILON);
            ps.o_pdf[p] = make_float4(no.x, no.y, no.z, bs.pdf);
            ps.d_flags[p] = make_float4(bs.wi.x, bs.wi.y, bs.wi.z, u2f(side > 0.0f ? FLAG_NEE_SIDE : 0u));
            T = Tn;
            alive = true;
          }
        }
      }
    }
    ps.thr[p] = make_float4(T.x, T.y, T.z, 0.0f);
    ps.rad[p] = make_float4(L.x, L.y, L.z, 0.0f);
    ps.rng[p] = rng;
  }
  const int slot = wave_append(alive, &ps.counters[CNT_QUEUE0 + (qin ^ 1)]);
  if (slot >= 0) ps.queue[qin ^ 1][slot] = p;
  const int sslot = wave_append(want_shadow, &ps.counters[CNT_SHADOW]);
  if (sslot >= 0) ps.shadow_queue[sslot] = p;
}

template <int W, int S>
__global__ __launch_bounds__(BLOCK) void k_shadow(SceneView sc, PathView ps,
                                                  const WNode<W>* __restrict__ nodes) {
  __shared__ uint2 s_stack[S * BLOCK];
  const uint32_t count = ps.counters[CNT_SHADOW];
  if (blockIdx.x == 0 && threadIdx.x == 0) ps.counters[CNT_STAT_SHADOW] += count;
  const uint32_t i = xcd_block(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
  if (i >= count) return;
  const uint32_t p = ps.shadow_queue[i];
  const float4 o = ps.sh_o[p], d = ps.sh_d[p];
  TraceRay ray = make_trace_ray(ld3(o), ld3(d), 0.0f, o.w);
  if (!trace_any<W, S>(nodes, sc.bvh_tris, sc.root, ray, LdsStack{s_stack + threadIdx.x, BLOCK})) {
    const float4 c = ps.sh_c[p];
    float4 l = ps.rad[p];
    ps.rad[p] = make_float4(l.x + c.x, l.y + c.y, l.z + c.z, 0.0f);
  }
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
  out_rgba[dst] = unorm8(tm.x * k) | (unorm8(tm.y * k) << 8) | (unorm8(tm.z * k) << 16) | (255u << 24);
}

template <int W>
void launch_bounce_kernels(const SceneView& sc, const PathView& ps, uint32_t n_paths, int qin,
                           const FrameView& fv, uint32_t bounce, hipStream_t s) {
  const dim3 grid((n_paths + BLOCK - 1) / BLOCK);
  const WNode<W>* nodes = static_cast<const WNode<W>*>(sc.nodes);
  APT_FOR_STACK_DEPTH(sc.lds_stack, {
    hipLaunchKernelGGL((k_extend<W, S_>), grid, dim3(BLOCK), 0, s, sc, ps, qin, nodes);
    hipLaunchKernelGGL(k_shade, grid, dim3(BLOCK), 0, s, fv, sc, ps, qin, bounce);
    hipLaunchKernelGGL((k_shadow<W, S_>), grid, dim3(BLOCK), 0, s, sc, ps, nodes);
  })
}

}  // namespace

apt_status integrate_pass(const FrameView& fv, const SceneView& sc, const PathView& ps,
                          uint32_t n_paths, uint32_t sample_base, uint32_t s_count,
                          float4* frame_sum, hipStream_t s) {
  const dim3 grid((n_paths + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_raygen, grid, dim3(BLOCK), 0, s, fv, sc, ps, n_paths, sample_base);
  for (uint32_t b = 0; b < fv.max_bounces; ++b) {
    const int qin = (int)(b & 1u);
    switch (sc.width) {
      case 2: launch_bounce_kernels<2>(sc, ps, n_paths, qin, fv, b, s); break;
      case 4: launch_bounce_kernels<4>(sc, ps, n_paths, qin, fv, b, s); break;
      case 8: launch_bounce_kernels<8>(sc, ps, n_paths, qin, fv, b, s); break;
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
