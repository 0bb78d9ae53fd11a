// Per-path device code of the wavefront integrator, shared by the renderer's
// kernels (integrator.hip) and the serving loop's kernels (server.hip).
#pragma once

#include "integrator.h"
#include "shading.h"
#include "trace_device.h"

namespace apt {

constexpr float TRACE_EPSILON = 1e-4f;  // helpers/trace.wgsl:10
constexpr float SUN_DISTANCE = 1e6f;    // shared/sky_bindings.wgsl:4

constexpr uint32_t FLAG_NEE_SIDE = 1u;  // previous vertex: light sampling could reach wi

// Work-group-level compaction into two queues at once: one atomic per queue
// per work-group instead of one per wave (a single counter that every wave of
// a launch hits serialises those atomics).  Every thread of the work-group
// must call it; returns the slots (or -1) through sa / sb.
template <int BLOCK>
__device__ __forceinline__ void block_append2(bool a, uint32_t* ca, int& sa, bool b, uint32_t* cb, int& sb) {
  constexpr int NW = BLOCK / 64;
  __shared__ uint32_t s_n[2][NW];
  __shared__ uint32_t s_base[2];
  const uint32_t lane = __lane_id(), w = threadIdx.x / 64;
  const unsigned long long ma = __ballot(a), mb = __ballot(b);
  __syncthreads();  // a previous call's reads of s_n are done (loops call it repeatedly)
  if (lane == 0) {
    s_n[0][w] = (uint32_t)__popcll(ma);
    s_n[1][w] = (uint32_t)__popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ta = 0, tb = 0;
    for (int k = 0; k < NW; ++k) {
      ta += s_n[0][k];
      tb += s_n[1][k];
    }
    s_base[0] = ta ? atomicAdd(ca, ta) : 0u;
    s_base[1] = tb ? atomicAdd(cb, tb) : 0u;
  }
  __syncthreads();
  uint32_t oa = s_base[0], ob = s_base[1];
  for (uint32_t k = 0; k < w; ++k) {
    oa += s_n[0][k];
    ob += s_n[1][k];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  sa = a ? (int)(oa + (uint32_t)__popcll(ma & below)) : -1;
  sb = b ? (int)(ob + (uint32_t)__popcll(mb & below)) : -1;
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

// Primary ray of a pixel sample: raygen.wgsl:21-40 with a jittered pixel
// position (the reference samples the pixel centre).
__device__ __forceinline__ void camera_ray(const SceneView& sc, uint32_t x, uint32_t y,
                                           uint32_t width, uint32_t height, float jx, float jy,
                                           v3& o, v3& d) {
  const float ux = ((float)x + jx) / (float)width * 2.0f - 1.0f;
  const float uy = -(((float)y + jy) / (float)height * 2.0f - 1.0f);
  v3 dl = normalize(mk(ux * sc.cam_tan_half * sc.cam_aspect, uy * sc.cam_tan_half, -1.0f));
  d = sc.cam_right * dl.x + sc.cam_up * dl.y + sc.cam_back * dl.z;
  o = sc.cam_pos;
}

// Starts path slot p for pixel `pixel` (x, y) of a width x height frame and
// sample seed `sample_seed` (frame * spp + sample + base seed).
__device__ __forceinline__ void start_path(const SceneView& sc, const PathView& ps, uint32_t p,
                                           uint32_t x, uint32_t y, uint32_t width,
                                           uint32_t height, uint32_t sample_seed) {
  uint32_t rng = path_seed(y * width + x, sample_seed);
  const float jx = rng_next(rng), jy = rng_next(rng);
  v3 o, d;
  camera_ray(sc, x, y, width, height, jx, jy, o, d);
  ps.o_pdf[p] = make_float4(o.x, o.y, o.z, 0.0f);
  ps.d_flags[p] = make_float4(d.x, d.y, d.z, u2f(0u));
  ps.thr[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
  ps.rad[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ps.rng[p] = rng;
}

// One path vertex: the closest hit in ps.hit[p] is shaded (sky or emission,
// next-event estimation, BSDF sampling, roulette).  Sets `alive` when the path
// continues and `want_shadow` when a shadow ray with its contribution was
// written to the slot.
__device__ __forceinline__ void shade_path(const SceneView& sc, const PathView& ps, uint32_t p,
                                           uint32_t bounce, uint32_t max_bounces,
                                           uint32_t rr_start, bool& alive, bool& want_shadow) {
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
    if (bounce + 1 < max_bounces) {
      const float r0 = rng_next(rng), r1 = rng_next(rng), r2 = rng_next(rng);
      BsdfSample bs = lambert ? lambert_sample(albedo, sf.f, wo, r0, r1)
                              : disney_sample(bd, sf, wo, t, back_face, r0, r1, r2);
      if (bs.pdf > 0.0f) {
        const float cosw = fabsf(dot(bs.wi, ns));
        v3 Tn = T * bs.f * (cosw / bs.pdf);
        bool ok = Tn.x == Tn.x && Tn.y == Tn.y && Tn.z == Tn.z && maxc(Tn) > 0.0f;
        if (ok && bounce + 1 >= rr_start) {
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

}  // namespace apt
