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

// ---------------------------------------------------------------------------
// Textured materials (shared/material/material_pool_bindings.wgsl:19-154).
// ---------------------------------------------------------------------------

// Bilinear filtering with repeat addressing at level 0 (the reference's
// samplers, src/scene_resources/material_pool.rs:68-74 and sky.rs:44-50) on a
// width x height grid: texel centres at (i + 1/2) / W; the four texels around
// (u W - 1/2, v H - 1/2), indices wrapped modulo the size, and the weights of
// the second column and row.
struct Bilinear {
  uint32_t i00, i10, i01, i11;  // texel indices, row major
  float ax, ay;
};

__device__ __forceinline__ Bilinear bilinear_repeat(uint32_t width, uint32_t height, float u, float v) {
  const float x = (u == u ? u : 0.0f) * (float)width - 0.5f, y = (v == v ? v : 0.0f) * (float)height - 0.5f;
  const float fx = fminf(fmaxf(floorf(x), -1099511627776.0f), 1099511627776.0f);  // +-2^40
  const float fy = fminf(fmaxf(floorf(y), -1099511627776.0f), 1099511627776.0f);
  Bilinear b;
  b.ax = fminf(fmaxf(x - fx, 0.0f), 1.0f);
  b.ay = fminf(fmaxf(y - fy, 0.0f), 1.0f);
  const int64_t W = width, H = height;
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
  const uint32_t x0 = (uint32_t)(((ix % W) + W) % W), x1 = (uint32_t)((((ix + 1) % W) + W) % W);
  const uint32_t y0 = (uint32_t)(((iy % H) + H) % H), y1 = (uint32_t)((((iy + 1) % H) + H) % H);
  b.i00 = y0 * width + x0;
  b.i10 = y0 * width + x1;
  b.i01 = y1 * width + x0;
  b.i11 = y1 * width + x1;
  return b;
}

__device__ __forceinline__ float bilerp(const Bilinear& b, float c00, float c10, float c01, float c11) {
  return (c00 * (1.0f - b.ax) + c10 * b.ax) * (1.0f - b.ay) + (c01 * (1.0f - b.ax) + c11 * b.ax) * b.ay;
}

// A material texture (RGBA8 texels as c / 255).
__device__ __forceinline__ float4 tex_sample(const SceneView& sc, uint32_t id, float u, float v) {
  const TexDesc d = sc.tex[id];
  const Bilinear b = bilinear_repeat(d.width, d.height, u, v);
  const auto texel = [&](uint32_t i) {
    const uint32_t c = d.texels[i];
    return make_float4((float)(c & 0xFFu) / 255.0f, (float)((c >> 8) & 0xFFu) / 255.0f,
                       (float)((c >> 16) & 0xFFu) / 255.0f, (float)(c >> 24) / 255.0f);
  };
  const float4 c00 = texel(b.i00), c10 = texel(b.i10), c01 = texel(b.i01), c11 = texel(b.i11);
  return make_float4(bilerp(b, c00.x, c10.x, c01.x, c11.x), bilerp(b, c00.y, c10.y, c01.y, c11.y),
                     bilerp(b, c00.z, c10.z, c01.z, c11.z), bilerp(b, c00.w, c10.w, c01.w, c11.w));
}

// The sky along direction d without the sun (Sky::sky, shared/sky_bindings.wgsl:68-70):
// the panorama at unit_vector_to_panorama_coords(d) (assets/shaders/math.wgsl:101-106)
// when the scene has one, else the analytic gradient.
__device__ __forceinline__ v3 sky_radiance(const SceneView& sc, v3 d) {
  if (!sc.sky_tex) return sky_gradient(sc.sky, d);
  const float u = (atan2f(d.z, d.x) + PI) / TWO_PI;
  const float v = acosf(clampf(d.y, -1.0f, 1.0f)) / PI;
  const Bilinear b = bilinear_repeat(sc.sky_w, sc.sky_h, u, v);
  const float4 c00 = sc.sky_tex[b.i00], c10 = sc.sky_tex[b.i10], c01 = sc.sky_tex[b.i01], c11 = sc.sky_tex[b.i11];
  return mk(bilerp(b, c00.x, c10.x, c01.x, c11.x), bilerp(b, c00.y, c10.y, c01.y, c11.y),
            bilerp(b, c00.z, c10.z, c01.z, c11.z));
}

// srgb_to_linear (assets/shaders/color.wgsl:9-15), per channel.
__device__ __forceinline__ float srgb_linear(float c) {
  return c < 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

// A material's factors at a surface point: the descriptor's values scaled by
// its textures where it has them (MaterialDescriptor::color, emission,
// metallic_roughness, clearcoat, clearcoat_roughness, transmission, sheen,
// sheen_tint).
struct MatEval {
  v3 color;
  float luminance;  // the colour's alpha
  float metallic, roughness;
  v3 emission;
  float transmission, sheen;
  v3 sheen_tint;
  float clearcoat, clearcoat_roughness;
};

__device__ __forceinline__ MatEval mat_eval(const SceneView& sc, const apt_material& m, float2 uv) {
  MatEval e;
  e.color = mk(m.color[0], m.color[1], m.color[2]);
  e.luminance = 1.0f;
  e.metallic = m.metallic;
  e.roughness = m.roughness;
  e.emission = mk(m.emission[0], m.emission[1], m.emission[2]);
  e.transmission = m.transmission;
  e.sheen = m.sheen;
  e.sheen_tint = mk(m.sheen_tint[0], m.sheen_tint[1], m.sheen_tint[2]);
  e.clearcoat = m.clearcoat;
  e.clearcoat_roughness = m.clearcoat_roughness;
  if (!sc.tex) return e;
  if (m.color_texture != APT_INVALID_TEXTURE) {
    const float4 t = tex_sample(sc, m.color_texture, uv.x, uv.y);
    e.color = e.color * mk(srgb_linear(t.x), srgb_linear(t.y), srgb_linear(t.z));
    e.luminance = t.w;
  }
  if (m.emission_texture != APT_INVALID_TEXTURE && dot(e.emission, e.emission) > 0.0f) {
    const float4 t = tex_sample(sc, m.emission_texture, uv.x, uv.y);
    e.emission = e.emission * mk(t.x, t.y, t.z);
  }
  if (m.metallic_roughness_texture != APT_INVALID_TEXTURE && (e.metallic > 0.0f || e.roughness > 0.0f)) {
    const float4 t = tex_sample(sc, m.metallic_roughness_texture, uv.x, uv.y);
    e.metallic *= t.z;
    e.roughness *= t.y;
  }
  if (m.clearcoat_texture != APT_INVALID_TEXTURE && e.clearcoat > 0.0f)
    e.clearcoat *= tex_sample(sc, m.clearcoat_texture, uv.x, uv.y).x;
  if (m.clearcoat_roughness_texture != APT_INVALID_TEXTURE && e.clearcoat_roughness > 0.0f)
    e.clearcoat_roughness *= tex_sample(sc, m.clearcoat_roughness_texture, uv.x, uv.y).y;
  if (m.transmission_texture != APT_INVALID_TEXTURE && e.transmission > 0.0f)
    e.transmission *= tex_sample(sc, m.transmission_texture, uv.x, uv.y).x;
  if (m.sheen_texture != APT_INVALID_TEXTURE && e.sheen > 0.0f)
    e.sheen *= tex_sample(sc, m.sheen_texture, uv.x, uv.y).x;
  if (m.sheen_tint_texture != APT_INVALID_TEXTURE && dot(e.sheen_tint, e.sheen_tint) > 0.0f) {
    const float4 t = tex_sample(sc, m.sheen_tint_texture, uv.x, uv.y);
    e.sheen_tint = e.sheen_tint * mk(srgb_linear(t.x), srgb_linear(t.y), srgb_linear(t.z));
  }
  return e;
}

// Emission of a material at a surface point (MaterialDescriptor::emission).
__device__ __forceinline__ v3 mat_emission(const SceneView& sc, const apt_material& m, float2 uv) {
  v3 e = mk(m.emission[0], m.emission[1], m.emission[2]);
  if (sc.tex && m.emission_texture != APT_INVALID_TEXTURE && dot(e, e) > 0.0f) {
    const float4 t = tex_sample(sc, m.emission_texture, uv.x, uv.y);
    e = e * mk(t.x, t.y, t.z);
  }
  return e;
}

// DisneyBsdf::from_material (disney_bsdf.wgsl:226-244) of the evaluated factors.
__device__ __forceinline__ Disney disney_from_eval(const apt_material& m, const MatEval& e) {
  Disney d = disney_from_material(m);
  d.color = e.color;
  d.metallic = e.metallic;
  d.luminance = e.luminance;
  d.specular = fmaxf(m.specular, 1.0f - e.roughness);
  d.roughness = e.roughness;
  d.sheen = e.sheen;
  d.sheen_tint = e.sheen_tint;
  d.clearcoat = e.clearcoat;
  d.clearcoat_gloss = 1.0f - e.clearcoat_roughness;
  d.transmission = e.transmission;
  return d;
}

// Interpolated uv of a hit (0 without uvs); bu, bv weight vertices 1 and 2.
__device__ __forceinline__ float2 hit_uv(const SceneView& sc, uint32_t prim, float bu, float bv) {
  if (!sc.tri_uv) return make_float2(0.0f, 0.0f);
  const float2* t = sc.tri_uv + 3 * (size_t)prim;
  const float b0 = 1.0f - bu - bv;
  return make_float2(t[0].x * b0 + t[1].x * bu + t[2].x * bv, t[0].y * b0 + t[1].y * bu + t[2].y * bv);
}

// Normal mapping (MaterialDescriptor::apply_normal_mapping and
// apply_clearcoat_normal_mapping, material_pool_bindings.wgsl:89-123, in the
// frame of VertexPoolBindings::load_tbn, vertex_pool_bindings.wgsl:48-58):
// n is the interpolated vertex normal before the face flip.  n_map bends it
// by the normal texture; n_cc is the clearcoat lobe's normal -- the clearcoat
// normal texture's, n_map when both textures are the same, else n itself
// (trace.wgsl:176-191).
__device__ __forceinline__ void map_normals(const SceneView& sc, const apt_material& m, uint32_t prim, float bu,
                                            float bv, float2 uv, v3 n, v3& n_map, v3& n_cc) {
  n_map = n;
  n_cc = n;
  if (!sc.tex || !sc.tri_tan) return;
  const bool nt = m.normal_texture != APT_INVALID_TEXTURE && m.normal_scale > 0.0f;
  const bool ct = m.clearcoat_normal_texture != APT_INVALID_TEXTURE;
  if (!nt && !ct) return;
  const float b0 = 1.0f - bu - bv;
  const float4* tt = sc.tri_tan + 3 * (size_t)prim;
  const float4* tb = sc.tri_bit + 3 * (size_t)prim;
  const v3 T = normalize(ld3(tt[0]) * b0 + ld3(tt[1]) * bu + ld3(tt[2]) * bv);
  const v3 B = normalize(ld3(tb[0]) * b0 + ld3(tb[1]) * bu + ld3(tb[2]) * bv);
  if (nt) {
    const float4 t = tex_sample(sc, m.normal_texture, uv.x, uv.y);
    v3 nm = normalize(T * (t.x * 2.0f - 1.0f) + B * (t.y * 2.0f - 1.0f) + n * (t.z * 2.0f - 1.0f));
    if (m.normal_scale < 1.0f) nm = normalize(n * (1.0f - m.normal_scale) + nm * m.normal_scale);
    n_map = nm;
  }
  if (ct) {
    if (m.clearcoat_normal_texture == m.normal_texture) {
      n_cc = n_map;
    } else {
      const float4 t = tex_sample(sc, m.clearcoat_normal_texture, uv.x, uv.y);
      n_cc = normalize(T * (t.x * 2.0f - 1.0f) + B * (t.y * 2.0f - 1.0f) + n * (t.z * 2.0f - 1.0f));
    }
  }
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
    v3 Le = sky_radiance(sc, rd);
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
    // shading normals: normal mapping on the interpolated normal, then the
    // face flip of all of them (trace.wgsl:143-191)
    const float2 uv = hit_uv(sc, prim, bu, bv);
    v3 n_map, n_cc;
    map_normals(sc, mat, prim, bu, bv, uv, ns, n_map, n_cc);
    const bool back_face = dot(wo, ns) < 0.0f;
    if (back_face) {
      n_map = -n_map;
      n_cc = -n_cc;
    }
    ns = n_map;
    const MatEval me = mat_eval(sc, mat, uv);
    // Emission (one-sided: the emitter's front is its winding normal).
    if (has_emission(mat) && dot(ng, rd) < 0.0f) {
      v3 Le = me.emission;
      float w = 1.0f;
      if (pdf_prev > 0.0f && (flags & FLAG_NEE_SIDE) && sc.n_emissive > 0) {
        float cos_l = -dot(ng, rd);
        float p_l = (1.0f - p_sun) / (float)sc.n_emissive / tri_area(g) * (t * t) / cos_l;
        w = power_heuristic(pdf_prev, p_l);
      }
      L += T * Le * w;
    }
    const bool lambert = mat.bsdf_model == APT_BSDF_LAMBERT;
    const v3 albedo = me.color;
    const Disney bd = disney_from_eval(mat, me);
    ShadeFrame sf;
    sf.f = onb(ns);
    sf.cc = sc.tri_tan ? onb(n_cc) : sf.f;

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
        Le = mat_emission(sc, lm, hit_uv(sc, lp, su * (1.0f - u_b), su * u_b));
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
