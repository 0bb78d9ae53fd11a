// ReSTIR DI / GI integrator and the denoise/post chain for gfx950.
//
// Restates the reference's frame (crates/appearance-path-tracer-gpu/src/lib.rs:
// 354-576) as per-pixel kernels over 16x16 tiles, one sample at a time:
//   raygen                 pixel-centre camera rays (raygen.wgsl:21-40)
//   per bounce:
//     extend               closest hit, bounce-0 emission, sky on miss,
//                          G-buffer at bounce 0 (trace.wgsl:62-263)
//     di_ris               RIS over light candidates plus one BSDF candidate,
//                          visibility of the winner (helpers/nee.wgsl:122-237)
//     gi_ris               roulette, one BSDF-sampled candidate whose target is
//                          an approximate one-bounce estimate
//                          (helpers/inline_path_tracer.wgsl:51-296)
//     bounce 0 only:       DI temporal + spatial reuse (restir_di_temporal.wgsl,
//                          restir_di_spatial.wgsl), GI temporal reuse
//                          (restir_gi_temporal.wgsl)
//     apply_di             shade the DI reservoir (apply_di.wgsl:48-111)
//     apply_gi             continue the path toward the GI sample
//                          (apply_gi.wgsl:44-99)
//   post                   demodulate, firefly filter, SVGF temporal + a-trous,
//                          TAA, remodulate (demodulate_radiance.wgsl,
//                          firefly_filter.wgsl, svgf_temporal.wgsl,
//                          svgf_atrous.wgsl, taa.wgsl)
// Each pixel carries one RNG chain through all of its kernels in the
// reference's draw order.  Deviations from the reference (light pdfs in solid
// angle with plain emission, balance-heuristic MIS on each candidate's own
// pdfs, uniform barycentrics, one-sided emitters, the GI target floor, the
// roulette probability, out-of-place filters, clamped neighbour indices) are
// noted where they occur.
#include <hip/hip_fp16.h>

#include "pathtrace_device.h"
#include "restir.h"

#ifndef APT_RESTIR_INST
#define APT_RESTIR_INST 0
#endif
#ifndef APT_RESTIR_TEX
#define APT_RESTIR_TEX 0
#endif

namespace apt {

namespace {

constexpr int TILE = 16;
constexpr int BLOCK = TILE * TILE;
// The kernels that shade and trace one pixel per thread run 8 x 8 tiles, one
// wave per work-group: a wave's slot is handed to the next tile as soon as
// its own longest ray is done, not when the slowest of four waves is (+2.4 %
// on the 1080p frame; the light per-pixel passes stay on 16 x 16 tiles).
constexpr int TILE8 = 8;
constexpr int BLOCK8 = TILE8 * TILE8;
constexpr float RESTIR_DI_EPSILON = 1e-6f;            // di_reservoir.wgsl:4
constexpr float GOLDEN_ANGLE = 2.39996322972865332f;  // pi (3 - sqrt 5)
constexpr float GI_SKY_DISTANCE = 1000.0f;  // first hit of an escaped GI ray (inline_path_tracer.wgsl:246)
constexpr float GI_OFFSET = 1e-4f;          // GiReservoir::phat origin offset (inline_path_tracer.wgsl:323)
// Target floor of the GI reservoir, times luma(T f cos): keeps p^ > 0 wherever
// the continuation can carry light.  The reference's p^ is zero whenever its
// one-sample estimate is occluded, and such paths are dropped (a bias).
constexpr float GI_PHAT_FLOOR = 1e-3f;

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
// Pixel of this thread in a frame of T x T tiles (one work-group per tile).
template <int T = TILE>
__device__ __forceinline__ bool tile_pixel(const RestirParams& p, uint32_t& x, uint32_t& y,
                                           uint32_t& pix) {
  const uint32_t tiles_x = (p.width + T - 1) / T;
  const uint32_t t = xcd_block(blockIdx.x, gridDim.x);
  x = (t % tiles_x) * T + threadIdx.x % T;
  y = (t / tiles_x) * T + threadIdx.x / T;
  pix = y * p.width + x;
  return x < p.width && y < p.rows;
}

// Adds the wave's traced rays to one of RS_CTR_SLOTS counters of a kind
// (chosen by work-group): one address for every wave of a light kernel
// serialises its atomics.
__device__ __forceinline__ void count_rays(bool traced, uint32_t* ctr) {
  const unsigned long long m = __ballot(traced);
  if (m != 0ull && (int)__lane_id() == __ffsll(m) - 1)
    atomicAdd(ctr + (blockIdx.x % RS_CTR_SLOTS), (uint32_t)__popcll(m));
}

__device__ __forceinline__ uint32_t h16(float a) { return (uint32_t)__half_as_ushort(__float2half_rn(a)); }
__device__ __forceinline__ float f16lo(uint32_t v) { return __half2float(__ushort_as_half((unsigned short)(v & 0xFFFFu))); }
__device__ __forceinline__ float f16hi(uint32_t v) { return __half2float(__ushort_as_half((unsigned short)(v >> 16))); }

__device__ __forceinline__ float safe_div(float a, float b) { return b > 0.0f ? a / b : 0.0f; }  // math.wgsl:20-26
__device__ __forceinline__ v3 safe_div3(v3 a, v3 b) { return mk(safe_div(a.x, b.x), safe_div(a.y, b.y), safe_div(a.z, b.z)); }
__device__ __forceinline__ v3 f4v(float4 a) { return mk(a.x, a.y, a.z); }
__device__ __forceinline__ float4 v4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ bool finite3(v3 a) { return a.x == a.x && a.y == a.y && a.z == a.z; }
__device__ __forceinline__ v3 lerp3(v3 a, v3 b, float t) { return a * (1.0f - t) + b * t; }

// G-buffer texel as the reuse passes see it: FP16 depth and normal.
struct Tex {
  v3 pos;
  float depth;
  v3 n;
  v3 albedo;
};

__device__ __forceinline__ GTexel pack_texel(v3 pos, float depth, v3 n, v3 albedo, uint32_t matbits) {
  GTexel t;
  t.pos_depth = make_float4(pos.x, pos.y, pos.z, depth);
  t.packed = make_uint4(h16(depth) | (h16(n.x) << 16), h16(n.y) | (h16(n.z) << 16), pack_rgb9e5(albedo), matbits);
  return t;
}

__device__ __forceinline__ Tex unpack_texel(const GTexel& g) {
  Tex t;
  t.pos = f4v(g.pos_depth);
  t.depth = f16lo(g.packed.x);
  t.n = mk(f16hi(g.packed.x), f16lo(g.packed.y), f16hi(g.packed.y));
  t.albedo = unpack_rgb9e5(g.packed.z);
  return t;
}

// GBuffer::sample_prev_gbuffer (gbuffer_bindings.wgsl:20-51): bilinear texel
// of the previous frame at a pixel-index position inside the slab.
__device__ __forceinline__ Tex sample_prev(const GTexel* g, const RestirParams& p, float px, float py) {
  const uint32_t ix = (uint32_t)floorf(px), iy = (uint32_t)floorf(py);
  const float fx = px - floorf(px), fy = py - floorf(py);
  const uint32_t x1 = min(ix + 1, p.width - 1), y1 = min(iy + 1, p.rows - 1);
  const Tex t00 = unpack_texel(g[iy * p.width + ix]), t10 = unpack_texel(g[iy * p.width + x1]);
  const Tex t01 = unpack_texel(g[y1 * p.width + ix]), t11 = unpack_texel(g[y1 * p.width + x1]);
  Tex r;
  r.pos = lerp3(lerp3(t00.pos, t10.pos, fx), lerp3(t01.pos, t11.pos, fx), fy);
  r.depth = lerpf(lerpf(t00.depth, t10.depth, fx), lerpf(t01.depth, t11.depth, fx), fy);
  r.n = normalize(lerp3(lerp3(t00.n, t10.n, fx), lerp3(t01.n, t11.n, fx), fy));
  r.albedo = lerp3(lerp3(t00.albedo, t10.albedo, fx), lerp3(t01.albedo, t11.albedo, fx), fy);
  return r;
}

// GBufferTexel::is_disoccluded (gbuffer_bindings.wgsl:53-57); true when the
// texels match.
__device__ __forceinline__ bool texel_similar(const Tex& a, const Tex& prev) {
  return fabsf(a.depth - prev.depth) <= 0.1f * prev.depth && dot(a.n, prev.n) > 0.906f;
}

// PackedGBufferTexel::depth_cs (gbuffer.wgsl:43-46) with near 0.001, far 10000.
__device__ __forceinline__ float depth_cs(float depth) {
  const float zn = 0.001f, zf = 10000.0f;
  const float zl = (depth - zn) / zf;
  return (zn * zf) / (zf - zl * (zf - zn));
}

// Pixel-index position (pixel centres at integers, slab rows) of a world point
// in the previous camera; stands in for the velocity texture
// (gbuffer.wgsl:25-53): prev = id - resolution * velocity.
__device__ __forceinline__ bool reproject(const RestirParams& p, v3 x, float& px, float& py) {
  const CamView& c = p.prev_cam;
  const v3 d = x - c.pos;
  const float lz = dot(d, c.back);
  if (!(lz < 0.0f)) return false;
  const float nx = dot(d, c.right) / (-lz * c.tan_half * c.aspect);
  const float ny = dot(d, c.up) / (-lz * c.tan_half);
  px = (nx + 1.0f) * 0.5f * (float)p.width - 0.5f;
  py = (1.0f - ny) * 0.5f * (float)p.height - 0.5f - (float)p.row_start;
  return px >= 0.0f && py >= 0.0f && px <= (float)(p.width - 1) && py <= (float)(p.rows - 1);
}

// ---------------------------------------------------------------------------
// shading point and BSDF
// ---------------------------------------------------------------------------
struct Vtx {
  v3 x, n, wo, T;
  v3 ncc;     // clearcoat lobe normal (n without clearcoat normal mapping)
  float2 uv;  // texture coordinates (textured scenes)
  float t;
  uint32_t mat;
  bool back;
};

__device__ __forceinline__ bool load_vtx(const SceneView& sc, const RestirBuffers& b, uint32_t i, Vtx& v) {
  const float4 p4 = b.vtx_p[i];
  if (!(p4.w > 0.0f)) return false;
  const float4 n4 = b.vtx_n[i];
  v.x = f4v(p4);
  v.t = p4.w;
  v.n = f4v(n4);
  const uint32_t bits = f2u(n4.w);
  v.mat = bits & 0x7FFFFFFFu;
  v.back = (bits >> 31) != 0u;
  v.wo = -f4v(b.ray_d[i]);
  v.T = f4v(b.vtx_t[i]);
  if (sc.tex) {
    const float4 m4 = b.vtx_m[i];
    v.uv = make_float2(m4.x, m4.y);
    v.ncc = sc.tri_tan ? unpack_oct30(f2u(m4.z)) : v.n;
  } else {
    v.uv = make_float2(0.0f, 0.0f);
    v.ncc = v.n;
  }
  return true;
}

struct Bsdf {
  Disney d;
  v3 albedo;
  ShadeFrame sf;
  bool lambert, back;
  float t;
};

// The BSDF at a shading point: material `mat` with its textures at uv, the
// shading frame of n and the clearcoat frame of ncc.
__device__ __forceinline__ Bsdf make_bsdf_at(const SceneView& sc, uint32_t mat, v3 n, v3 ncc, float2 uv, bool back,
                                             float t) {
  const apt_material& m = sc.mats[mat];
  Bsdf b;
  b.lambert = m.bsdf_model == APT_BSDF_LAMBERT;
  if (sc.tex) {
    const MatEval e = mat_eval(sc, m, uv);
    b.albedo = e.color;
    b.d = disney_from_eval(m, e);
  } else {
    b.albedo = mk(m.color[0], m.color[1], m.color[2]);
    b.d = disney_from_material(m);
  }
  b.sf.f = onb(n);
  b.sf.cc = sc.tri_tan ? onb(ncc) : b.sf.f;
  b.back = back;
  b.t = t;
  return b;
}

__device__ __forceinline__ Bsdf make_bsdf(const SceneView& sc, const Vtx& v) {
  return make_bsdf_at(sc, v.mat, v.n, v.ncc, v.uv, v.back, v.t);
}

__device__ __forceinline__ v3 bsdf_eval(const Bsdf& b, v3 wo, v3 wi, float& pdf) {
  return b.lambert ? lambert_eval(b.albedo, b.sf.f, wo, wi, pdf) : disney_eval(b.d, b.sf, wo, wi, b.back, pdf);
}

__device__ __forceinline__ BsdfSample bsdf_sample(const Bsdf& b, v3 wo, float r0, float r1, float r2) {
  return b.lambert ? lambert_sample(b.albedo, b.sf.f, wo, r0, r1)
                   : disney_sample(b.d, b.sf, wo, b.t, b.back, r0, r1, r2);
}

// ---------------------------------------------------------------------------
// light samples (light_sample.wgsl, helpers/nee.wgsl:16-115)
// ---------------------------------------------------------------------------
// Nee::sample_uniform: r0 picks the sun (probability sun_pick) or the
// emissive triangles, r2 the triangle (r1 would pick an instance; the scene
// here is one instance), (r3, r4) the point: cone coordinates for the sun,
// uniform barycentrics (sqrt warp) for a triangle.  The reference uses (r3, r4)
// as barycentrics directly, which leaves the triangle when r3 + r4 > 1.
__device__ __forceinline__ void sample_light(const SceneView& sc, float r0, float r2, float r3, float r4,
                                             uint32_t& light, float& u, float& v) {
  const float p_sun = sun_pick(sc);
  if (p_sun > 0.0f && (r0 < p_sun || sc.n_emissive == 0)) {
    light = LIGHT_SUN;
    u = r3;
    v = r4;
  } else if (sc.n_emissive > 0) {
    uint32_t k = (uint32_t)(r2 * (float)sc.n_emissive);
    k = k < sc.n_emissive ? k : sc.n_emissive - 1;
    light = sc.emissive[k];
    const float su = sqrtf(r3);
    u = su * (1.0f - r4);
    v = su * r4;
  } else {
    light = LIGHT_EMPTY;
    u = v = 0.0f;
  }
}

// Sky::inverse_direction_to_sun (sky_bindings.wgsl:57-72).
__device__ __forceinline__ void sun_uv(const SceneView& sc, v3 w, float& u, float& v) {
  const v3 s = sun_to(sc.sky);
  const v3 bt = perpendicular(s);
  const v3 tg = cross(bt, s);
  const float z = dot(w, s), x = dot(w, bt), y = dot(w, tg);
  const float h = cosf(sun_half_angle(sc.sky));
  u = fmodf(atan2f(y, x) / TWO_PI, 1.0f);
  v = (z - h) / (1.0f - h);
}

// LightSample::load_eval_data (nee.wgsl:16-57) evaluated from shading point x:
// the radiance toward x (triangles emit from their front side only; plain
// emission instead of the reference's solid-angle x 10 scale), direction,
// distance and the solid-angle density with which sample_light produces it.
struct LightHit {
  v3 Le, wi;
  float dist, pdf;
};

__device__ __forceinline__ bool eval_light(const SceneView& sc, uint32_t light, float u, float v, v3 x,
                                           LightHit& e) {
  const float p_sun = sun_pick(sc);
  if (light == LIGHT_SUN) {
    const v3 dir = normalize(perturb_direction(u, v, sun_to(sc.sky), sun_half_angle(sc.sky)));
    const v3 dv = dir * SUN_DISTANCE - x;
    e.dist = length(dv);
    e.wi = normalize(dv);
    e.Le = mk(sc.sky.sun_color[0], sc.sky.sun_color[1], sc.sky.sun_color[2]) * sun_intensity(sc.sky, dir);
    e.pdf = p_sun > 0.0f ? p_sun / sun_solid_angle(sc.sky) : 0.0f;
    return true;
  }
  if (light == LIGHT_EMPTY || sc.n_emissive == 0) return false;
  const TriGeom g = load_tri(sc, light);
  const v3 pt = g.p0 * (1.0f - u - v) + g.p1 * u + g.p2 * v;
  const v3 nl = normalize(cross(g.p1 - g.p0, g.p2 - g.p0));
  const v3 dv = pt - x;
  e.dist = length(dv);
  e.wi = normalize(dv);
  const float cos_l = -dot(nl, e.wi);
  const apt_material& lm = sc.mats[sc.tri_mat[light]];
  if (cos_l > 0.0f) {
    e.Le = mat_emission(sc, lm, hit_uv(sc, light, u, v));  // at the light point (nee.wgsl:39-51)
    e.pdf = (1.0f - p_sun) / (float)sc.n_emissive / tri_area(g) * (e.dist * e.dist) / cos_l;
  } else {
    e.Le = mk(0.0f, 0.0f, 0.0f);
    e.pdf = 0.0f;
  }
  return true;
}

// A DI reservoir's contribution weight is in solid angle at the shading point
// that made it: its light pdfs are p_A d^2 / cos_l there (solid-angle light
// pdfs, a documented deviation from helpers/nee.wgsl:83-85, whose area-measure
// pdfs need no conversion).  Reused at another shading point, the weight is
// scaled by the ratio of the two points' solid-angle pdfs of the same light
// point; a sun sample is a direction and keeps its weight.
__device__ __forceinline__ float di_jacobian(const SceneView& sc, uint32_t light, float u, float v, v3 here,
                                             v3 there) {
  if (light == LIGHT_SUN || light == LIGHT_EMPTY) return 1.0f;
  LightHit a, b;
  if (!eval_light(sc, light, u, v, here, a) || !eval_light(sc, light, u, v, there, b)) return 0.0f;
  return a.pdf > 0.0f ? b.pdf / a.pdf : 0.0f;
}

// trace_shadow_ray (helpers/trace.wgsl:20-84): opaque any-hit from the point
// offset along the normal, tmax = distance - 2 eps.
template <int W, int S, bool INST>
__device__ __forceinline__ bool visible(const SceneView& sc, const LdsStack& st, v3 x, v3 n, v3 wi, float dist,
                                        uint32_t* ctr) {
  TraceRay ray = make_trace_ray(x + n * TRACE_EPSILON, wi, 0.0f, dist - 2.0f * TRACE_EPSILON);
  count_rays(true, ctr + RS_CTR_SLOTS);
  return !trace_any<W, S, INST>(static_cast<const WNode<W>*>(sc.nodes), sc.bvh_tris, sc.root, ray, st, sc.insts);
}

template <int W, int S, bool INST>
__device__ __forceinline__ HitRec closest(const SceneView& sc, const LdsStack& st, v3 o, v3 d, uint32_t* ctr) {
  TraceRay ray = make_trace_ray(o, d, 0.0f, 1e30f);
  count_rays(true, ctr);
  return trace_closest<W, S, INST>(static_cast<const WNode<W>*>(sc.nodes), sc.bvh_tris, sc.root, ray, st,
                                   sc.insts);
}

// LightSample::phat (di_reservoir.wgsl:58-97): luma(f Le cos), with the
// visibility test when `vis`.
template <int W, int S, bool INST>
__device__ __forceinline__ float di_phat(const SceneView& sc, const LdsStack& st, const Vtx& v, const Bsdf& bs,
                                         uint32_t light, float u, float w, bool vis, uint32_t* ctr) {
  LightHit e;
  if (!eval_light(sc, light, u, w, v.x, e)) return 0.0f;
  const float c = dot(e.wi, v.n);
  if (!(c > 0.0f)) return 0.0f;
  // the BSDF first: nothing of it stays live through the shadow ray, and a
  // zero target needs no ray
  float pdf;
  const float ph = luma(bsdf_eval(bs, v.wo, e.wi, pdf) * e.Le * c);
  if (vis && ph != 0.0f && !visible<W, S, INST>(sc, st, v.x, v.n, e.wi, e.dist, ctr)) return 0.0f;
  return ph;
}

// DiReservoir::update (di_reservoir.wgsl:46-56); one draw per call.
__device__ __forceinline__ bool res_update(float& wsum, float& M, float w, uint32_t& rng) {
  wsum += w;
  M += 1.0f;
  return rng_next(rng) <= w / wsum;
}

__device__ __forceinline__ DiRes di_empty() { return DiRes{LIGHT_EMPTY, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f}; }

// The two updates of DiReservoir::combine / combine_unbiased (:99-133);
// `kept` tells whose sample the result holds: 0 none (empty), 1 r1's, 2 r2's.
__device__ __forceinline__ DiRes di_merge(const DiRes& r1, const DiRes& r2, uint32_t& rng, int& kept) {
  DiRes c = di_empty();
  kept = 0;
  if (res_update(c.wsum, c.M, r1.phat * r1.W * r1.M, rng)) {
    c.light = r1.light; c.u = r1.u; c.v = r1.v; c.phat = r1.phat;
    kept = 1;
  }
  if (res_update(c.wsum, c.M, r2.phat * r2.W * r2.M, rng)) {
    c.light = r2.light; c.u = r2.u; c.v = r2.v; c.phat = r2.phat;
    kept = 2;
  }
  c.M = r1.M + r2.M;
  return c;
}

// InlinePathTracer::trace with the approx config (max_bounces 1, no RIS):
// radiance toward x carried by direction wi with path throughput Tf, and the
// first hit (or a far point on an escaped ray).  Next-event estimation at the
// hit takes one light sample (5 draws), visibility tested.
template <int W, int S, bool INST>
__device__ __forceinline__ v3 approx_trace(const SceneView& sc, const LdsStack& st, v3 x, v3 wi, v3 Tf, v3& first,
                                           uint32_t& rng, uint32_t* ctr) {
  const v3 o = x + wi * TRACE_EPSILON;
  const HitRec h = closest<W, S, INST>(sc, st, o, wi, ctr);
  if (h.prim == INVALID) {
    first = x + wi * GI_SKY_DISTANCE;
    return Tf * sky_radiance(sc, wi);
  }
  const TriGeom g = load_tri(sc, h.prim);
  const v3 x2 = o + wi * h.t;
  first = x2;
  v3 n2 = normalize(cross(g.p1 - g.p0, g.p2 - g.p0));
  if (sc.tri_nrm) {
    const float4* tn = sc.tri_nrm + 3 * (size_t)h.prim;
    n2 = normalize(ld3(tn[0]) * (1.0f - h.u - h.v) + ld3(tn[1]) * h.u + ld3(tn[2]) * h.v);
  }
  const v3 wo2 = -wi;
  const uint32_t mat2 = sc.tri_mat[h.prim];
  const float2 uv2 = hit_uv(sc, h.prim, h.u, h.v);
  v3 nm2, ncc2;
  map_normals(sc, sc.mats[mat2], h.prim, h.u, h.v, uv2, n2, nm2, ncc2);
  const bool back = dot(wo2, n2) < 0.0f;
  n2 = back ? -nm2 : nm2;
  ncc2 = back ? -ncc2 : ncc2;
  const float r0 = rng_next(rng);
  rng_next(rng);  // r1: instance pick
  const float r2 = rng_next(rng), r3 = rng_next(rng), r4 = rng_next(rng);
  uint32_t light;
  float lu, lv;
  sample_light(sc, r0, r2, r3, r4, light, lu, lv);
  LightHit e;
  if (!eval_light(sc, light, lu, lv, x2, e) || !(e.pdf > 0.0f)) return mk(0.0f, 0.0f, 0.0f);
  const float c = dot(e.wi, n2);
  if (!(c > 0.0f) || !(maxc(e.Le) > 0.0f)) return mk(0.0f, 0.0f, 0.0f);
  if (!visible<W, S, INST>(sc, st, x2, n2, e.wi, e.dist, ctr)) return mk(0.0f, 0.0f, 0.0f);
  const Bsdf b2 = make_bsdf_at(sc, mat2, n2, ncc2, uv2, back, h.t);
  float pdf;
  return Tf * bsdf_eval(b2, wo2, e.wi, pdf) * e.Le * (c / e.pdf);
}

// GiReservoir::phat (inline_path_tracer.wgsl:299-331): replays the approximate
// estimate of sample x2 from shading point v with the stored RNG state.
template <int W, int S, bool INST>
__device__ __forceinline__ float gi_phat(const SceneView& sc, const LdsStack& st, const Vtx& v, const Bsdf& bs,
                                         v3 T, const GiRes& g, uint32_t* ctr) {
  const v3 wi = normalize(f4v(g.x2_rng) - v.x);
  float pdf;
  const v3 f = bsdf_eval(bs, v.wo, wi, pdf);
  const float c = fabsf(dot(wi, v.n));
  const v3 Tf = T * f * c;
  uint32_t r = f2u(g.x2_rng.w);
  v3 first;
  const v3 L = approx_trace<W, S, INST>(sc, st, v.x + wi * GI_OFFSET, wi, Tf, first, r, ctr);
  return luma(L) + GI_PHAT_FLOOR * luma(Tf);
}

// ---------------------------------------------------------------------------
// kernels: one sample
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_rs_raygen(RestirParams p, SceneView sc, RestirBuffers b) {
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  v3 o, d;
  camera_ray(sc, x, y + p.row_start, p.width, p.height, 0.5f, 0.5f, o, d);
  b.ray_o[pix] = v4(o, 0.0f);
  b.ray_d[pix] = v4(d, 0.0f);
  b.thr[pix] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  b.rng[pix] = path_seed(pix, p.seed);  // trace.wgsl:76
}

// Waves per SIMD the shading-heavy kernels that also trace (RIS, temporal and
// spatial reuse: a Disney BSDF evaluation and a visibility ray per candidate)
// are compiled for: at most 128 registers, four waves to hide the traversal's
// memory latency.  Left to itself the compiler uses ~210 registers and runs
// two waves, 15% slower end to end on the 1080p ReSTIR frame.
constexpr int RS_TRACE_WAVES = 4;
// The DI RIS (33 BSDF evaluations, its rays traced by the pool) fits four
// waves per SIMD without spills once its first round (area candidate 0 and
// the BSDF candidate) is peeled off the loop; five spill.
constexpr int RS_RIS_WAVES = 4;
// The textured (material textures, normal maps) and two-level (instance
// traversal) variants hold more live state; at four waves they spill (28 to
// 70 registers), so they run three.
constexpr int RS_WIDE_WAVES = 3;

// ---- ray stages --------------------------------------------------------------
// The rays that dominate the frame (the path's extension ray, the BSDF
// candidate of the DI RIS and the visibility of its winner) are written per
// pixel, traced by the persistent pool kernel over the slab's pixels (pixels
// without a ray hold an empty one, tmax < 0, which ends at the root), and the
// next kernel reads each pixel's result.  The traversal is the one every other
// stage runs, so the results equal tracing in place.

__global__ void k_rq_reset(RestirBuffers b) { pool_zero(b.rq_ctr + RQ_POOL, threadIdx.x); }

// Stores pixel pix's ray of the stage (rq_o.w = 1 marks a ray, else empty).
__device__ __forceinline__ void rq_put(const RestirBuffers& b, bool want, uint32_t pix, v3 o, v3 d, float tmax) {
  b.rq_o[pix] = v4(o, want ? 1.0f : 0.0f);
  b.rq_d[pix] = v4(d, want ? tmax : -1.0f);
}

// Rays of a stage: rq_o/rq_d per pixel (EXT: the live paths' extension rays,
// read straight from the path state, safe_origin along the ray as in
// trace.wgsl:91-98); results to rq_hit per pixel.
template <int W, bool EXT>
struct RqSrc {
  static constexpr bool kOneScene = true;
  const WNode<W>* nodes;
  const float4* tris;
  uint32_t root;
  RestirBuffers b;
  __device__ __forceinline__ void fetch(uint32_t i, TraceRay& ray) const {
    if (EXT) {
      const bool live = b.thr[i].w > 0.0f;
      const v3 d = f4v(b.ray_d[i]);
      ray = make_trace_ray(f4v(b.ray_o[i]) + d * TRACE_EPSILON, d, 0.0f, live ? 1e30f : -1.0f);
    } else {
      const float4 o = b.rq_o[i], d = b.rq_d[i];
      ray = make_trace_ray(f4v(o), f4v(d), 0.0f, d.w);
    }
  }
  __device__ __forceinline__ void closest(uint32_t i, const HitRec& h) const {
    b.rq_hit[i] = make_float4(h.t, h.u, h.v, u2f(h.prim));
  }
  __device__ __forceinline__ void any(uint32_t i, bool hit) const {
    b.rq_hit[i] = make_float4(0.0f, 0.0f, 0.0f, hit ? 1.0f : 0.0f);
  }
};

template <int W, int S, bool ANY, bool EXT, bool INST>
__global__ __launch_bounds__(BLOCK, INST ? 6 : 8) void k_rs_trace(SceneView sc, RestirBuffers b, uint32_t n) {
  __shared__ uint2 s_stack[pool_lds_entries(S) * BLOCK];
  __shared__ float4 s_top[TOP_BYTES / 16];
  const WNode<W>* nodes = static_cast<const WNode<W>*>(sc.nodes);
  const uint32_t ktop = stage_top<W>(nodes, sc.n_nodes, s_top);
  RqSrc<W, EXT> src{nodes, sc.bvh_tris, sc.root, b};
  trace_pool<W, S, ANY, RqSrc<W, EXT>, INST>(src, n, b.rq_ctr + RQ_POOL, min(sc.refill, 64u), s_stack + threadIdx.x,
                                             BLOCK, s_top, ktop, sc.insts);
}

template <bool TEX>
__global__ __launch_bounds__(BLOCK) void k_rs_extend(RestirParams p, SceneView sc_, RestirBuffers b, uint32_t bounce) {
  const SceneView sc = scene_view<TEX>(sc_);
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  const float4 t4 = b.thr[pix];
  count_rays(t4.w > 0.0f, b.counters);
  if (!(t4.w > 0.0f)) {
    b.vtx_p[pix].w = -1.0f;
    return;
  }
  const v3 T = f4v(t4);
  const v3 d = f4v(b.ray_d[pix]);
  const v3 o = f4v(b.ray_o[pix]) + d * TRACE_EPSILON;  // safe_origin along the ray (trace.wgsl:91-98)
  const float4 hq = b.rq_hit[pix];                     // k_rs_trace<.., EXT>
  HitRec h;
  h.t = hq.x;
  h.u = hq.y;
  h.v = hq.z;
  h.prim = f2u(hq.w);
  const float4 L = b.radiance[pix];
  if (h.prim == INVALID) {
    // sky without the sun disk, which the light samples cover (trace.wgsl:237-243)
    const v3 sky = T * sky_radiance(sc, d);
    b.radiance[pix] = make_float4(L.x + sky.x, L.y + sky.y, L.z + sky.z, 0.0f);
    b.thr[pix].w = 0.0f;
    b.vtx_p[pix].w = -1.0f;
    if (bounce == 0)
      b.gbuf[p.cur][pix] = pack_texel(mk(0.0f, 0.0f, 0.0f), 0.0f, mk(0.0f, 0.0f, 0.0f), mk(1.0f, 1.0f, 1.0f), 0u);
    return;
  }
  const TriGeom g = load_tri(sc, h.prim);
  const v3 hp = o + d * h.t;
  const v3 ng = normalize(cross(g.p1 - g.p0, g.p2 - g.p0));
  v3 ns = ng;
  if (sc.tri_nrm) {
    const float4* tn = sc.tri_nrm + 3 * (size_t)h.prim;
    ns = normalize(ld3(tn[0]) * (1.0f - h.u - h.v) + ld3(tn[1]) * h.u + ld3(tn[2]) * h.v);
  }
  const uint32_t mid = sc.tri_mat[h.prim];
  const apt_material& m = sc.mats[mid];
  const float2 uv = hit_uv(sc, h.prim, h.u, h.v);
  // normal mapping on the interpolated normal, then the face flip (trace.wgsl:143-191)
  v3 n_map, n_cc;
  map_normals(sc, m, h.prim, h.u, h.v, uv, ns, n_map, n_cc);
  const bool back = dot(-d, ns) < 0.0f;
  ns = back ? -n_map : n_map;
  n_cc = back ? -n_cc : n_cc;
  // emission at the first vertex only (trace.wgsl:139-141), front side only
  if (bounce == 0 && has_emission(m) && dot(ng, d) < 0.0f) {
    const v3 e = T * mat_emission(sc, m, uv);
    b.radiance[pix] = make_float4(L.x + e.x, L.y + e.y, L.z + e.z, 0.0f);
  }
  const uint32_t bits = mid | (back ? 0x80000000u : 0u);
  b.vtx_p[pix] = v4(hp, h.t);
  b.vtx_n[pix] = v4(ns, u2f(bits));
  b.vtx_t[pix] = v4(T, 0.0f);
  if (sc.tex) b.vtx_m[pix] = make_float4(uv.x, uv.y, u2f(pack_oct30(n_cc)), 0.0f);
  if (bounce == 0) {
    // albedo = colour + emission (trace.wgsl:218-223)
    v3 alb;
    if (sc.tex) {
      const MatEval e = mat_eval(sc, m, uv);
      alb = e.color + e.emission;
    } else {
      alb = mk(m.color[0] + m.emission[0], m.color[1] + m.emission[1], m.color[2] + m.emission[2]);
    }
    b.gbuf[p.cur][pix] = pack_texel(hp, h.t, ns, alb, bits);
  }
}

// Nee::sample_ris (nee.wgsl:122-237) with di_candidates area candidates and
// one BSDF candidate.  MIS is the balance heuristic over each candidate's own
// densities under both strategies (the reference mixes the densities of two
// different candidates), so W = wsum / p^ is unbiased.
template <bool TEX>
__global__ __launch_bounds__(BLOCK8, TEX ? RS_WIDE_WAVES : RS_RIS_WAVES) void k_rs_di_ris(RestirParams p, SceneView sc_, RestirBuffers b) {
  const SceneView sc = scene_view<TEX>(sc_);
  uint32_t x, y, pix;
  if (!tile_pixel<TILE8>(p, x, y, pix)) return;
  Vtx v;
  if (!load_vtx(sc, b, pix, v)) {
    rq_put(b, false, pix, mk(0.0f, 0.0f, 0.0f), mk(0.0f, 0.0f, 1.0f), 0.0f);
    return;
  }
  const Bsdf bs = make_bsdf(sc, v);
  uint32_t rng = b.rng[pix];
  const float p_sun = sun_pick(sc);
  const float na = (float)p.di_candidates, nb = 1.0f;
  DiRes r = di_empty();
  // One area-light candidate (five draws): light, (u, v), target and the two
  // densities of its direction.
  const auto light_candidate = [&](uint32_t& light, float& u, float& w, float& phat, float& pl, float& pb) {
    light = LIGHT_EMPTY;
    u = w = phat = pl = pb = 0.0f;
    const float r0 = rng_next(rng);
    rng_next(rng);  // r1: instance pick
    const float r2 = rng_next(rng), r3 = rng_next(rng), r4 = rng_next(rng);
    sample_light(sc, r0, r2, r3, r4, light, u, w);
    LightHit e;
    if (eval_light(sc, light, u, w, v.x, e)) {
      pl = e.pdf;
      const float c = dot(e.wi, v.n);
      if (c > 0.0f && pl > 0.0f) {
        const v3 f = bsdf_eval(bs, v.wo, e.wi, pb);
        if (pb > 0.0f) phat = luma(f * e.Le * c);
      }
    }
  };
  const auto update = [&](uint32_t light, float u, float w_, float phat, float pl, float pb) {
    const float w = pl > 0.0f ? phat / (na * pl + nb * pb) : 0.0f;
    if (res_update(r.wsum, r.M, w, rng)) {
      r.light = light; r.u = u; r.v = w_; r.phat = phat;
    }
  };
  // First round: area candidate 0, then the BSDF candidate (three draws),
  // then their two reservoir updates; later rounds: one area candidate each.
  uint32_t a_light = LIGHT_EMPTY;
  float a_u = 0.0f, a_v = 0.0f, a_phat = 0.0f, a_pl = 0.0f, a_pb = 0.0f;
  if (p.di_candidates > 0u) light_candidate(a_light, a_u, a_v, a_phat, a_pl, a_pb);
  uint32_t b_light = LIGHT_EMPTY;
  float b_u = 0.0f, b_v = 0.0f, b_phat = 0.0f, b_pl = 0.0f, b_pb = 0.0f;
  {
    const float r0 = rng_next(rng), r1 = rng_next(rng), r2 = rng_next(rng);
    const BsdfSample s = bsdf_sample(bs, v.wo, r0, r1, r2);
    b_pb = s.pdf;
    const v3 contrib = s.f * fabsf(dot(s.wi, v.n));
    if (s.pdf > 0.0f && dot(contrib, contrib) > 0.0f) {
      const float4 hq = b.rq_hit[pix];  // the ray k_rs_di_bsdf_ray queued for this sample
      HitRec h;
      h.u = hq.y;
      h.v = hq.z;
      h.prim = f2u(hq.w);
      if (h.prim != INVALID) {
        if (has_emission(sc.mats[sc.tri_mat[h.prim]])) {
          b_light = h.prim;
          b_u = h.u;
          b_v = h.v;
        }
      } else if (p_sun > 0.0f) {
        b_light = LIGHT_SUN;
        sun_uv(sc, s.wi, b_u, b_v);
      }
      LightHit e;
      if (b_light != LIGHT_EMPTY && eval_light(sc, b_light, b_u, b_v, v.x, e)) {
        b_pl = e.pdf;
        b_phat = luma(contrib * e.Le);
      }
    }
  }
  if (p.di_candidates > 0u) update(a_light, a_u, a_v, a_phat, a_pl, a_pb);
  update(b_light, b_u, b_v, b_phat, b_pl, b_pb);
  for (uint32_t i = 1; i < p.di_candidates; ++i) {
    light_candidate(a_light, a_u, a_v, a_phat, a_pl, a_pb);
    update(a_light, a_u, a_v, a_phat, a_pl, a_pb);
  }
  // W = wsum / phat once the winner's shadow ray (queued here) finds it
  // visible: k_rs_di_vis
  r.W = 0.0f;
  LightHit e{};
  const bool shoot = r.phat > RESTIR_DI_EPSILON && eval_light(sc, r.light, r.u, r.v, v.x, e);
  b.di[pix] = r;
  b.rng[pix] = rng;
  count_rays(shoot, b.counters + RS_CTR_SLOTS);
  rq_put(b, shoot, pix, v.x + v.n * TRACE_EPSILON, e.wi, e.dist - 2.0f * TRACE_EPSILON);
}

// The BSDF candidate of k_rs_di_ris: the same draws (after the first area
// candidate's five), the same sample, its closest-hit ray queued.
template <bool TEX>
__global__ __launch_bounds__(BLOCK) void k_rs_di_bsdf_ray(RestirParams p, SceneView sc_, RestirBuffers b) {
  const SceneView sc = scene_view<TEX>(sc_);
  if (blockIdx.x == 0) pool_zero(b.rq_ctr + RQ_POOL, threadIdx.x);  // for the trace that follows
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  Vtx v;
  const bool have = load_vtx(sc, b, pix, v);
  bool want = false;
  v3 wi = mk(0.0f, 0.0f, 0.0f);
  if (have) {
    const Bsdf bs = make_bsdf(sc, v);
    uint32_t rng = b.rng[pix];
    if (p.di_candidates > 0u)
      for (int k = 0; k < 5; ++k) rng_next(rng);
    const float r0 = rng_next(rng), r1 = rng_next(rng), r2 = rng_next(rng);
    const BsdfSample s = bsdf_sample(bs, v.wo, r0, r1, r2);
    const v3 contrib = s.f * fabsf(dot(s.wi, v.n));
    want = s.pdf > 0.0f && dot(contrib, contrib) > 0.0f;
    wi = s.wi;
  }
  count_rays(want, b.counters);
  rq_put(b, want, pix, have ? v.x + v.n * TRACE_EPSILON : wi, wi, 1e30f);
}

// trace_shadow_ray's verdict on the winners' rays: visible ones get
// W = wsum / phat (nee.wgsl:226-235).
__global__ __launch_bounds__(BLOCK) void k_rs_di_vis(RestirBuffers b, uint32_t npix) {
  const uint32_t pix = blockIdx.x * BLOCK + threadIdx.x;
  if (pix >= npix || b.rq_o[pix].w == 0.0f || b.rq_hit[pix].w != 0.0f) return;
  DiRes r = b.di[pix];
  r.W = (1.0f / r.phat) * r.wsum;
  b.di[pix] = r;
}

// Roulette and InlinePathTracer::sample_ris (inline_path_tracer.wgsl:258-296):
// one BSDF candidate whose target is the approximate one-bounce estimate.
// Roulette keeps the path with probability min(1, max T) (the reference
// divides by max T even above 1), and a path it ends keeps this vertex's DI.
template <int W, int S, bool INST, bool TEX>
__global__ __launch_bounds__(BLOCK8) void k_rs_gi_ris(RestirParams p, SceneView sc_, RestirBuffers b, uint32_t bounce) {
  const SceneView sc = scene_view<TEX>(sc_);
  __shared__ uint2 s_stack[S * BLOCK8];
  uint32_t x, y, pix;
  if (!tile_pixel<TILE8>(p, x, y, pix)) return;
  Vtx v;
  if (!load_vtx(sc, b, pix, v)) return;
  const LdsStack st{s_stack + threadIdx.x, BLOCK8};
  uint32_t rng = b.rng[pix];
  v3 T = v.T;
  GiRes g;
  g.x2_rng = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  g.mwsp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  bool alive = true;
  if (bounce > 1) {  // trace.wgsl:221-230
    const float q = fminf(1.0f, maxc(T));
    if (q < rng_next(rng)) alive = false;
    else T = T * (1.0f / q);
  }
  if (alive) {
    const Bsdf bs = make_bsdf(sc, v);
    const float r0 = rng_next(rng), r1 = rng_next(rng), r2 = rng_next(rng);
    const BsdfSample s = bsdf_sample(bs, v.wo, r0, r1, r2);
    float M = 0.0f, wsum = 0.0f, phat = 0.0f;
    if (s.pdf > 1e-6f) {
      const float c = fabsf(dot(s.wi, v.n));
      const v3 Tf = T * s.f * c;
      const uint32_t phat_rng = rng;
      v3 first;
      const v3 L = approx_trace<W, S, INST>(sc, st, v.x, s.wi, Tf, first, rng, b.counters);
      const float ph = luma(L) + GI_PHAT_FLOOR * luma(Tf);
      if (res_update(wsum, M, ph / s.pdf, rng)) {
        g.x2_rng = v4(first, u2f(phat_rng));
        phat = ph;
      }
    }
    const float Wc = (phat > 0.0f && M * wsum > 0.0f) ? (1.0f / phat) * (1.0f / M * wsum) : 0.0f;
    g.mwsp = make_float4(M, Wc, wsum, phat);
  }
  b.thr[pix] = v4(T, alive ? 1.0f : 0.0f);
  b.gi[pix] = g;
  b.rng[pix] = rng;
}

// restir_di_temporal.wgsl:68-176 (nearest previous pixel, stochastic
// rounding), with the visibility-tested target.
template <int W, int S, bool INST, bool TEX>
__global__ __launch_bounds__(BLOCK8, RS_TRACE_WAVES) void k_rs_di_temporal(RestirParams p, SceneView sc_, RestirBuffers b) {
  const SceneView sc = scene_view<TEX>(sc_);
  __shared__ uint2 s_stack[S * BLOCK8];
  uint32_t x, y, pix;
  if (!tile_pixel<TILE8>(p, x, y, pix)) return;
  Vtx v;
  if (!load_vtx(sc, b, pix, v)) {
    b.di_tmp[pix] = di_empty();
    b.di_hist[p.cur][pix] = di_empty();
    return;
  }
  const LdsStack st{s_stack + threadIdx.x, BLOCK8};
  DiRes r = b.di[pix];
  uint32_t rng = b.rng[pix];
  float px, py;
  if (reproject(p, v.x, px, py)) {
    const Tex cur = unpack_texel(b.gbuf[p.cur][pix]);
    const Tex prev = sample_prev(b.gbuf[p.cur ^ 1], p, px, py);
    if (texel_similar(cur, prev)) {
      const float ux = rng_next(rng), uy = rng_next(rng);
      const uint32_t qx = min((uint32_t)floorf(px + ux), p.width - 1);
      const uint32_t qy = min((uint32_t)floorf(py + uy), p.rows - 1);
      DiRes pr = b.di_hist[p.cur ^ 1][qy * p.width + qx];
      pr.W *= di_jacobian(sc, pr.light, pr.u, pr.v, v.x, f4v(b.gbuf[p.cur ^ 1][qy * p.width + qx].pos_depth));
      pr.M = fminf(pr.M, (float)p.di_mcap * r.M);
      const Bsdf bs = make_bsdf(sc, v);
      pr.phat = di_phat<W, S, INST>(sc, st, v, bs, pr.light, pr.u, pr.v, true, b.counters);
      int kept;
      DiRes c = di_merge(r, pr, rng, kept);
      if (c.phat > RESTIR_DI_EPSILON && c.M * c.wsum > RESTIR_DI_EPSILON)
        c.W = (1.0f / c.phat) * (1.0f / c.M * c.wsum);
      r = c;
    }
  }
  b.di_tmp[pix] = r;
  if (p.neighbours == 0) {
    b.di[pix] = r;
    b.di_hist[p.cur][pix] = r;
  }
  b.rng[pix] = rng;
}

// Before the spatial pass: each pixel's temporal sample at the pixel itself.
// combine_unbiased needs that sample's visibility-tested target at the pixel
// (while the pixel keeps its own sample) and at every pixel whose reservoir
// takes it from this one (as a neighbour): the same number, so it is
// evaluated once here, the BSDF part into di_self and the shadow ray (the
// one visible() would trace) queued for the pool kernel.
template <bool TEX>
__global__ __launch_bounds__(BLOCK) void k_rs_di_self(RestirParams p, SceneView sc_, RestirBuffers b) {
  const SceneView sc = scene_view<TEX>(sc_);
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  Vtx v;
  float ph = 0.0f;
  bool shoot = false;
  v3 o = mk(0.0f, 0.0f, 0.0f), wi = mk(0.0f, 0.0f, 1.0f);
  float tmax = 0.0f;
  if (load_vtx(sc, b, pix, v)) {
    const DiRes r = b.di_tmp[pix];
    LightHit e;
    if (eval_light(sc, r.light, r.u, r.v, v.x, e)) {
      const float c = dot(e.wi, v.n);
      if (c > 0.0f) {
        const Bsdf bs = make_bsdf(sc, v);
        float pdf;
        ph = luma(bsdf_eval(bs, v.wo, e.wi, pdf) * e.Le * c);
        shoot = ph != 0.0f;
        o = v.x + v.n * TRACE_EPSILON;
        wi = e.wi;
        tmax = e.dist - 2.0f * TRACE_EPSILON;
      }
    }
  }
  b.di_self[pix] = ph;
  count_rays(shoot, b.counters + RS_CTR_SLOTS);
  rq_put(b, shoot, pix, o, wi, tmax);
}

// di_phat(.., vis = true) > 0 of pixel i's own temporal sample at i.
__device__ __forceinline__ bool di_self_visible(const RestirBuffers& b, uint32_t i) {
  return b.di_self[i] > 0.0f && b.rq_hit[i].w == 0.0f;
}

// interleaved_gradient_noise_animated (shared/sampling.wgsl:64-74).
__device__ __forceinline__ float fractf(float a) { return a - floorf(a); }
__device__ __forceinline__ float ign_animated(uint32_t x, uint32_t y, uint32_t frame) {
  const uint32_t id = frame % 64u;
  const float fx = (float)x + 5.588238f * (float)id, fy = (float)y + 5.588238f * (float)id;
  return fractf(52.9829189f * fractf(0.06711056f * fx + 0.00583715f * fy));
}

// mirror (restir_di_spatial.wgsl:59-61), clamped into the frame (the
// reference's returns `max` for x = +-max).
__device__ __forceinline__ uint32_t mirror_px(int x, int m) {
  const int r = ((x + m) % (2 * m)) - m;
  const int a = r < 0 ? -r : r;
  return (uint32_t)(a < m ? a : m - 1);
}

// restir_di_spatial.wgsl:70-162 (one pass, unbiased combination).
template <int W, int S, bool INST, bool TEX>
__global__ __launch_bounds__(BLOCK8, TEX || INST ? RS_WIDE_WAVES : RS_TRACE_WAVES) void k_rs_di_spatial(RestirParams p, SceneView sc_, RestirBuffers b) {
  const SceneView sc = scene_view<TEX>(sc_);
  __shared__ uint2 s_stack[S * BLOCK8];
  uint32_t x, y, pix;
  if (!tile_pixel<TILE8>(p, x, y, pix)) return;
  Vtx v;
  if (!load_vtx(sc, b, pix, v)) {
    b.di[pix] = di_empty();
    return;
  }
  const LdsStack st{s_stack + threadIdx.x, BLOCK8};
  DiRes r = b.di_tmp[pix];
  uint32_t rng = b.rng[pix];
  // The pixel's BSDF (about 50 registers) is rebuilt from its material at
  // each use instead of staying live through the taps' shadow rays.
  const auto pixel_bsdf = [&]() {
    uint32_t m = v.mat;
    asm volatile("" : "+v"(m));
    return make_bsdf_at(sc, m, v.n, v.ncc, v.uv, v.back, v.t);
  };
  const Tex ct = unpack_texel(b.gbuf[p.cur][pix]);
  const float cd = depth_cs(ct.depth);
  float radius = p.radius > 0.0f ? p.radius : (float)(p.width + p.rows) / 2.0f * 0.05f;
  const uint32_t pass_seed = p.seed;  // seed * spatial_pass_count + pass index, one pass
  const float r_off = ign_animated(x, y, pass_seed);
  const uint32_t aseed = hash_combine(x >> 2, hash_combine(y >> 2, pass_seed));
  const float a_off = (float)aseed * (1.0f / (float)0xFFFFFFFFu) * TWO_PI;
  const uint32_t n = p.neighbours;
  // Whether r's sample has a visibility-tested target > 0 at this pixel (0 /
  // 1).  It is a function of the sample alone, so it carries over while r
  // keeps its sample, and a kept neighbour sample brings its own (nr.phat
  // below): each shadow ray at this pixel is traced once.  The pixel's own
  // sample's was evaluated by k_rs_di_self.
  int r_vis = di_self_visible(b, pix) ? 1 : 0;
  for (uint32_t i = 0; i < n; ++i) {
    const float ang = (float)i * GOLDEN_ANGLE + a_off;
    const float rad = powf((float)i / (float)n, 0.5f) * radius + r_off;
    const int ox = (int)(rad * cosf(ang)), oy = (int)(rad * sinf(ang));
    const uint32_t nx = mirror_px((int)x + ox, (int)p.width), ny = mirror_px((int)y + oy, (int)p.rows);
    const uint32_t q = ny * p.width + nx;
    if (q == pix) continue;
    const GTexel ng = b.gbuf[p.cur][q];
    bool valid = ng.pos_depth.w != 0.0f;
    if (valid) {
      const Tex nt = unpack_texel(ng);
      valid = fabsf(cd - depth_cs(nt.depth)) / cd < 0.1f && dot(ct.n, nt.n) > 0.906f;
    }
    if (valid) {
      DiRes nr = b.di_tmp[q];
      nr.W *= di_jacobian(sc, nr.light, nr.u, nr.v, v.x, f4v(ng.pos_depth));
      nr.phat = di_phat<W, S, INST>(sc, st, v, pixel_bsdf(), nr.light, nr.u, nr.v, true, b.counters);
      valid = nr.phat > 0.0f;
      // DiReservoir::combine_unbiased: Z counts the M of the domains where the
      // chosen sample has a (visibility-tested) target > 0
      int kept;
      DiRes c = di_merge(r, nr, rng, kept);
      float z = 0.0f;
      int c_vis = 0;  // the merged (empty) reservoir's sample has no target
      if (kept == 2) {
        c_vis = nr.phat > 0.0f ? 1 : 0;
      } else if (kept == 1) {
        c_vis = r_vis;
      }
      if (c_vis) z += r.M;
      // The merged sample's visibility-tested target at the neighbour: the
      // neighbour's own sample's was evaluated by k_rs_di_self (the
      // neighbour's vertex is its G-buffer texel's position), another sample
      // is evaluated here; an empty merge has no target.
      if (kept == 2) {
        if (di_self_visible(b, q)) z += nr.M;
      } else if (kept == 1) {
        Vtx vn;
        if (load_vtx(sc, b, q, vn)) {
          vn.x = f4v(ng.pos_depth);  // the neighbour's G-buffer position
          const Bsdf bn = make_bsdf(sc, vn);
          if (di_phat<W, S, INST>(sc, st, vn, bn, c.light, c.u, c.v, true, b.counters) > 0.0f) z += nr.M;
        }
      }
      if (c.phat > RESTIR_DI_EPSILON && z * c.wsum > RESTIR_DI_EPSILON) c.W = (1.0f / c.phat) * (1.0f / z * c.wsum);
      r = c;
      r_vis = c_vis;
    }
    if (!valid) radius = fmaxf(radius * 0.5f, 3.0f);
  }
  b.di[pix] = r;
  b.di_hist[p.cur][pix] = r;
  b.rng[pix] = rng;
}

// restir_gi_temporal.wgsl:69-115 (M capped at gi_mcap, biased combination).
template <int W, int S, bool INST, bool TEX>
__global__ __launch_bounds__(BLOCK8) void k_rs_gi_temporal(RestirParams p, SceneView sc_, RestirBuffers b) {
  const SceneView sc = scene_view<TEX>(sc_);
  __shared__ uint2 s_stack[S * BLOCK8];
  uint32_t x, y, pix;
  if (!tile_pixel<TILE8>(p, x, y, pix)) return;
  Vtx v;
  GiRes r;
  r.x2_rng = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  r.mwsp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (!load_vtx(sc, b, pix, v)) {
    if (p.gi_spatial_passes == 0) b.gi_hist[p.cur][pix] = r;
    return;
  }
  r = b.gi[pix];
  const float4 t4 = b.thr[pix];
  if (t4.w > 0.0f) {
    const LdsStack st{s_stack + threadIdx.x, BLOCK8};
    uint32_t rng = b.rng[pix];
    float px, py;
    if (reproject(p, v.x, px, py)) {
      const float ux = rng_next(rng), uy = rng_next(rng);
      const uint32_t qx = min((uint32_t)floorf(px + ux), p.width - 1);
      const uint32_t qy = min((uint32_t)floorf(py + uy), p.rows - 1);
      const Tex cur = unpack_texel(b.gbuf[p.cur][pix]);
      const Tex prev = sample_prev(b.gbuf[p.cur ^ 1], p, px, py);
      if (texel_similar(cur, prev)) {
        const GiRes pr = b.gi_hist[p.cur ^ 1][qy * p.width + qx];
        const float pM = fminf(pr.mwsp.x, (float)p.gi_mcap);
        const Bsdf bs = make_bsdf(sc, v);
        const float pp = gi_phat<W, S, INST>(sc, st, v, bs, f4v(t4), pr, b.counters);
        // GiReservoir::combine (gi_reservoir.wgsl:66-78)
        float wsum = 0.0f, M = 0.0f, phat = 0.0f;
        GiRes c;
        c.x2_rng = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (res_update(wsum, M, r.mwsp.w * r.mwsp.y * r.mwsp.x, rng)) {
          c.x2_rng = r.x2_rng;
          phat = r.mwsp.w;
        }
        if (res_update(wsum, M, pp * pr.mwsp.y * pM, rng)) {
          c.x2_rng = pr.x2_rng;
          phat = pp;
        }
        M = r.mwsp.x + pM;
        const float Wc = (phat > 0.0f && M * wsum > 0.0f) ? (1.0f / phat) * (1.0f / M * wsum) : 0.0f;
        c.mwsp = make_float4(M, Wc, wsum, phat);
        r = c;
      }
    }
    b.rng[pix] = rng;
  }
  b.gi[pix] = r;
  if (p.gi_spatial_passes == 0) b.gi_hist[p.cur][pix] = r;  // else the last spatial pass writes it
}

// restir_gi_spatial.wgsl:70-165: pass `pass` of p.gi_spatial_passes, the
// unbiased combination over five taps (NUM_SAMPLES); the radius is 5 % of
// (width + rows) / 2 whatever the configured pixel radius, the seeds mix
// seed * 3 + pass, and the angle hash uses pixel / 4 on the first pass,
// pixel / 2 after.  Reads `in`, writes `out`; the last pass writes the
// history.  A tap's targets are approximate one-bounce paths (gi_phat); Z
// evaluates the chosen sample's at both domains only when W could be
// nonzero (a sample chosen, a positive weight sum).
template <int W, int S, bool INST, bool TEX>
__global__ __launch_bounds__(BLOCK8) void k_rs_gi_spatial(RestirParams p, SceneView sc_, RestirBuffers b, uint32_t pass,
                                                         const GiRes* __restrict__ in, GiRes* __restrict__ out) {
  const SceneView sc = scene_view<TEX>(sc_);
  __shared__ uint2 s_stack[S * BLOCK8];
  uint32_t x, y, pix;
  if (!tile_pixel<TILE8>(p, x, y, pix)) return;
  const bool last = pass + 1 == p.gi_spatial_passes;
  GiRes r = in[pix];
  Vtx v;
  if (!load_vtx(sc, b, pix, v)) {
    out[pix] = r;
    if (last) {
      GiRes e;
      e.x2_rng = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      e.mwsp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      b.gi_hist[p.cur][pix] = e;
    }
    return;
  }
  const float4 t4 = b.thr[pix];
  if (t4.w > 0.0f) {
    const LdsStack st{s_stack + threadIdx.x, BLOCK8};
    uint32_t rng = b.rng[pix];
    const v3 T = f4v(t4);
    const Tex ct = unpack_texel(b.gbuf[p.cur][pix]);
    const float cd = depth_cs(ct.depth);
    float radius = (float)(p.width + p.rows) / 2.0f * 0.05f;
    const uint32_t pass_seed = p.seed * 3u + pass;
    const float r_off = ign_animated(x, y, pass_seed);
    const uint32_t sh = pass == 0 ? 2u : 1u;
    const uint32_t aseed = hash_combine(x >> sh, hash_combine(y >> sh, pass_seed));
    const float a_off = (float)aseed * (1.0f / (float)0xFFFFFFFFu) * TWO_PI;
    constexpr uint32_t n = 5;
    for (uint32_t i = 0; i < n; ++i) {
      const float ang = (float)i * GOLDEN_ANGLE + a_off;
      const float rad = powf((float)i / (float)n, 0.5f) * radius + r_off;
      const int ox = (int)(rad * cosf(ang)), oy = (int)(rad * sinf(ang));
      const uint32_t nx = mirror_px((int)x + ox, (int)p.width), ny = mirror_px((int)y + oy, (int)p.rows);
      const uint32_t q = ny * p.width + nx;
      if (q == pix) continue;
      const GTexel ng = b.gbuf[p.cur][q];
      bool valid = ng.pos_depth.w != 0.0f;
      if (valid) {
        const Tex nt = unpack_texel(ng);
        valid = fabsf(cd - depth_cs(nt.depth)) / cd < 0.1f && dot(ct.n, nt.n) > 0.906f;
      }
      if (valid) {
        GiRes nr = in[q];
        nr.mwsp.w = gi_phat<W, S, INST>(sc, st, v, make_bsdf(sc, v), T, nr, b.counters);
        valid = nr.mwsp.w > 0.0f;
        // GiReservoir::combine_unbiased (inline_path_tracer.wgsl:333-355)
        float wsum = 0.0f, M = 0.0f, phat = 0.0f;
        GiRes c;
        c.x2_rng = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        bool took_n = false;
        if (res_update(wsum, M, r.mwsp.w * r.mwsp.y * r.mwsp.x, rng)) {
          c.x2_rng = r.x2_rng;
          phat = r.mwsp.w;
        }
        if (res_update(wsum, M, nr.mwsp.w * nr.mwsp.y * nr.mwsp.x, rng)) {
          c.x2_rng = nr.x2_rng;
          phat = nr.mwsp.w;
          took_n = true;
        }
        M = r.mwsp.x + nr.mwsp.x;
        float z = 0.0f;
        if (phat > 0.0f && wsum > 0.0f) {
          // the neighbour's sample at this pixel is nr's phat, just evaluated
          if (took_n || gi_phat<W, S, INST>(sc, st, v, make_bsdf(sc, v), T, c, b.counters) > 0.0f) z += r.mwsp.x;
          Vtx vn;
          if (load_vtx(sc, b, q, vn)) {
            vn.x = f4v(ng.pos_depth);  // the neighbour's G-buffer position
            if (gi_phat<W, S, INST>(sc, st, vn, make_bsdf(sc, vn), f4v(b.thr[q]), c, b.counters) > 0.0f)
              z += nr.mwsp.x;
          }
        }
        const float Wc = (phat > 0.0f && z * wsum > 0.0f) ? (1.0f / phat) * (1.0f / z * wsum) : 0.0f;
        c.mwsp = make_float4(M, Wc, wsum, phat);
        r = c;
      }
      if (!valid) radius = fmaxf(radius * 0.5f, 3.0f);
    }
    b.rng[pix] = rng;
  }
  out[pix] = r;
  if (last) b.gi_hist[p.cur][pix] = r;
}

// apply_di.wgsl:48-111: the reservoir's sample, visibility tested, weighted by
// W, with the throughput at the vertex.
template <int W, int S, bool INST, bool TEX>
__global__ __launch_bounds__(BLOCK8) void k_rs_apply_di(RestirParams p, SceneView sc_, RestirBuffers b) {
  const SceneView sc = scene_view<TEX>(sc_);
  __shared__ uint2 s_stack[S * BLOCK8];
  uint32_t x, y, pix;
  if (!tile_pixel<TILE8>(p, x, y, pix)) return;
  Vtx v;
  if (!load_vtx(sc, b, pix, v)) return;
  const DiRes r = b.di[pix];
  if (r.W == 0.0f) return;
  LightHit e;
  if (!eval_light(sc, r.light, r.u, r.v, v.x, e)) return;
  const float c = dot(e.wi, v.n);
  if (!(c > 0.0f)) return;
  const LdsStack st{s_stack + threadIdx.x, BLOCK8};
  if (!visible<W, S, INST>(sc, st, v.x, v.n, e.wi, e.dist, b.counters)) return;
  const Bsdf bs = make_bsdf(sc, v);
  float pdf;
  const v3 add = v.T * bsdf_eval(bs, v.wo, e.wi, pdf) * e.Le * (c * r.W);
  if (!finite3(add)) return;
  const float4 L = b.radiance[pix];
  b.radiance[pix] = make_float4(L.x + add.x, L.y + add.y, L.z + add.z, 0.0f);
}

// apply_gi.wgsl:44-99: the next ray heads for the GI sample point.
template <bool TEX>
__global__ __launch_bounds__(BLOCK) void k_rs_apply_gi(RestirParams p, SceneView sc_, RestirBuffers b) {
  const SceneView sc = scene_view<TEX>(sc_);
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  Vtx v;
  if (!load_vtx(sc, b, pix, v)) return;
  const float4 t4 = b.thr[pix];
  if (!(t4.w > 0.0f)) return;
  const GiRes g = b.gi[pix];
  if (g.mwsp.y > 0.0f) {
    const v3 wi = normalize(f4v(g.x2_rng) - v.x);
    const Bsdf bs = make_bsdf(sc, v);
    float pdf;
    const v3 T = f4v(t4) * bsdf_eval(bs, v.wo, wi, pdf) * (fabsf(dot(v.n, wi)) * g.mwsp.y);
    const bool ok = finite3(T) && maxc(T) > 0.0f;
    b.thr[pix] = v4(T, ok ? 1.0f : 0.0f);
    b.ray_o[pix] = v4(v.x, 0.0f);
    b.ray_d[pix] = v4(wi, 0.0f);
  } else {
    b.thr[pix].w = 0.0f;
  }
}

// ---------------------------------------------------------------------------
// kernels: post chain
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_rs_demod(RestirParams p, RestirBuffers b) {
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  const v3 alb = unpack_rgb9e5(b.gbuf[p.cur][pix].packed.z);
  b.demod[p.cur][pix] = v4(safe_div3(f4v(b.radiance[pix]), alb), 0.0f);
}

// firefly_filter.wgsl: clamp to the luma range of the 3x3 neighbourhood
// (out of place; the reference filters in place while neighbours read).
__global__ __launch_bounds__(BLOCK) void k_rs_firefly(RestirParams p, RestirBuffers b, const float4* in,
                                                      float4* out) {
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  const float4 c4 = in[pix];
  if (b.gbuf[p.cur][pix].pos_depth.w == 0.0f) {
    out[pix] = c4;
    return;
  }
  const v3 cur = f4v(c4);
  float min_l = 3.402823466e38f, max_l = 0.0f;
  v3 min_c = cur, max_c = mk(0.0f, 0.0f, 0.0f);
  const float cl = luma(cur);
  for (int dx = -1; dx <= 1; ++dx) {
    for (int dy = -1; dy <= 1; ++dy) {
      if (dx == 0 && dy == 0) continue;
      const int sx = (int)x + dx, sy = (int)y + dy;
      if (sx < 0 || sy < 0 || sx >= (int)p.width || sy >= (int)p.rows) continue;
      const uint32_t q = (uint32_t)sy * p.width + (uint32_t)sx;
      if (b.gbuf[p.cur][q].pos_depth.w == 0.0f) continue;
      const v3 nc = f4v(in[q]);
      const float nl = luma(nc);
      if (nl < min_l) {
        min_l = nl;
        min_c = nc;
      } else if (nl > max_l) {
        max_l = nl;
        max_c = nc;
      }
    }
  }
  v3 rec = cur;
  if (cl < min_l) rec = min_c;
  else if (cl > max_l) rec = max_c;
  if (min_l > max_l) rec = cur;
  out[pix] = v4(rec, 0.0f);
}

constexpr uint32_t SVGF_SKY = 0xFFFFFFFFu;  // guide word of a sky texel (an FP16 NaN depth)

// svgf_temporal.wgsl:87-139.
__global__ __launch_bounds__(BLOCK) void k_svgf_temporal(RestirParams p, RestirBuffers b, const float4* in) {
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  const float4 rad4 = in[pix];
  const GTexel cg = b.gbuf[p.cur][pix];
  if (cg.pos_depth.w == 0.0f) {
    b.svgf_col[0][pix] = v4(f4v(rad4), 0.0f);  // variance 0
    b.svgf_guide[pix] = make_uint2(SVGF_SKY, 0u);
    b.svgf_mom[p.cur][pix] = make_float2(0.0f, 0.0f);
    return;
  }
  b.svgf_guide[pix] = make_uint2(cg.packed.x, cg.packed.y);
  const Tex cur = unpack_texel(cg);
  const v3 rad = f4v(rad4);
  const float l = luma(rad);
  v3 tc = mk(0.0f, 0.0f, 0.0f);
  float m1 = 0.0f, m2 = 0.0f;
  uint32_t count = 0;
  float px, py;
  if (reproject(p, cur.pos, px, py)) {
    const Tex prev = sample_prev(b.gbuf[p.cur ^ 1], p, px, py);
    if (texel_similar(cur, prev)) {
      const uint32_t ix = (uint32_t)floorf(px), iy = (uint32_t)floorf(py);
      const float fx = px - floorf(px), fy = py - floorf(py);
      const uint32_t x1 = min(ix + 1, p.width - 1), y1 = min(iy + 1, p.rows - 1);
      const uint32_t i00 = iy * p.width + ix, i10 = iy * p.width + x1, i01 = y1 * p.width + ix, i11 = y1 * p.width + x1;
      const float4* h = b.svgf_hist;
      tc = lerp3(lerp3(f4v(h[i00]), f4v(h[i10]), fx), lerp3(f4v(h[i01]), f4v(h[i11]), fx), fy);
      const float2* hm = b.svgf_mom[p.cur ^ 1];
      m1 = lerpf(lerpf(hm[i00].x, hm[i10].x, fx), lerpf(hm[i01].x, hm[i11].x, fx), fy);
      m2 = lerpf(lerpf(hm[i00].y, hm[i10].y, fx), lerpf(hm[i01].y, hm[i11].y, fx), fy);
      count = min(b.svgf_count[pix] + 1, p.svgf_history);  // this pixel's count (:126)
    }
  }
  const float alpha = 1.0f / (1.0f + (float)count);
  m1 = lerpf(m1, l, alpha);
  m2 = lerpf(m2, l * l, alpha);
  tc = lerp3(tc, rad, alpha);
  // the variance rides in w for the a-trous passes (0 when there are none:
  // the image goes on as it is)
  b.svgf_col[0][pix] = v4(tc, p.atrous_passes ? fmaxf(m2 - m1 * m1, 0.0f) : 0.0f);
  b.svgf_mom[p.cur][pix] = make_float2(m1, m2);
  b.svgf_count[pix] = count;
}

// svgf_atrous.wgsl:48-117: pass `pass` reads ping-pong slot pass & 1.  Sky
// pixels pass through (the reference leaves them unwritten).  The edge stops
// read the 8-byte guide k_svgf_temporal wrote (the G-buffer's FP16 depth and
// normal, the values unpack_texel decodes) and the variance rides in the
// colour's w, so a tap reads 24 bytes; the last pass writes w = 0.
__global__ __launch_bounds__(BLOCK) void k_svgf_atrous(RestirParams p, RestirBuffers b, uint32_t pass) {
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  const uint32_t src = pass & 1u, dst = src ^ 1u;
  const bool last = pass + 1u == p.atrous_passes;
  const float4 c4 = b.svgf_col[src][pix];
  const uint2 cg = b.svgf_guide[pix];
  if (cg.x == SVGF_SKY) {
    b.svgf_col[dst][pix] = v4(f4v(c4), 0.0f);
    if (pass == 0) b.svgf_hist[pix] = v4(f4v(c4), 0.0f);
    return;
  }
  const float cdepth = f16lo(cg.x);
  const v3 cn = mk(f16hi(cg.x), f16lo(cg.y), f16hi(cg.y));
  const float KW[3] = {3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};  // svgf_atrous.wgsl:11-15
  const int step = 1 << pass;
  v3 sum = f4v(c4);
  float vsum = c4.w;
  float wsum = 1.0f;
  const float cl = luma(sum);
  const float inv_l = 1.0f / fmaxf(4.0f * sqrtf(c4.w), 1e-8f);  // PHI_LUMA
  for (int dy = -2; dy <= 2; ++dy) {
    for (int dx = -2; dx <= 2; ++dx) {
      if (dx == 0 && dy == 0) continue;
      const int sx = (int)x + dx * step, sy = (int)y + dy * step;
      if (sx < 0 || sy < 0 || sx >= (int)p.width || sy >= (int)p.rows) continue;
      const uint32_t q = (uint32_t)sy * p.width + (uint32_t)sx;
      const uint2 g = b.svgf_guide[q];
      if (g.x == SVGF_SKY) continue;
      const float4 cq4 = b.svgf_col[src][q];
      const v3 cq = f4v(cq4);
      const v3 nq = mk(f16hi(g.x), f16lo(g.y), f16hi(g.y));
      // PHI_NORMAL: max(n . nq, 0)^128 by seven squarings
      float wn = fmaxf(dot(cn, nq), 0.0f);
      for (int k = 0; k < 7; ++k) wn *= wn;
      const float dd = fabsf(cdepth - f16lo(g.x));
      // PHI_DEPTH and PHI_LUMA in one exponential
      const float e = -(dd * dd) * (1.0f / (2.0f * (0.1f * 0.1f))) - fabsf(cl - luma(cq)) * inv_l;
      float w = wn * expf(e);
      w *= KW[dx < 0 ? -dx : dx] * KW[dy < 0 ? -dy : dy];
      sum += cq * w;
      vsum += cq4.w * (w * w);
      wsum += w;
    }
  }
  sum = sum / wsum;
  vsum = vsum / (wsum * wsum);
  b.svgf_col[dst][pix] = v4(sum, last ? 0.0f : vsum);
  if (pass == 0) b.svgf_hist[pix] = v4(sum, 0.0f);
}

// taa.wgsl: Mitchell-weighted neighbourhood, history clipped to the
// neighbourhood's mean +- deviation (out of place).
__device__ __forceinline__ float mitchell_1d(float x0, float B, float C) {
  const float x = fabsf(2.0f * x0);
  const float k = 1.0f / 6.0f;
  if (x > 1.0f)
    return ((-B - 6.0f * C) * x * x * x + (6.0f * B + 30.0f * C) * x * x + (-12.0f * B - 48.0f * C) * x +
            (8.0f * B + 24.0f * C)) * k;
  return ((12.0f - 9.0f * B - 6.0f * C) * x * x * x + (-18.0f + 12.0f * B + 6.0f * C) * x * x + (6.0f - 2.0f * B)) * k;
}

__device__ __forceinline__ v3 clip_aabb(v3 lo, v3 hi, v3 h) {
  const v3 c = (hi + lo) * 0.5f;
  const v3 e = (hi - lo) * 0.5f;
  const v3 r = h - c;
  const v3 ru = mk(fabsf(r.x / e.x), fabsf(r.y / e.y), fabsf(r.z / e.z));
  const float m = fmaxf(ru.x, fmaxf(ru.y, ru.z));
  return m > 1.0f ? c + r / m : h;
}

__global__ __launch_bounds__(BLOCK) void k_rs_taa(RestirParams p, RestirBuffers b, const float4* in,
                                                  const float4* prev, float4* out) {
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  const GTexel cg = b.gbuf[p.cur][pix];
  if (cg.pos_depth.w == 0.0f) {
    out[pix] = in[pix];
    return;
  }
  const v3 cur = f4v(in[pix]);
  const float m0 = mitchell_1d(0.0f, 0.33f, 0.33f);
  float wsum = m0 * m0;
  v3 rec = cur * wsum;
  v3 fm = cur, sm = cur * cur;
  float n = 1.0f;
  for (int dx = -1; dx <= 1; ++dx) {
    for (int dy = -1; dy <= 1; ++dy) {
      if (dx == 0 && dy == 0) continue;
      const int sx = (int)x + dx, sy = (int)y + dy;
      if (sx < 0 || sy < 0 || sx >= (int)p.width || sy >= (int)p.rows) continue;
      const float4 q4 = in[(uint32_t)sy * p.width + (uint32_t)sx];
      const v3 sc = mk(fmaxf(q4.x, 0.0f), fmaxf(q4.y, 0.0f), fmaxf(q4.z, 0.0f));
      float w = mitchell_1d((float)dx, 0.33f, 0.33f) * mitchell_1d((float)dy, 0.33f, 0.33f);
      w *= 1.0f / (1.0f + luma(sc));
      rec += sc * w;
      wsum += w;
      fm += sc;
      sm += sc * sc;
      n += 1.0f;
    }
  }
  rec = rec / fmaxf(wsum, 1e-5f);
  v3 hist;
  uint32_t prev_id;
  float px, py;
  const Tex ct = unpack_texel(cg);
  if (reproject(p, ct.pos, px, py)) {
    const uint32_t fx0 = min((uint32_t)floorf(px), p.width - 1), fy0 = min((uint32_t)floorf(py), p.rows - 1);
    const uint32_t cx = min((uint32_t)ceilf(px), p.width - 1), cy = min((uint32_t)ceilf(py), p.rows - 1);
    prev_id = (uint32_t)floorf(py) * p.width + (uint32_t)floorf(px);
    // the reference's history00/10/01/11 index rows by the first letter
    const v3 h00 = f4v(prev[fy0 * p.width + fx0]), h10 = f4v(prev[cy * p.width + fx0]);
    const v3 h01 = f4v(prev[fy0 * p.width + cx]), h11 = f4v(prev[cy * p.width + cx]);
    const float tx = px - floorf(px), ty = py - floorf(py);
    const v3 a = h00 * (1.0f - tx) + h10 * tx;  // bilinear(tx, ty, c00, c10, c01, c11) (taa.wgsl:48-52)
    const v3 c = h01 * (1.0f - tx) + h11 * tx;
    hist = a * (1.0f - ty) + c * ty;
  } else {
    hist = f4v(prev[pix]);
    prev_id = pix;
  }
  const Tex pt = unpack_texel(b.gbuf[p.cur ^ 1][prev_id]);
  const float cdc = depth_cs(ct.depth), pdc = depth_cs(pt.depth);
  const bool valid_depth = fabsf(cdc - pdc) / cdc < 0.1f;
  const v3 mean = fm / n;
  v3 sd = mk(fabsf(sm.x - fm.x * fm.x / n), fabsf(sm.y - fm.y * fm.y / n), fabsf(sm.z - fm.z * fm.z / n));
  sd = sd / (n - 1.0f);
  sd = mk(sqrtf(sd.x), sqrtf(sd.y), sqrtf(sd.z));
  const v3 clipped = clip_aabb(mean - sd, mean + sd, hist);
  const float blend = 1.0f - 0.8f * (valid_depth ? 1.0f : 0.0f);  // history_influence 0.8 (lib.rs:537)
  const float cw = clampf(blend * (1.0f / (1.0f + luma(rec))), 0.0f, 1.0f);
  const float hw = clampf((1.0f - blend) * (1.0f / (1.0f + luma(clipped))), 0.0f, 1.0f);
  out[pix] = v4((rec * cw + clipped * hw) / (cw + hw), 0.0f);
}

// Remodulation (demodulate_radiance.wgsl with remodulate) into the frame sum;
// the final demodulated image stays in demod[cur] for the next frame's TAA.
__global__ __launch_bounds__(BLOCK) void k_rs_remod(RestirParams p, RestirBuffers b, const float4* fin,
                                                    float4* frame_sum) {
  uint32_t x, y, pix;
  if (!tile_pixel(p, x, y, pix)) return;
  const float4 c = fin[pix];
  const v3 alb = unpack_rgb9e5(b.gbuf[p.cur][pix].packed.z);
  const v3 r = f4v(c) * alb;
  const float4 s = frame_sum[pix];
  frame_sum[pix] = make_float4(s.x + r.x, s.y + r.y, s.z + r.z, s.w);
  if (fin != b.demod[p.cur]) b.demod[p.cur][pix] = c;
}

template <int W, int S, bool INST, bool TEX>
apt_status sample_t(const RestirParams& p, const SceneView& sc, const RestirBuffers& b, hipStream_t s) {
  const uint32_t tiles = ((p.width + TILE - 1) / TILE) * ((p.rows + TILE - 1) / TILE);
  const uint32_t tiles8 = ((p.width + TILE8 - 1) / TILE8) * ((p.rows + TILE8 - 1) / TILE8);
  const dim3 grid(tiles), block(BLOCK), grid8(tiles8), block8(BLOCK8);
  hipLaunchKernelGGL(k_rs_raygen, grid, block, 0, s, p, sc, b);
  for (uint32_t bounce = 0; bounce < p.max_bounces; ++bounce) {
    const bool more = bounce + 1 < p.max_bounces;
    const uint32_t npix = p.width * p.rows;
    const dim3 pool(std::min<uint32_t>(pool_blocks(), (npix + 63) / 64));
    hipLaunchKernelGGL(k_rq_reset, dim3(1), dim3(64), 0, s, b);
    hipLaunchKernelGGL((k_rs_trace<W, S, false, true, INST>), pool, block, 0, s, sc, b, npix);
    hipLaunchKernelGGL(k_rs_extend<TEX>, grid, block, 0, s, p, sc, b, bounce);
    hipLaunchKernelGGL(k_rs_di_bsdf_ray<TEX>, grid, block, 0, s, p, sc, b);  // also resets the pool counters
    hipLaunchKernelGGL((k_rs_trace<W, S, false, false, INST>), pool, block, 0, s, sc, b, npix);
    hipLaunchKernelGGL(k_rq_reset, dim3(1), dim3(64), 0, s, b);
    hipLaunchKernelGGL(k_rs_di_ris<TEX>, grid8, block8, 0, s, p, sc, b);
    hipLaunchKernelGGL((k_rs_trace<W, S, true, false, INST>), pool, block, 0, s, sc, b, npix);
    hipLaunchKernelGGL(k_rs_di_vis, dim3((npix + BLOCK - 1) / BLOCK), block, 0, s, b, npix);
    if (more) hipLaunchKernelGGL((k_rs_gi_ris<W, S, INST, TEX>), grid8, block8, 0, s, p, sc, b, bounce);
    if (bounce == 0 && p.restir_di) {
      hipLaunchKernelGGL((k_rs_di_temporal<W, S, INST, TEX>), grid8, block8, 0, s, p, sc, b);
      if (p.neighbours > 0) {
        hipLaunchKernelGGL(k_rq_reset, dim3(1), dim3(64), 0, s, b);
        hipLaunchKernelGGL(k_rs_di_self<TEX>, grid, block, 0, s, p, sc, b);
        hipLaunchKernelGGL((k_rs_trace<W, S, true, false, INST>), pool, block, 0, s, sc, b, npix);
        hipLaunchKernelGGL((k_rs_di_spatial<W, S, INST, TEX>), grid8, block8, 0, s, p, sc, b);
      }
    }
    if (bounce == 0 && more && p.restir_gi) {
      hipLaunchKernelGGL((k_rs_gi_temporal<W, S, INST, TEX>), grid8, block8, 0, s, p, sc, b);
      for (uint32_t pass = 0; pass < p.gi_spatial_passes; ++pass) {
        const GiRes* in = pass % 2 == 0 ? b.gi : b.gi_tmp;
        GiRes* out = pass % 2 == 0 ? b.gi_tmp : b.gi;
        hipLaunchKernelGGL((k_rs_gi_spatial<W, S, INST, TEX>), grid8, block8, 0, s, p, sc, b, pass, in, out);
      }
      if (p.gi_spatial_passes % 2 == 1)
        APT_HIP(hipMemcpyAsync(b.gi, b.gi_tmp, (size_t)npix * sizeof(GiRes), hipMemcpyDeviceToDevice, s));
    }
    hipLaunchKernelGGL((k_rs_apply_di<W, S, INST, TEX>), grid8, block8, 0, s, p, sc, b);
    if (more) hipLaunchKernelGGL(k_rs_apply_gi<TEX>, grid, block, 0, s, p, sc, b);
  }
  APT_HIP(hipGetLastError());
  return APT_OK;
}

}  // namespace

// This file is compiled four times, once per variant: two-level scene or not
// (APT_RESTIR_INST), textured scene or not (APT_RESTIR_TEX); each translation
// unit defines its variant's restir_sample_variant, and the plain one also the
// dispatch and the post chain.
template <>
apt_status restir_sample_variant<APT_RESTIR_INST != 0, APT_RESTIR_TEX != 0>(const RestirParams& p,
                                                                          const SceneView& sc,
                                                                          const RestirBuffers& b, hipStream_t s) {
  constexpr bool INST = APT_RESTIR_INST != 0, TEX = APT_RESTIR_TEX != 0;
  switch (sc.width) {
    case 2: APT_FOR_STACK_DEPTH(sc.lds_stack, return (sample_t<2, S_, INST, TEX>(p, sc, b, s)));
    case 4: APT_FOR_STACK_DEPTH(sc.lds_stack, return (sample_t<4, S_, INST, TEX>(p, sc, b, s)));
    case 8: APT_FOR_STACK_DEPTH(sc.lds_stack, return (sample_t<8, S_, INST, TEX>(p, sc, b, s)));
    default: break;
  }
  set_error("restir: unsupported BVH width %u", sc.width);
  return APT_ERR_INVALID_ARG;
}

#if !APT_RESTIR_INST && !APT_RESTIR_TEX
apt_status restir_sample(const RestirParams& p, const SceneView& sc, const RestirBuffers& b, hipStream_t s) {
  if (scene_textured(sc))
    return sc.insts ? restir_sample_variant<true, true>(p, sc, b, s) : restir_sample_variant<false, true>(p, sc, b, s);
  return sc.insts ? restir_sample_variant<true, false>(p, sc, b, s) : restir_sample_variant<false, false>(p, sc, b, s);
}
#endif

#if !APT_RESTIR_INST && !APT_RESTIR_TEX
apt_status restir_post(const RestirParams& p, const RestirBuffers& b, bool firefly, bool svgf, bool taa,
                       float4* frame_sum, hipStream_t s) {
  const uint32_t tiles = ((p.width + TILE - 1) / TILE) * ((p.rows + TILE - 1) / TILE);
  const dim3 grid(tiles), block(BLOCK);
  hipLaunchKernelGGL(k_rs_demod, grid, block, 0, s, p, b);
  const float4* img = b.demod[p.cur];
  if (firefly) {
    hipLaunchKernelGGL(k_rs_firefly, grid, block, 0, s, p, b, img, b.work);
    img = b.work;
  }
  if (svgf) {
    hipLaunchKernelGGL(k_svgf_temporal, grid, block, 0, s, p, b, img);
    for (uint32_t i = 0; i < p.atrous_passes; ++i) hipLaunchKernelGGL(k_svgf_atrous, grid, block, 0, s, p, b, i);
    img = b.svgf_col[p.atrous_passes & 1u];
  }
  if (taa && !p.first_frame) {
    float4* out = img == b.work ? b.demod[p.cur] : b.work;
    hipLaunchKernelGGL(k_rs_taa, grid, block, 0, s, p, b, img, b.demod[p.cur ^ 1], out);
    img = out;
  }
  hipLaunchKernelGGL(k_rs_remod, grid, block, 0, s, p, b, img, frame_sum);
  APT_HIP(hipGetLastError());
  return APT_OK;
}

#endif  // the plain variant

}  // namespace apt
