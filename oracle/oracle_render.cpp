// ORACLE - test infrastructure only (see oracle_core.cpp).
//
// CPU path tracer, templated on the arithmetic type: float is the FP32 CPU
// restatement the GPU integrator is compared with; double is the FP64
// high-spp reference.  Structure follows the reference's CPU integrator
// PathIntegrator::li (crates/appearance-path-tracer/src/path_integrator.rs:103-307:
// NEE with a power-heuristic MIS, BSDF sampling, Russian roulette) with the
// GPU path's shading model (Disney BSDF, shared/material/disney_bsdf.wgsl and
// ggxmdf.wgsl; sun and sky, shared/sky_bindings.wgsl; light selection,
// helpers/nee.wgsl:102-115) and these documented choices, identical to the
// product's: emission on hit (one-sided), solid-angle light pdfs, jittered
// pixel positions, p = min(1, max(T)) roulette, the Disney evaluate() pdf
// weighted once and with the back-face eta.
//
// Pins (tests/test_oracle_render.py, tests/test_oracle_disney.py): closed
// furnace (L = Le / (1 - a)), a lit-plane closed form, FP32 vs FP64 agreement,
// Disney invariants (pdf normalisation, sample/eval pdf agreement, energy
// bound, Burley and Lambert special cases).
#include <atomic>
#include <thread>

#include "oracle_pt.h"

namespace ocr {

// One camera path: the same estimator, draw order and epsilons as libapt's
// wavefront integrator (raygen -> [shade: emission, NEE, BSDF, roulette]).
template <typename R>
V<R> path(const Scene& sc, const Cfg& cf, uint32_t x, uint32_t y, uint32_t sample, uint64_t* nrays) {
  const Sky& k = *sc.sky;
  const R eps = R(1e-4);
  const uint32_t pixel = y * cf.width + x;
  const uint32_t seed = cf.frame_idx * cf.spp + sample + cf.seed;
  uint32_t rng = orc::pcg_hash(pixel ^ orc::xor_shift(seed));
  R jx = draw<R>(rng), jy = draw<R>(rng);
  if (!cf.jitter) jx = jy = R(0.5);
  R ux = ((R)x + jx) / (R)cf.width * R(2) - R(1);
  R uy = -(((R)y + jy) / (R)cf.height * R(2) - R(1));
  const float* m = sc.cam->m;
  R th = std::tan(R(0.5) * (R)sc.cam->fov_deg * C<R>::pi / R(180));
  R asp = (R)cf.width / (R)cf.height;
  V<R> dl = unit(V<R>(ux * th * asp, uy * th, R(-1)));
  V<R> cr = unit(V<R>(m[0], m[1], m[2])), cu = unit(V<R>(m[4], m[5], m[6])), cb = unit(V<R>(m[8], m[9], m[10]));
  V<R> ro(m[12], m[13], m[14]);
  V<R> rd = cr * dl.x + cu * dl.y + cb * dl.z;
  if (!cf.jitter) ro = ro + rd * eps;  // the ReSTIR integrator's camera ray (safe_origin along it)
  V<R> T(1, 1, 1), L;
  R pdf_prev = 0;
  bool nee_side = false;
  const uint32_t ne = (uint32_t)sc.emissive.size();
  const R p_sun = (k.sun_enabled > 0.0f) ? (ne > 0 ? R(0.5) : R(1)) : R(0);
  const V<R> sun_l(-(R)k.sun_direction[0], -(R)k.sun_direction[1], -(R)k.sun_direction[2]);
  const R sun_ang = (R)k.sun_size * R(0.1);
  const R sun_sa = R(2) * C<R>::pi * (R(1) - std::cos(sun_ang));
  for (uint32_t bounce = 0; bounce < cf.max_bounces; ++bounce) {
    R t, bu, bv;
    uint32_t prim;
    if (nrays) nrays[0]++;
    if (!closest(sc, ro, rd, t, bu, bv, prim)) {
      V<R> Le = sky_L<R>(k, sc.at, rd);
      R si = sun_I<R>(k, rd);
      if (si > R(0)) {
        R w = R(1);
        if (pdf_prev > R(0) && nee_side) w = power_h(pdf_prev, p_sun / sun_sa);
        Le = Le + f3<R>(k.sun_color) * (si * w);
      }
      L = L + T * Le;
      break;
    }
    V<R> a, b, c;
    tri_pts(sc, prim, a, b, c);
    const Mat& mt = sc.mats[sc.mat_ids ? sc.mat_ids[prim] : 0];
    V<R> hp = ro + rd * t;
    V<R> ng = unit(crossp(b - a, c - a));
    V<R> ns = ng;
    if (sc.nrm) {
      const float* n = sc.nrm + 9 * (size_t)prim;
      ns = unit(f3<R>(n) * (R(1) - bu - bv) + f3<R>(n + 3) * bu + f3<R>(n + 6) * bv);
    }
    V<R> wo = -rd;
    // textures at the hit; normal mapping on the interpolated normal, then the
    // face flip of the mapped and clearcoat normals (inline_path_tracer.wgsl:103-165)
    R tu, tv;
    hit_uv(sc.at, prim, bu, bv, tu, tv);
    V<R> nmap, ncc;
    mapped_normals(sc.at, mt, prim, bu, bv, tu, tv, ns, nmap, ncc);
    bool back = dotp(wo, ns) < R(0);
    ns = back ? -nmap : nmap;
    if (back) ncc = -ncc;
    const MatEval<R> me = eval_material<R>(sc.at, mt, tu, tv);
    bool emits = mt.emission[0] > 0.0f || mt.emission[1] > 0.0f || mt.emission[2] > 0.0f;
    if (emits && dotp(ng, rd) < R(0)) {
      R w = R(1);
      if (pdf_prev > R(0) && nee_side && ne > 0) {
        R p_l = (R(1) - p_sun) / (R)ne / area_of(a, b, c) * (t * t) / (-dotp(ng, rd));
        w = power_h(pdf_prev, p_l);
      }
      L = L + T * me.emission * w;
    }
    const bool lambert = mt.bsdf_model == 1u;
    const V<R> albedo = me.color;
    const Disney<R> bd = from_eval<R>(mt, me);
    const Frame<R> fr = basis(ns), fc = basis(ncc);
    // NEE
    R u_sel = draw<R>(rng), u_a = draw<R>(rng), u_b = draw<R>(rng);
    if (p_sun > R(0) || ne > 0) {
      V<R> wi, Le;
      R p_l = 0, dist = 0;
      if (u_sel < p_sun || ne == 0) {  // draws can be exactly 1.0
        wi = unit(cone_dir(u_a, u_b, sun_l, sun_ang));
        Le = f3<R>(k.sun_color) * sun_I<R>(k, wi);
        p_l = p_sun / sun_sa;
        dist = R(1e6);
      } else {
        R us = (u_sel - p_sun) / (R(1) - p_sun);
        uint32_t kk = std::min((uint32_t)std::max(R(0), us * (R)ne), ne - 1);
        uint32_t lp = sc.emissive[kk];
        V<R> la, lb, lc;
        tri_pts(sc, lp, la, lb, lc);
        R su = std::sqrt(u_a);
        V<R> xl = la * (R(1) - su) + lb * (su * (R(1) - u_b)) + lc * (su * u_b);
        V<R> nl = unit(crossp(lb - la, lc - la));
        V<R> dv = xl - hp;
        dist = len(dv);
        wi = dv * (R(1) / dist);
        R cos_l = -dotp(nl, wi);
        const Mat& lm = sc.mats[sc.mat_ids ? sc.mat_ids[lp] : 0];
        R lu, lv;  // the emission texture at the light point (helpers/nee.wgsl:39-51)
        hit_uv(sc.at, lp, su * (R(1) - u_b), su * u_b, lu, lv);
        Le = eval_material<R>(sc.at, lm, lu, lv).emission;
        p_l = cos_l > R(0) ? (R(1) - p_sun) / (R)ne / area_of(la, lb, lc) * (dist * dist) / cos_l : R(0);
      }
      R cos_i = dotp(wi, ns);
      if (p_l > R(0) && cos_i > R(0) && mx(Le) > R(0)) {
        R p_b = 0;
        V<R> f;
        if (lambert) {
          R co = dotp(ns, wo);
          if (co * cos_i > R(0)) {
            f = albedo * (R(1) / C<R>::pi);
            p_b = std::fabs(cos_i) / C<R>::pi;
          }
        } else {
          f = disney_f(bd, fr, fc, wo, wi, back, p_b);
        }
        R w = power_h(p_l, p_b);
        V<R> cc = T * f * Le * (cos_i * w / p_l);
        if (mx(cc) > R(0) && cc.x == cc.x && cc.y == cc.y && cc.z == cc.z) {
          if (nrays) nrays[1]++;
          if (!occluded(sc, hp + ns * eps, wi, dist - R(2) * eps)) L = L + cc;
        }
      }
    }
    if (bounce + 1 >= cf.max_bounces) break;
    R r0 = draw<R>(rng), r1 = draw<R>(rng), r2 = draw<R>(rng);
    Smp<R> bs;
    if (lambert) {
      V<R> l = cos_hemi(r0, r1);
      if (dotp(fr.n, wo) < R(0)) l.z = -l.z;
      bs.wi = fr.world(l);
      bs.pdf = std::fabs(l.z) / C<R>::pi;
      bs.f = albedo * (R(1) / C<R>::pi);
      if (!(bs.pdf > R(0))) bs.pdf = R(0);
    } else {
      bs = disney_sample(bd, fr, fc, wo, t, back, r0, r1, r2);
    }
    if (!(bs.pdf > R(0))) break;
    V<R> Tn = T * bs.f * (std::fabs(dotp(bs.wi, ns)) / bs.pdf);
    if (!(Tn.x == Tn.x && Tn.y == Tn.y && Tn.z == Tn.z && mx(Tn) > R(0))) break;
    if (bounce + 1 >= cf.rr_start) {
      R q = std::min(R(1), mx(Tn));
      R u = draw<R>(rng);
      if (u > q) break;
      Tn = Tn * (R(1) / q);
    }
    R side = dotp(bs.wi, ns) > R(0) ? R(1) : R(-1);
    ro = hp + ns * (side * eps);
    rd = bs.wi;
    pdf_prev = bs.pdf;
    nee_side = side > R(0);
    T = Tn;
  }
  return L;
}

template <typename R>
void render(const Scene& sc, const Cfg& cf, const uint32_t* pixels, uint32_t n_pixels, float* out,
            int threads, uint64_t* counts, size_t n_counts) {
  // Work items are 64x64 tiles of the slab (the reference CPU tracer's block
  // loop, crates/appearance-path-tracer/src/lib.rs:107-190), or chunks of an
  // explicit pixel list.
  const bool list = pixels != nullptr;
  const uint32_t tiles_x = (cf.width + 63) / 64, tiles_y = (cf.rows + 63) / 64;
  const uint32_t items = list ? (n_pixels + 255) / 256 : tiles_x * tiles_y;
  std::atomic<uint32_t> next{0};
  // per thread: extension rays, shadow rays, then the box and triangle tests
  // of the closest-hit and of the any-hit queries (FP32 only)
  constexpr size_t NC = 6;
  std::vector<uint64_t> cnt((size_t)threads * NC, 0);
  auto shade_px = [&](uint32_t x, uint32_t yl, float* o, uint64_t* nr) {
    V<R> acc;
    for (uint32_t s = 0; s < cf.spp; ++s) acc = acc + path<R>(sc, cf, x, cf.row_start + yl, s, nr);
    R inv = R(1) / (R)cf.spp;
    o[0] = (float)(acc.x * inv);
    o[1] = (float)(acc.y * inv);
    o[2] = (float)(acc.z * inv);
  };
  auto work = [&](int tid) {
    uint64_t* nr = &cnt[NC * (size_t)tid];
    orc::oracle_set_tally(nr + 2);
    while (true) {
      uint32_t it = next.fetch_add(1);
      if (it >= items) break;
      if (list) {
        for (uint32_t i = it * 256; i < std::min(n_pixels, it * 256 + 256); ++i) {
          uint32_t p = pixels[i];
          shade_px(p % cf.width, p / cf.width, out + 3 * (size_t)i, nr);
        }
      } else {
        uint32_t tx = it % tiles_x, ty = it / tiles_x;
        for (uint32_t yl = ty * 64; yl < std::min(cf.rows, ty * 64 + 64); ++yl)
          for (uint32_t x = tx * 64; x < std::min(cf.width, tx * 64 + 64); ++x)
            shade_px(x, yl, out + 3 * ((size_t)yl * cf.width + x), nr);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  orc::oracle_set_tally(nullptr);
  if (counts) {
    for (size_t k = 0; k < n_counts; ++k) {
      counts[k] = 0;
      for (int t = 0; t < threads; ++t) counts[k] += cnt[NC * (size_t)t + k];
    }
  }
}

}  // namespace ocr

extern "C" {

// Renders rows [row_start, row_start+rows) of a width x height frame, spp
// samples per pixel, and writes the linear mean radiance (3 floats per
// pixel, row order) to out; cfg10 ends with the jitter switch (0: pixel
// centres).  If pixels != NULL only those slab-local pixel
// indices are rendered, in list order.  precision: 32 or 64.
// counts[0..1] receive extension and shadow rays traced.  attribs: an
// ocr::Attribs (uvs, tangent frames, textures) or NULL.
int or_render(const float* pos, const float* nrm, const uint32_t* mat_ids, const void* mats,
              uint32_t n_tris, void* bvh, const void* sky, const void* cam, const uint32_t* cfg10,
              const uint32_t* pixels, uint32_t n_pixels, int precision, int threads, float* out,
              uint64_t* counts, const void* attribs) {
  return or_render_counted(pos, nrm, mat_ids, mats, n_tris, bvh, sky, cam, cfg10, pixels, n_pixels, precision,
                           threads, out, counts, 2, attribs);
}

// or_render with up to six counters: extension rays, shadow rays, box tests
// and triangle tests of the extension (closest-hit) rays, box tests and
// triangle tests of the shadow (any-hit) rays; the tests are counted in FP32
// only.  The per-sample figures of the algorithmic-bytes roofline.
int or_render_counted(const float* pos, const float* nrm, const uint32_t* mat_ids, const void* mats,
                      uint32_t n_tris, void* bvh, const void* sky, const void* cam, const uint32_t* cfg10,
                      const uint32_t* pixels, uint32_t n_pixels, int precision, int threads, float* out,
                      uint64_t* counts, uint32_t n_counts, const void* attribs) {
  if (n_counts > 6) return -1;
  ocr::Scene sc;
  sc.at = static_cast<const ocr::Attribs*>(attribs);
  sc.pos = pos;
  sc.nrm = nrm;
  sc.mat_ids = mat_ids;
  sc.mats = static_cast<const ocr::Mat*>(mats);
  sc.n_tris = n_tris;
  sc.sky = static_cast<const ocr::Sky*>(sky);
  sc.cam = static_cast<const ocr::Cam*>(cam);
  sc.bvh = static_cast<const orc::Bvh*>(bvh);
  for (uint32_t t = 0; t < n_tris; ++t) {
    const ocr::Mat& m = sc.mats[mat_ids ? mat_ids[t] : 0];
    if (m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f) {
      const float* p = pos + 9 * (size_t)t;
      float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]}, e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
      float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      if (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > 0.0f) sc.emissive.push_back(t);
    }
  }
  ocr::Cfg cf{cfg10[0], cfg10[1], cfg10[2], cfg10[3], cfg10[4], cfg10[5], cfg10[6], cfg10[7], cfg10[8], cfg10[9]};
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (precision == 64) ocr::render<double>(sc, cf, pixels, n_pixels, out, threads, counts, n_counts);
  else ocr::render<float>(sc, cf, pixels, n_pixels, out, threads, counts, n_counts);
  return 0;
}

// Disney BSDF probes for the invariance tests (FP64).  frame = normal only;
// ncc (NULL: n) is the clearcoat lobe's normal.
// eval: returns pdf, writes f.  sample: writes wi, f, returns pdf.
double or_disney_eval(const void* mat, const double* n, const double* wo, const double* wi, int back,
                      double* f, const double* ncc) {
  ocr::Disney<double> b = ocr::from_mat<double>(*static_cast<const ocr::Mat*>(mat));
  ocr::Frame<double> fr = ocr::basis(ocr::V<double>(n[0], n[1], n[2]));
  ocr::Frame<double> fc = ncc ? ocr::basis(ocr::V<double>(ncc[0], ncc[1], ncc[2])) : fr;
  double pdf;
  ocr::V<double> v = ocr::disney_f(b, fr, fc, ocr::V<double>(wo[0], wo[1], wo[2]),
                                   ocr::V<double>(wi[0], wi[1], wi[2]), back != 0, pdf);
  f[0] = v.x;
  f[1] = v.y;
  f[2] = v.z;
  return pdf;
}

double or_disney_sample(const void* mat, const double* n, const double* wo, int back, double r0, double r1,
                        double r2, double* wi, double* f, const double* ncc) {
  ocr::Disney<double> b = ocr::from_mat<double>(*static_cast<const ocr::Mat*>(mat));
  ocr::Frame<double> fr = ocr::basis(ocr::V<double>(n[0], n[1], n[2]));
  ocr::Frame<double> fc = ncc ? ocr::basis(ocr::V<double>(ncc[0], ncc[1], ncc[2])) : fr;
  ocr::Smp<double> s = ocr::disney_sample(b, fr, fc, ocr::V<double>(wo[0], wo[1], wo[2]), 0.0, back != 0, r0, r1,
                                          r2);
  wi[0] = s.wi.x;
  wi[1] = s.wi.y;
  wi[2] = s.wi.z;
  f[0] = s.f.x;
  f[1] = s.f.y;
  f[2] = s.f.z;
  return s.pdf;
}

// Texture probe: textureSampleLevel of one texture (apt_texture layout) at
// (u, v) in FP32; out4 = RGBA.
void or_tex_sample(const void* tex, float u, float v, float* out4) {
  ocr::tex_sample<float>(*static_cast<const ocr::Texture*>(tex), u, v, out4);
}

// Shading-point probe (FP32): triangle prim of a scene with attributes
// `attribs` (an ocr::Attribs), hit at barycentrics (bu, bv) of vertices 1 and
// 2, interpolated normal n, material mat.  out (26 floats): uv, mapped normal,
// clearcoat normal, then the evaluated material: colour, luminance, metallic,
// roughness, emission, transmission, sheen, sheen tint, clearcoat, clearcoat
// roughness.
void or_shading_probe(const void* mat, const void* attribs, uint32_t prim, float bu, float bv, const float* n,
                      float* out) {
  const ocr::Mat& m = *static_cast<const ocr::Mat*>(mat);
  const ocr::Attribs* at = static_cast<const ocr::Attribs*>(attribs);
  float u, v;
  ocr::hit_uv<float>(at, prim, bu, bv, u, v);
  ocr::V<float> nmap, ncc;
  ocr::mapped_normals<float>(at, m, prim, bu, bv, u, v, ocr::V<float>(n[0], n[1], n[2]), nmap, ncc);
  const ocr::MatEval<float> e = ocr::eval_material<float>(at, m, u, v);
  const float vals[26] = {u, v, nmap.x, nmap.y, nmap.z, ncc.x, ncc.y, ncc.z, e.color.x, e.color.y, e.color.z,
                          e.luminance, e.metallic, e.roughness, e.emission.x, e.emission.y, e.emission.z,
                          e.transmission, e.sheen, e.sheen_tint.x, e.sheen_tint.y, e.sheen_tint.z, e.clearcoat,
                          e.clearcoat_roughness, 0.0f, 0.0f};
  for (int k = 0; k < 26; ++k) out[k] = vals[k];
}

// Sky probe: the sky (without the sun) seen along d, FP32; attribs may carry
// a panorama.  out: RGB.
void or_sky_probe(const void* sky, const void* attribs, const float* d, float* out) {
  const ocr::V<float> v = ocr::sky_L<float>(*static_cast<const ocr::Sky*>(sky), static_cast<const ocr::Attribs*>(attribs),
                                            ocr::V<float>(d[0], d[1], d[2]));
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
}

}  // extern "C"
