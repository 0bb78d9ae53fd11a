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
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"

namespace orc {
uint32_t pcg_hash(uint32_t x);
uint32_t xor_shift(uint32_t x);
struct Bvh;
}

namespace ocr {

// ---- POD mirrors of the ABI structs (160 B material, 80 B sky, 80 B camera) ----
struct Mat {
  float color[3]; uint32_t color_texture;
  float metallic, roughness; uint32_t mr_tex; float normal_scale;
  float emission[3]; uint32_t normal_texture;
  uint32_t emission_texture; float transmission, eta, subsurface;
  float absorption[3]; float specular;
  float specular_tint[3]; float anisotropic;
  float sheen; uint32_t sheen_texture; float clearcoat; uint32_t clearcoat_texture;
  float clearcoat_roughness; uint32_t cr_tex; float alpha_cutoff; uint32_t sheen_tint_texture;
  uint32_t clearcoat_normal_texture; uint32_t bsdf_model; uint32_t pad1, pad2;
  float sheen_tint[3]; uint32_t transmission_texture;
};
static_assert(sizeof(Mat) == 160, "material layout");

struct Sky {
  float sun_direction[3]; float sun_size;
  float sun_color[3]; float sun_intensity;
  float zenith[3]; float sun_enabled;
  float horizon[3]; float pad0;
  float ground[3]; float pad1;
};
static_assert(sizeof(Sky) == 80, "sky layout");

struct Cam {
  float m[16];
  float fov_deg, near_plane, far_plane;
  uint32_t pad;
};
static_assert(sizeof(Cam) == 80, "camera layout");

template <typename R>
struct V {
  R x, y, z;
  V() : x(0), y(0), z(0) {}
  V(R a, R b, R c) : x(a), y(b), z(c) {}
  V operator+(const V& o) const { return V(x + o.x, y + o.y, z + o.z); }
  V operator-(const V& o) const { return V(x - o.x, y - o.y, z - o.z); }
  V operator-() const { return V(-x, -y, -z); }
  V operator*(R s) const { return V(x * s, y * s, z * s); }
  V operator*(const V& o) const { return V(x * o.x, y * o.y, z * o.z); }
  R operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
};
template <typename R> R dotp(const V<R>& a, const V<R>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> V<R> crossp(const V<R>& a, const V<R>& b) {
  return V<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename R> R len(const V<R>& a) { return std::sqrt(dotp(a, a)); }
template <typename R> V<R> unit(const V<R>& a) { return a * (R(1) / std::sqrt(dotp(a, a))); }
template <typename R> R mx(const V<R>& a) { return std::max(a.x, std::max(a.y, a.z)); }
template <typename R> V<R> f3(const float* p) { return V<R>((R)p[0], (R)p[1], (R)p[2]); }
template <typename R> R lerp(R a, R b, R t) { return a * (R(1) - t) + b * t; }
template <typename R> R cl(R x, R lo, R hi) { return std::min(std::max(x, lo), hi); }
template <typename R> R sq(R x) { return x * x; }

template <typename R> struct C {
  static constexpr R pi = (R)3.14159265358979323846;
};

// ---- RNG (random.wgsl) ------------------------------------------------------
template <typename R>
R draw(uint32_t& st) {
  st = orc::pcg_hash(st);
  if constexpr (sizeof(R) == 4) return (float)st / 4294967296.0f;
  else return (double)st / 4294967296.0;
}

// ---- frames -----------------------------------------------------------------
template <typename R>
struct Frame {
  V<R> t1, t2, n;
  V<R> world(const V<R>& l) const { return t1 * l.x + t2 * l.y + n * l.z; }
  V<R> local(const V<R>& w) const { return V<R>(dotp(t1, w), dotp(t2, w), dotp(n, w)); }
};

template <typename R>
Frame<R> basis(const V<R>& n) {  // math.wgsl:71-86
  Frame<R> f;
  if (n.z < R(0)) {
    R a = R(1) / (R(1) - n.z), b = n.x * n.y * a;
    f.t1 = V<R>(R(1) - n.x * n.x * a, -b, n.x);
    f.t2 = V<R>(b, n.y * n.y * a - R(1), -n.y);
  } else {
    R a = R(1) / (R(1) + n.z), b = -n.x * n.y * a;
    f.t1 = V<R>(R(1) - n.x * n.x * a, b, -n.x);
    f.t2 = V<R>(b, R(1) - n.y * n.y * a, -n.y);
  }
  f.n = n;
  return f;
}

template <typename R>
V<R> perp(const V<R>& u) {  // math.wgsl:89-99
  R ax = std::fabs(u.x), ay = std::fabs(u.y), az = std::fabs(u.z);
  uint32_t xm = ((ax - ay) < 0 && (ax - az) < 0) ? 1u : 0u;
  uint32_t ym = (ay - az) < 0 ? (1u ^ xm) : 0u;
  uint32_t zm = 1u ^ (xm | ym);
  return crossp(u, V<R>((R)xm, (R)ym, (R)zm));
}

template <typename R>
V<R> cone_dir(R u0, R u1, const V<R>& dir, R angle) {  // sampling.wgsl:52-61
  R h = std::cos(angle), phi = R(2) * C<R>::pi * u0;
  R z = h + (R(1) - h) * u1;
  R s = std::sqrt(std::max(R(0), R(1) - z * z));
  V<R> b = perp(dir), t = crossp(b, dir);
  return b * (std::cos(phi) * s) + t * (std::sin(phi) * s) + dir * z;
}

template <typename R>
V<R> cos_hemi(R u0, R u1) {  // CPU tracer sampling/mod.rs:31-53 (concentric disk)
  R ox = R(2) * u0 - R(1), oy = R(2) * u1 - R(1), dx = 0, dy = 0;
  if (ox != R(0) || oy != R(0)) {
    R r, th;
    if (std::fabs(ox) > std::fabs(oy)) {
      r = ox;
      th = (C<R>::pi * R(0.25)) * (oy / ox);
    } else {
      r = oy;
      th = (C<R>::pi * R(0.5)) - (C<R>::pi * R(0.25)) * (ox / oy);
    }
    dx = r * std::cos(th);
    dy = r * std::sin(th);
  }
  return V<R>(dx, dy, std::sqrt(std::max(R(0), R(1) - dx * dx - dy * dy)));
}

template <typename R> R power_h(R a, R b) {
  R a2 = a * a, b2 = b * b;
  return a2 + b2 > R(0) ? a2 / (a2 + b2) : R(0);
}

// ---- microfacet distributions (ggxmdf.wgsl) ---------------------------------
template <typename R>
void alpha_of(R rough, R aniso, R& ax, R& ay) {
  R s = rough * rough;
  R asp = std::sqrt(R(1) + aniso * (aniso < R(0) ? R(0.9) : R(-0.9)));
  ax = std::max(R(0.001), s / asp);
  ay = std::max(R(0.001), s * asp);
}

template <typename R>
R D_ggx(R ax, R ay, const V<R>& m) {
  if (m.z == R(0)) return ax * ax / C<R>::pi;
  R c2 = m.z * m.z, s = std::sqrt(std::max(R(0), R(1) - c2)), t2 = (R(1) - c2) / c2, sr;
  if (ax == ay || s == R(0)) sr = R(1) / (ax * ax);
  else sr = sq(m.x / (s * ax)) + sq(m.y / (s * ay));
  return R(1) / (C<R>::pi * ax * ay * sq(c2) * sq(R(1) + t2 * sr));
}

template <typename R>
R Lambda_ggx(R ax, R ay, const V<R>& v) {
  if (v.z == R(0)) return R(0);
  R c2 = v.z * v.z, s = std::sqrt(std::max(R(0), R(1) - c2)), pr;
  if (ax == ay || s == R(0)) pr = ax;
  else pr = std::sqrt(sq((v.x * ax) / s) + sq((v.y * ay) / s));
  R t2 = (s * s) / c2;
  return (R(-1) + std::sqrt(R(1) + pr * pr * t2)) * R(0.5);
}

template <typename R>
R G_ggx(R ax, R ay, const V<R>& wi, const V<R>& wo) {
  return R(1) / (R(1) + Lambda_ggx(ax, ay, wo) + Lambda_ggx(ax, ay, wi));
}

template <typename R>
V<R> sample_ggx(R ax, R ay, const V<R>& v, R r0, R r1) {
  R sg = v.z < R(0) ? R(-1) : R(1);
  V<R> st = unit(V<R>(sg * v.x * ax, sg * v.y * ay, sg * v.z));
  V<R> t1 = v.z < R(0.9999) ? unit(crossp(st, V<R>(0, 0, 1))) : V<R>(1, 0, 0);
  V<R> t2 = crossp(t1, st);
  R a = R(1) / (R(1) + st.z), r = std::sqrt(r0);
  R phi = r1 < a ? r1 / a * C<R>::pi : C<R>::pi + (r1 - a) / (R(1) - a) * C<R>::pi;
  R p1 = r * std::cos(phi), p2 = r * std::sin(phi);
  if (r1 >= a) p2 *= st.z;
  V<R> h = t1 * p1 + t2 * p2 + st * std::sqrt(std::max(R(0), R(1) - p1 * p1 - p2 * p2));
  return unit(V<R>(h.x * ax, h.y * ay, std::max(R(0), h.z)));
}

template <typename R>
R pdf_ggx(R ax, R ay, const V<R>& v, const V<R>& m) {
  if (v.z == R(0)) return R(0);
  R g1 = R(1) / (R(1) + Lambda_ggx(ax, ay, v));
  return g1 * std::fabs(dotp(v, m)) * D_ggx(ax, ay, m) / std::fabs(v.z);
}

template <typename R>
R D_gtr1(R a0, const V<R>& m) {
  R al = cl(a0, R(0.001), R(0.999)), a2 = al * al;
  return (a2 - R(1)) / (C<R>::pi * std::log(a2)) * (R(1) / (R(1) + (a2 - R(1)) * m.z * m.z));
}

template <typename R>
R Lambda_gtr1(R a0, const V<R>& v) {
  if (v.z == R(0)) return R(0);
  R c2 = v.z * v.z, s = std::sqrt(std::max(R(0), R(1) - c2));
  if (s == R(0)) return R(0);
  R cot2 = c2 / (s * s), cot = std::sqrt(cot2), al2 = sq(cl(a0, R(0.001), R(0.999)));
  R a = std::sqrt(cot2 + al2), b = std::sqrt(cot2 + R(1));
  return (a - b + cot * (std::log(cot + b) - std::log(cot + a))) / (cot * std::log(al2));
}

template <typename R>
R G_gtr1(R a0, const V<R>& wi, const V<R>& wo) {
  return R(1) / (R(1) + Lambda_gtr1(a0, wo) + Lambda_gtr1(a0, wi));
}

template <typename R>
V<R> sample_gtr1(R a0, R r0, R r1) {
  R al = cl(a0, R(0.001), R(0.999)), a2 = al * al;
  R c2 = (R(1) - std::pow(a2, R(1) - r0)) / (R(1) - a2);
  R s = std::sqrt(std::max(R(0), R(1) - c2)), phi = R(2) * C<R>::pi * r1;
  return V<R>(std::cos(phi) * s, std::sin(phi) * s, std::sqrt(c2));
}

// ---- Disney BSDF (disney_bsdf.wgsl) -------------------------------------------
template <typename R>
struct Disney {
  V<R> color, transmittance, spec_tint, sheen_tint;
  R metallic, subsurface, luminance, specular, roughness, anisotropic, sheen, clearcoat,
      clearcoat_gloss, transmission, eta;
};

template <typename R>
Disney<R> from_mat(const Mat& m) {
  Disney<R> d;
  d.color = f3<R>(m.color);
  d.metallic = m.metallic;
  d.transmittance = f3<R>(m.absorption);
  d.subsurface = m.subsurface;
  d.luminance = R(1);
  d.specular = std::max((R)m.specular, R(1) - (R)m.roughness);
  d.roughness = m.roughness;
  d.spec_tint = f3<R>(m.specular_tint);
  d.anisotropic = m.anisotropic;
  d.sheen = m.sheen;
  d.sheen_tint = f3<R>(m.sheen_tint);
  d.clearcoat = m.clearcoat;
  d.clearcoat_gloss = R(1) - (R)m.clearcoat_roughness;
  d.transmission = m.transmission;
  d.eta = m.eta;
  return d;
}

template <typename R> R schlick(R u) {
  R m = cl(R(1) - u, R(0), R(1));
  return sq(sq(m)) * m;
}

template <typename R>
R fres(R ci, R eta, R& ct) {
  R st2 = (R(1) - ci * ci) * (eta * eta);
  if (st2 > R(1)) {
    ct = R(0);
    return R(1);
  }
  ct = std::min(std::sqrt(std::max(R(1) - st2, R(0))), R(1));
  R a = std::fabs(ci);
  if (a == R(0) && ct == R(0)) return R(1);
  R k0 = eta * ct, k1 = eta * a;
  return R(0.5) * (sq((a - k0) / (a + k0)) + sq((ct - k1) / (ct + k1)));
}

template <typename R>
V<R> spec_F(const Disney<R>& b, const V<R>& o, const V<R>& h) {
  V<R> v = b.spec_tint * (b.specular * R(0.08));
  v = v * (R(1) - b.metallic) + b.color * b.metallic;
  R f = schlick(std::fabs(dotp(o, h)));
  return v * (R(1) - f) + V<R>(f, f, f);
}

template <typename R>
V<R> cc_F(const Disney<R>& b, const V<R>& o, const V<R>& h) {
  R f = lerp(R(0.04), R(1), schlick(std::fabs(dotp(o, h)))) * R(0.25) * b.clearcoat;
  return V<R>(f, f, f);
}

template <typename R> R cc_alpha(const Disney<R>& b) { return lerp(R(0.1), R(0.001), b.clearcoat_gloss); }

template <typename R>
R diffuse_lobe(const Disney<R>& b, const V<R>& n, const V<R>& wo, const V<R>& wi, const V<R>& m, V<R>& val) {
  R con = dotp(n, wo), cin = dotp(n, wi), cih = dotp(wi, m);
  R fl = schlick(cin), fv = schlick(con), fd = 0;
  if (b.subsurface != R(1)) {
    R f90 = R(0.5) + R(2) * cih * cih * b.roughness;
    fd = lerp(R(1), f90, fl) * lerp(R(1), f90, fv);
  }
  if (b.subsurface > R(0)) {
    R f90 = cih * cih * b.roughness;
    R fss = lerp(R(1), f90, fl) * lerp(R(1), f90, fv);
    R ss = R(1.25) * (fss * (R(1) / (std::fabs(con) + std::fabs(cin)) - R(0.5)) + R(0.5));
    fd = lerp(fd, ss, b.subsurface);
  }
  val = b.color * (fd / C<R>::pi * (R(1) - b.metallic));
  return std::fabs(cin) / C<R>::pi;
}

template <typename R>
R sheen_lobe(const Disney<R>& b, const V<R>& wi, const V<R>& m, V<R>& val) {
  val = b.sheen_tint * (schlick(dotp(wi, m)) * b.sheen * (R(1) - b.metallic));
  return R(1) / (R(2) * C<R>::pi);
}

template <typename R>
R mf_lobe(const Disney<R>& b, R ax, R ay, const V<R>& wo, const V<R>& wi, const V<R>& m, bool cc, V<R>& val) {
  val = V<R>();
  if (wo.z == R(0) || wi.z == R(0)) return R(0);
  R coh = dotp(wo, m);
  if (coh == R(0)) return R(0);
  R den = std::fabs(R(4) * wo.z * wi.z);
  if (cc) {
    val = cc_F(b, wo, m) * (D_gtr1(ax, m) * G_gtr1(ax, wi, wo) / den);
    return D_gtr1(ax, m) * std::fabs(m.z) / std::fabs(R(4) * coh);
  }
  val = spec_F(b, wo, m) * (D_ggx(ax, ay, m) * G_ggx(ax, ay, wi, wo) / den);
  return pdf_ggx(ax, ay, wo, m) / std::fabs(R(4) * coh);
}

template <typename R>
bool weights(const Disney<R>& b, R w[4]) {
  w[0] = lerp(b.luminance, R(0), b.metallic);
  w[1] = lerp(b.sheen, R(0), b.metallic);
  w[2] = lerp(b.specular, R(1), b.metallic);
  w[3] = b.clearcoat * R(0.25);
  R s = w[0] + w[1] + w[2] + w[3];
  if (s == R(0)) return false;
  for (int k = 0; k < 4; ++k) w[k] *= R(1) / s;
  return true;
}

template <typename R>
V<R> refl_eval(const V<R>& col, const V<R>& wo, const V<R>& wi, const V<R>& m, R ax, R ay, R f) {
  R den = std::fabs(R(4) * wo.z * wi.z);
  if (den == R(0)) return V<R>();
  return col * (f * D_ggx(ax, ay, m) * G_ggx(ax, ay, wi, wo) / den);
}

template <typename R>
V<R> refr_eval(R eta, const V<R>& col, const V<R>& wo, const V<R>& wi, const V<R>& m, R ax, R ay, R t) {
  if (wo.z == R(0) || wi.z == R(0)) return V<R>();
  R cih = dotp(m, wi), coh = dotp(m, wo);
  R dots = (cih * coh) / (wi.z * wo.z), sd = coh + eta * cih;
  if (std::fabs(sd) < R(1e-6)) return V<R>();
  return col * (std::fabs(dots) * t * D_ggx(ax, ay, m) * G_ggx(ax, ay, wi, wo) / (sd * sd) * (eta * eta));
}

template <typename R>
R refr_jac(const V<R>& wo, const V<R>& wi, const V<R>& m, R eta) {
  R cih = dotp(m, wi), coh = dotp(m, wo), sd = coh + eta * cih;
  if (std::fabs(sd) < R(1e-6)) return R(0);
  return std::fabs(cih) * sq(eta / sd);
}

template <typename R>
V<R> halfv(const V<R>& a, const V<R>& b) {
  V<R> h = unit(a + b);
  return h.z < R(0) ? -h : h;
}

template <typename R>
V<R> reflect_about(const V<R>& i, const V<R>& n) { return i - n * (R(2) * dotp(n, i)); }

// evaluate(): f and the density with which sample() produces wi.
template <typename R>
V<R> disney_f(const Disney<R>& b, const Frame<R>& fr, const V<R>& wo, const V<R>& wi, bool back, R& pdf) {
  pdf = R(0);
  if (b.transmission > R(0.5)) {
    V<R> ol = fr.local(wo), il = fr.local(wi);
    // relative index n_i / n_t with sample()'s conventions (see the product's
    // disney_eval for the deviation from the reference's evaluate())
    R eta = back ? R(1) / b.eta : b.eta;
    if (eta == R(1)) return V<R>();
    R ieta = R(1) / eta;
    R ax, ay, ct;
    alpha_of(b.roughness, b.anisotropic, ax, ay);
    V<R> val, m;
    R jac;
    if (il.z * ol.z >= R(0)) {
      m = halfv(ol, il);
      R cw = dotp(ol, m), f = fres(cw, eta, ct);
      val = refl_eval(b.color, ol, il, m, ax, ay, f);
      pdf = f;
      jac = cw == R(0) ? R(0) : R(1) / (R(4) * std::fabs(cw));
    } else {
      m = halfv(ol, il * ieta);
      R cw = dotp(ol, m), f = fres(cw, eta, ct);
      val = refr_eval(ieta, b.color, ol, il, m, ax, ay, R(1) - f);
      pdf = R(1) - f;
      jac = refr_jac(ol, il, m, ieta);
    }
    pdf *= jac * pdf_ggx(ax, ay, ol, m);
    return val;
  }
  R w[4];
  if (!weights(b, w)) return V<R>();
  V<R> v, c;
  if (w[0] + w[1] > R(0)) {
    V<R> m = unit(wi + wo);
    if (w[0] > R(0)) { pdf += w[0] * diffuse_lobe(b, fr.n, wo, wi, m, c); v = v + c; }
    if (w[1] > R(0)) { pdf += w[1] * sheen_lobe(b, wi, m, c); v = v + c; }
  }
  if (w[2] > R(0)) {
    V<R> ol = fr.local(wo), il = fr.local(wi);
    R ax, ay;
    alpha_of(b.roughness, b.anisotropic, ax, ay);
    R p = mf_lobe(b, ax, ay, ol, il, unit(ol + il), false, c);
    if (p > R(0)) { pdf += w[2] * p; v = v + c; }
  }
  if (w[3] > R(0)) {
    V<R> ol = fr.local(wo), il = fr.local(wi);
    R a = cc_alpha(b);
    R p = mf_lobe(b, a, a, ol, il, unit(ol + il), true, c);
    if (p > R(0)) { pdf += w[3] * p; v = v + c; }
  }
  return v;
}

template <typename R>
struct Smp {
  V<R> f, wi;
  R pdf = 0;
};

template <typename R>
Smp<R> disney_sample(const Disney<R>& b, const Frame<R>& fr, const V<R>& wo, R dist, bool back, R r0, R r1, R r2) {
  Smp<R> s;
  if (r0 < b.transmission) {
    R r3 = r0 / b.transmission;
    V<R> ol = fr.local(wo);
    R eta = back ? R(1) / b.eta : b.eta;
    if (eta == R(1)) return s;
    R ax, ay, ct;
    alpha_of(b.roughness, b.anisotropic, ax, ay);
    V<R> m = sample_ggx(ax, ay, ol, r1, r3);
    R cw = cl(dotp(ol, m), R(-1), R(1));
    R f = fres(cw, eta, ct);
    V<R> il, val;
    R pdf, jac;
    if (r2 < f) {
      il = reflect_about(-ol, m);
      if (il.z * ol.z <= R(0)) return s;
      val = refl_eval(b.color, ol, il, m, ax, ay, f);
      pdf = f;
      jac = cw == R(0) ? R(0) : R(1) / (R(4) * std::fabs(cw));
    } else {
      // refracted_direction (:105-113) is passed eta itself (:432)
      R rcp = R(1) / eta;
      il = cw > R(0) ? m * (eta * cw - ct) - ol * eta : m * (eta * cw + ct) - ol * eta;
      il = il * ((R(3) - dotp(il, il)) * R(0.5));
      if (il.z * ol.z > R(0)) return s;
      val = refr_eval(rcp, b.color, ol, il, m, ax, ay, R(1) - f);
      pdf = R(1) - f;
      jac = refr_jac(ol, il, m, rcp);
    }
    pdf *= jac * pdf_ggx(ax, ay, ol, m);
    if (!(pdf > R(1e-6))) return s;
    if (back) {
      val = val * V<R>(std::exp(-b.transmittance.x * dist * R(2)), std::exp(-b.transmittance.y * dist * R(2)),
                       std::exp(-b.transmittance.z * dist * R(2)));
    }
    s.f = val;
    s.wi = fr.world(il);
    s.pdf = pdf;
    return s;
  }
  R r3 = (r0 - b.transmission) / (R(1) - b.transmission);
  R w[4];
  if (!weights(b, w)) return s;
  R cx = w[0], cy = w[0] + w[1], cz = w[0] + w[1] + w[2];
  int lobe;
  V<R> wi, val;
  R prob;
  if (r3 < cy) {
    R rr = r3 / cy, t1 = R(2) * C<R>::pi * rr, t2 = std::sqrt(R(1) - r1);
    wi = fr.world(V<R>(std::cos(t1) * t2, std::sin(t1) * t2, std::sqrt(r1)));
    V<R> m = unit(wi + wo);
    if (r3 < cx) { lobe = 0; prob = w[0] * diffuse_lobe(b, fr.n, wo, wi, m, val); }
    else { lobe = 1; prob = w[1] * sheen_lobe(b, wi, m, val); }
  } else {
    bool cc = !(r3 < cz);
    lobe = cc ? 3 : 2;
    V<R> ol = fr.local(wo);
    if (ol.z == R(0)) return s;
    R ax, ay;
    if (cc) ax = ay = cc_alpha(b);
    else alpha_of(b.roughness, b.anisotropic, ax, ay);
    R rr = cc ? (r3 - cz) / (R(1) - cz) : (r3 - cy) / (cz - cy);
    V<R> m = cc ? sample_gtr1(ax, rr, r1) : sample_ggx(ax, ay, ol, rr, r1);
    V<R> il = reflect_about(-ol, m);
    if (il.z == R(0)) return s;
    R coh = dotp(ol, m);
    R cp = (cc ? D_gtr1(ax, m) * std::fabs(m.z) : pdf_ggx(ax, ay, ol, m)) / std::fabs(R(4) * coh);
    if (cp < R(1e-6)) return s;
    val = cc ? cc_F(b, ol, m) * (D_gtr1(ax, m) * G_gtr1(ax, il, ol))
             : spec_F(b, ol, m) * (D_ggx(ax, ay, m) * G_ggx(ax, ay, il, ol));
    val = val * (R(1) / std::fabs(R(4) * ol.z * il.z));
    prob = w[lobe] * cp;
    wi = fr.world(il);
  }
  // add the other lobes in the order diffuse, sheen, specular, clearcoat
  V<R> c;
  V<R> m = unit(wi + wo);
  if (lobe != 0 && w[0] > R(0)) { prob += w[0] * diffuse_lobe(b, fr.n, wo, wi, m, c); val = val + c; }
  if (lobe != 1 && w[1] > R(0)) { prob += w[1] * sheen_lobe(b, wi, m, c); val = val + c; }
  if (lobe != 2 && w[2] > R(0)) {
    V<R> ol = fr.local(wo), il = fr.local(wi);
    R ax, ay;
    alpha_of(b.roughness, b.anisotropic, ax, ay);
    prob += w[2] * mf_lobe(b, ax, ay, ol, il, unit(ol + il), false, c);
    val = val + c;
  }
  if (lobe != 3 && w[3] > R(0)) {
    V<R> ol = fr.local(wo), il = fr.local(wi);
    R a = cc_alpha(b);
    prob += w[3] * mf_lobe(b, a, a, ol, il, unit(ol + il), true, c);
    val = val + c;
  }
  if (!(prob > R(1e-6))) return s;
  s.f = val;
  s.wi = wi;
  s.pdf = prob;
  return s;
}

// ---- sky / sun (sky_bindings.wgsl) --------------------------------------------
template <typename R>
V<R> sky_rad(const Sky& k, const V<R>& d) {
  if (d.y >= R(0)) {
    R t = std::sqrt(std::min(d.y, R(1)));
    return V<R>(lerp((R)k.horizon[0], (R)k.zenith[0], t), lerp((R)k.horizon[1], (R)k.zenith[1], t),
                lerp((R)k.horizon[2], (R)k.zenith[2], t));
  }
  return f3<R>(k.ground);
}

template <typename R>
R sun_I(const Sky& k, const V<R>& d) {
  if (!(k.sun_enabled > 0.0f)) return R(0);
  V<R> l(-(R)k.sun_direction[0], -(R)k.sun_direction[1], -(R)k.sun_direction[2]);
  R cut = C<R>::pi / R(1.95);
  R it = (R)k.sun_intensity * std::max(R(0), R(1) - std::exp(-((cut - std::acos(cl(l.y, R(-1), R(1)))) / R(1.4))));
  return dotp(d, l) > std::cos((R)k.sun_size * R(0.1)) ? it : R(0);
}

}  // namespace ocr

// Traversal entry points of oracle_core.cpp with the precision of the caller.
namespace orc {
struct ClosestResult {
  double t, u, v;
  uint32_t prim;
};
bool oracle_closest_f(const Bvh* b, const float* ray8, float* hit4);
bool oracle_any_f(const Bvh* b, const float* ray8);
bool oracle_closest_d(const Bvh* b, const double* o, const double* d, double tmin, double tmax,
                      double* t, double* u, double* v, uint32_t* prim);
bool oracle_any_d(const Bvh* b, const double* o, const double* d, double tmin, double tmax);
}  // namespace orc

namespace ocr {

struct Scene {
  const float* pos;
  const float* nrm;
  const uint32_t* mat_ids;
  const Mat* mats;
  uint32_t n_tris;
  const Sky* sky;
  const Cam* cam;
  const orc::Bvh* bvh;
  std::vector<uint32_t> emissive;
};

struct Cfg {
  uint32_t width, height, row_start, rows, spp, max_bounces, rr_start, seed, frame_idx;
};

template <typename R>
bool closest(const Scene& sc, const V<R>& o, const V<R>& d, R& t, R& u, R& v, uint32_t& prim) {
  if constexpr (sizeof(R) == 4) {
    float ray[8] = {o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, 1e30f};
    float h[4];
    if (!orc::oracle_closest_f(sc.bvh, ray, h)) return false;
    t = h[0];
    u = h[1];
    v = h[2];
    std::memcpy(&prim, &h[3], 4);
    return true;
  } else {
    double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    return orc::oracle_closest_d(sc.bvh, oo, dd, 0.0, 1e30, &t, &u, &v, &prim);
  }
}

template <typename R>
bool occluded(const Scene& sc, const V<R>& o, const V<R>& d, R tmax) {
  if constexpr (sizeof(R) == 4) {
    float ray[8] = {o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, tmax};
    return orc::oracle_any_f(sc.bvh, ray);
  } else {
    double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    return orc::oracle_any_d(sc.bvh, oo, dd, 0.0, tmax);
  }
}

template <typename R>
void tri_pts(const Scene& sc, uint32_t p, V<R>& a, V<R>& b, V<R>& c) {
  a = f3<R>(sc.pos + 9 * (size_t)p);
  b = f3<R>(sc.pos + 9 * (size_t)p + 3);
  c = f3<R>(sc.pos + 9 * (size_t)p + 6);
}

template <typename R>
R area_of(const V<R>& a, const V<R>& b, const V<R>& c) { return R(0.5) * len(crossp(b - a, c - a)); }

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
  R ux = ((R)x + jx) / (R)cf.width * R(2) - R(1);
  R uy = -(((R)y + jy) / (R)cf.height * R(2) - R(1));
  const float* m = sc.cam->m;
  R th = std::tan(R(0.5) * (R)sc.cam->fov_deg * C<R>::pi / R(180));
  R asp = (R)cf.width / (R)cf.height;
  V<R> dl = unit(V<R>(ux * th * asp, uy * th, R(-1)));
  V<R> cr = unit(V<R>(m[0], m[1], m[2])), cu = unit(V<R>(m[4], m[5], m[6])), cb = unit(V<R>(m[8], m[9], m[10]));
  V<R> ro(m[12], m[13], m[14]);
  V<R> rd = cr * dl.x + cu * dl.y + cb * dl.z;
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
      V<R> Le = sky_rad<R>(k, rd);
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
    bool back = dotp(wo, ns) < R(0);
    if (back) ns = -ns;
    bool emits = mt.emission[0] > 0.0f || mt.emission[1] > 0.0f || mt.emission[2] > 0.0f;
    if (emits && dotp(ng, rd) < R(0)) {
      R w = R(1);
      if (pdf_prev > R(0) && nee_side && ne > 0) {
        R p_l = (R(1) - p_sun) / (R)ne / area_of(a, b, c) * (t * t) / (-dotp(ng, rd));
        w = power_h(pdf_prev, p_l);
      }
      L = L + T * f3<R>(mt.emission) * w;
    }
    const bool lambert = mt.bsdf_model == 1u;
    const V<R> albedo = f3<R>(mt.color);
    const Disney<R> bd = from_mat<R>(mt);
    const Frame<R> fr = basis(ns);
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
        Le = f3<R>(lm.emission);
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
          f = disney_f(bd, fr, wo, wi, back, p_b);
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
      bs = disney_sample(bd, fr, wo, t, back, r0, r1, r2);
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
            int threads, uint64_t* counts) {
  // Work items are 64x64 tiles of the slab (the reference CPU tracer's block
  // loop, crates/appearance-path-tracer/src/lib.rs:107-190), or chunks of an
  // explicit pixel list.
  const bool list = pixels != nullptr;
  const uint32_t tiles_x = (cf.width + 63) / 64, tiles_y = (cf.rows + 63) / 64;
  const uint32_t items = list ? (n_pixels + 255) / 256 : tiles_x * tiles_y;
  std::atomic<uint32_t> next{0};
  std::vector<uint64_t> cnt((size_t)threads * 2, 0);
  auto shade_px = [&](uint32_t x, uint32_t yl, float* o, uint64_t* nr) {
    V<R> acc;
    for (uint32_t s = 0; s < cf.spp; ++s) acc = acc + path<R>(sc, cf, x, cf.row_start + yl, s, nr);
    R inv = R(1) / (R)cf.spp;
    o[0] = (float)(acc.x * inv);
    o[1] = (float)(acc.y * inv);
    o[2] = (float)(acc.z * inv);
  };
  auto work = [&](int tid) {
    uint64_t* nr = &cnt[2 * (size_t)tid];
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
  if (counts) {
    counts[0] = counts[1] = 0;
    for (int t = 0; t < threads; ++t) {
      counts[0] += cnt[2 * (size_t)t];
      counts[1] += cnt[2 * (size_t)t + 1];
    }
  }
}

}  // namespace ocr

extern "C" {

// Renders rows [row_start, row_start+rows) of a width x height frame, spp
// samples per pixel, and writes the linear mean radiance (3 floats per
// pixel, row order) to out.  If pixels != NULL only those slab-local pixel
// indices are rendered, in list order.  precision: 32 or 64.
// counts[0..1] receive extension and shadow rays traced.
int or_render(const float* pos, const float* nrm, const uint32_t* mat_ids, const void* mats,
              uint32_t n_tris, void* bvh, const void* sky, const void* cam, const uint32_t* cfg9,
              const uint32_t* pixels, uint32_t n_pixels, int precision, int threads, float* out,
              uint64_t* counts) {
  ocr::Scene sc;
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
  ocr::Cfg cf{cfg9[0], cfg9[1], cfg9[2], cfg9[3], cfg9[4], cfg9[5], cfg9[6], cfg9[7], cfg9[8]};
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (precision == 64) ocr::render<double>(sc, cf, pixels, n_pixels, out, threads, counts);
  else ocr::render<float>(sc, cf, pixels, n_pixels, out, threads, counts);
  return 0;
}

// Disney BSDF probes for the invariance tests (FP64).  frame = normal only.
// eval: returns pdf, writes f.  sample: writes wi, f, returns pdf.
double or_disney_eval(const void* mat, const double* n, const double* wo, const double* wi, int back,
                      double* f) {
  ocr::Disney<double> b = ocr::from_mat<double>(*static_cast<const ocr::Mat*>(mat));
  ocr::Frame<double> fr = ocr::basis(ocr::V<double>(n[0], n[1], n[2]));
  double pdf;
  ocr::V<double> v = ocr::disney_f(b, fr, ocr::V<double>(wo[0], wo[1], wo[2]), ocr::V<double>(wi[0], wi[1], wi[2]),
                                   back != 0, pdf);
  f[0] = v.x;
  f[1] = v.y;
  f[2] = v.z;
  return pdf;
}

double or_disney_sample(const void* mat, const double* n, const double* wo, int back, double r0, double r1,
                        double r2, double* wi, double* f) {
  ocr::Disney<double> b = ocr::from_mat<double>(*static_cast<const ocr::Mat*>(mat));
  ocr::Frame<double> fr = ocr::basis(ocr::V<double>(n[0], n[1], n[2]));
  ocr::Smp<double> s = ocr::disney_sample(b, fr, ocr::V<double>(wo[0], wo[1], wo[2]), 0.0, back != 0, r0, r1, r2);
  wi[0] = s.wi.x;
  wi[1] = s.wi.y;
  wi[2] = s.wi.z;
  f[0] = s.f.x;
  f[1] = s.f.y;
  f[2] = s.f.z;
  return s.pdf;
}

}  // extern "C"
