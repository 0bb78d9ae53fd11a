// Shading library of the integrator: orthonormal bases, samplers, the Disney
// BSDF of the reference (AppleSeed-derived, shared/material/disney_bsdf.wgsl,
// ggxmdf.wgsl), the Lambertian BRDF of the CPU tracer
// (crates/appearance-path-tracer/src/reflectance/diffuse.rs) and the sky/sun.
// Licence notices.  The Disney BSDF below restates the reference's
// shared/material/disney_bsdf.wgsl, which carries this notice:
//
//   This code has been adapted from AppleSeed: https://appleseedhq.net
//   The AppleSeed software is released under the MIT license.
//   Copyright (c) 2014-2018 Esteban Tovagliari, The appleseedhq Organization.
//
// and the reference itself (Appearance) is:
//
//   MIT License. Copyright (c) 2025 Temporal Interactive
//
// Both under these terms:
//
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to permit
//   persons to whom the Software is furnished to do so, subject to the
//   following conditions:
//
//   The above copyright notice and this permission notice shall be included
//   in all copies or substantial portions of the Software.
//
//   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS
//   OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
//   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN
//   NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
//   DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR
//   OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE
//   USE OR OTHER DEALINGS IN THE SOFTWARE.
#pragma once

#include "apt_math.h"
#include "../../include/apt.h"

namespace apt {

// ---------------------------------------------------------------------------
// Frames and samplers
// ---------------------------------------------------------------------------
struct Frame {  // columns t1, t2, n: tangent -> world is t1*x + t2*y + n*z
  v3 t1, t2, n;
};

// math.wgsl:71-86 (Duff et al. 2017).
APT_HD Frame onb(v3 n) {
  Frame f;
  if (n.z < 0.0f) {
    float a = 1.0f / (1.0f - n.z);
    float b = n.x * n.y * a;
    f.t1 = mk(1.0f - n.x * n.x * a, -b, n.x);
    f.t2 = mk(b, n.y * n.y * a - 1.0f, -n.y);
  } else {
    float a = 1.0f / (1.0f + n.z);
    float b = -n.x * n.y * a;
    f.t1 = mk(1.0f - n.x * n.x * a, b, -n.x);
    f.t2 = mk(b, 1.0f - n.y * n.y * a, -n.y);
  }
  f.n = n;
  return f;
}
APT_HD v3 to_world(const Frame& f, v3 l) { return f.t1 * l.x + f.t2 * l.y + f.n * l.z; }
APT_HD v3 to_local(const Frame& f, v3 w) { return mk(dot(f.t1, w), dot(f.t2, w), dot(f.n, w)); }

// math.wgsl:89-99.
APT_HD v3 perpendicular(v3 u) {
  v3 a = mk(fabsf(u.x), fabsf(u.y), fabsf(u.z));
  uint32_t xm = ((a.x - a.y) < 0.0f && (a.x - a.z) < 0.0f) ? 1u : 0u;
  uint32_t ym = (a.y - a.z) < 0.0f ? (1u ^ xm) : 0u;
  uint32_t zm = 1u ^ (xm | ym);
  return cross(u, mk((float)xm, (float)ym, (float)zm));
}

// sampling.wgsl:52-61: uniform direction in the cone of half angle `angle`
// around `dir`.
APT_HD v3 perturb_direction(float u0, float u1, v3 dir, float angle) {
  float h = cosf(angle);
  float phi = TWO_PI * u0;
  float z = h + (1.0f - h) * u1;
  float sin_t = sqrtf(fmaxf(0.0f, 1.0f - z * z));
  float x = cosf(phi) * sin_t, y = sinf(phi) * sin_t;
  v3 b = perpendicular(dir);
  v3 t = cross(b, dir);
  return b * x + t * y + dir * z;
}

// Cosine hemisphere via the concentric disk, as the CPU tracer's
// sample_cosine_hemisphere (src/sampling/mod.rs:31-53).
APT_HD v3 cosine_hemisphere_concentric(float u0, float u1) {
  float ox = 2.0f * u0 - 1.0f, oy = 2.0f * u1 - 1.0f;
  float dx = 0.0f, dy = 0.0f;
  if (ox != 0.0f || oy != 0.0f) {
    float r, th;
    if (fabsf(ox) > fabsf(oy)) {
      r = ox;
      th = (PI * 0.25f) * (oy / ox);
    } else {
      r = oy;
      th = (PI * 0.5f) - (PI * 0.25f) * (ox / oy);
    }
    dx = r * cosf(th);
    dy = r * sinf(th);
  }
  return mk(dx, dy, sqrtf(fmaxf(0.0f, 1.0f - dx * dx - dy * dy)));
}

// Power heuristic with one sample per strategy (src/sampling/mod.rs:84-88).
APT_HD float power_heuristic(float a, float b) {
  float a2 = a * a, b2 = b * b;
  return a2 + b2 > 0.0f ? a2 / (a2 + b2) : 0.0f;
}

// ---------------------------------------------------------------------------
// GGX / GTR1 microfacet distributions (ggxmdf.wgsl)
// ---------------------------------------------------------------------------
struct Alpha2 {
  float x, y;
};

APT_HD Alpha2 alpha_from_roughness(float roughness, float aniso) {  // ggxmdf.wgsl:20-35
  float sq = roughness * roughness;
  float aspect = sqrtf(1.0f + aniso * (aniso < 0.0f ? 0.9f : -0.9f));
  return {fmaxf(0.001f, sq / aspect), fmaxf(0.001f, sq * aspect)};
}

APT_HD float ggx_d(Alpha2 a, v3 m) {  // ggxmdf.wgsl:46-63
  if (m.z == 0.0f) return a.x * a.x * INV_PI;
  float c2 = m.z * m.z;
  float s = sqrtf(fmaxf(0.0f, 1.0f - c2));
  float tan2 = (1.0f - c2) / c2;
  float sr;
  if (a.x == a.y || s == 0.0f) sr = 1.0f / (a.x * a.x);
  else sr = sqr(m.x / (s * a.x)) + sqr(m.y / (s * a.y));
  return 1.0f / (PI * a.x * a.y * sqr(c2) * sqr(1.0f + tan2 * sr));
}

APT_HD float ggx_lambda(Alpha2 a, v3 v) {  // ggxmdf.wgsl:65-83
  if (v.z == 0.0f) return 0.0f;
  float c2 = v.z * v.z;
  float s = sqrtf(fmaxf(0.0f, 1.0f - c2));
  float pr;
  if (a.x == a.y || s == 0.0f) pr = a.x;
  else pr = sqrtf(sqr((v.x * a.x) / s) + sqr((v.y * a.y) / s));
  float tan2 = (s * s) / c2;
  float a2r = pr * pr * tan2;
  return (-1.0f + sqrtf(1.0f + a2r)) * 0.5f;
}

APT_HD float ggx_g(Alpha2 a, v3 wi, v3 wo) { return 1.0f / (1.0f + ggx_lambda(a, wo) + ggx_lambda(a, wi)); }
APT_HD float ggx_g1(Alpha2 a, v3 v) { return 1.0f / (1.0f + ggx_lambda(a, v)); }

APT_HD v3 ggx_sample(Alpha2 a, v3 v, float r0, float r1) {  // ggxmdf.wgsl:94-137 (Heitz 2014)
  float sgn = v.z < 0.0f ? -1.0f : 1.0f;
  v3 st = normalize(mk(sgn * v.x * a.x, sgn * v.y * a.y, sgn * v.z));
  v3 t1 = v.z < 0.9999f ? normalize(cross(st, mk(0.0f, 0.0f, 1.0f))) : mk(1.0f, 0.0f, 0.0f);
  v3 t2 = cross(t1, st);
  float aa = 1.0f / (1.0f + st.z);
  float r = sqrtf(r0);
  float phi = r1 < aa ? r1 / aa * PI : PI + (r1 - aa) / (1.0f - aa) * PI;
  float p1 = r * cosf(phi);
  float p2 = r * sinf(phi);
  if (r1 >= aa) p2 *= st.z;
  v3 h = t1 * p1 + t2 * p2 + st * sqrtf(fmaxf(0.0f, 1.0f - p1 * p1 - p2 * p2));
  return normalize(mk(h.x * a.x, h.y * a.y, fmaxf(0.0f, h.z)));
}

APT_HD float ggx_pdf(Alpha2 a, v3 v, v3 m) {  // ggxmdf.wgsl:139-145
  if (v.z == 0.0f) return 0.0f;
  return ggx_g1(a, v) * fabsf(dot(v, m)) * ggx_d(a, m) / fabsf(v.z);
}

APT_HD float gtr1_d(float ax, v3 m) {  // ggxmdf.wgsl:156-162
  float al = clampf(ax, 0.001f, 0.999f);
  float a2 = al * al;
  float a = (a2 - 1.0f) / (PI * logf(a2));
  float b = 1.0f / (1.0f + (a2 - 1.0f) * m.z * m.z);
  return a * b;
}

APT_HD float gtr1_lambda(float ax, v3 v) {  // ggxmdf.wgsl:164-184
  if (v.z == 0.0f) return 0.0f;
  float c2 = v.z * v.z;
  float s = sqrtf(fmaxf(0.0f, 1.0f - c2));
  if (s == 0.0f) return 0.0f;
  float cot2 = c2 / (s * s);
  float cot = sqrtf(cot2);
  float al2 = sqr(clampf(ax, 0.001f, 0.999f));
  float a = sqrtf(cot2 + al2);
  float b = sqrtf(cot2 + 1.0f);
  float c = logf(cot + b);
  float d = logf(cot + a);
  return (a - b + cot * (c - d)) / (cot * logf(al2));
}

APT_HD float gtr1_g(float ax, v3 wi, v3 wo) { return 1.0f / (1.0f + gtr1_lambda(ax, wo) + gtr1_lambda(ax, wi)); }

APT_HD v3 gtr1_sample(float ax, float r0, float r1) {  // ggxmdf.wgsl:194-209
  float al = clampf(ax, 0.001f, 0.999f);
  float a2 = al * al;
  float c2 = (1.0f - powf(a2, 1.0f - r0)) / (1.0f - a2);
  float s = sqrtf(fmaxf(0.0f, 1.0f - c2));
  float phi = TWO_PI * r1;
  return mk(cosf(phi) * s, sinf(phi) * s, sqrtf(c2));
}

APT_HD float gtr1_pdf(float ax, v3 m) { return gtr1_d(ax, m) * fabsf(m.z); }

// ---------------------------------------------------------------------------
// Disney BSDF (disney_bsdf.wgsl:22-682)
// ---------------------------------------------------------------------------
APT_HD float schlick_fresnel(float u) {
  float m = clampf(1.0f - u, 0.0f, 1.0f);
  float m2 = m * m;
  return m2 * m2 * m;
}

APT_HD float fresnel_dielectric(float eta, float ci, float ct) {  // :28-36
  if (ci == 0.0f && ct == 0.0f) return 1.0f;
  float k0 = eta * ct, k1 = eta * ci;
  return 0.5f * (sqr((ci - k0) / (ci + k0)) + sqr((ct - k1) / (ct + k1)));
}

APT_HD float fresnel_reflectance(float ci, float eta, float& ct) {  // :38-47
  float st2 = (1.0f - ci * ci) * (eta * eta);
  if (st2 > 1.0f) {
    ct = 0.0f;
    return 1.0f;
  }
  ct = fminf(sqrtf(fmaxf(1.0f - st2, 0.0f)), 1.0f);
  return fresnel_dielectric(eta, fabsf(ci), ct);
}

struct Disney {
  v3 color;
  float metallic;
  v3 transmittance;
  float subsurface;
  float luminance;
  float specular;
  float roughness;
  v3 spec_tint;
  float anisotropic;
  float sheen;
  v3 sheen_tint;
  float clearcoat;
  float clearcoat_gloss;
  float transmission;
  float eta;
};

// Material::from_material_descriptor for untextured descriptors
// (material_pool_bindings.wgsl:128-154) followed by DisneyBsdf::from_material
// (disney_bsdf.wgsl:226-244).
APT_HD Disney disney_from_material(const apt_material& m) {
  Disney d;
  d.color = mk(m.color[0], m.color[1], m.color[2]);
  d.metallic = m.metallic;
  d.transmittance = mk(m.absorption[0], m.absorption[1], m.absorption[2]);
  d.subsurface = m.subsurface;
  d.luminance = 1.0f;
  d.specular = fmaxf(m.specular, 1.0f - m.roughness);
  d.roughness = m.roughness;
  d.spec_tint = mk(m.specular_tint[0], m.specular_tint[1], m.specular_tint[2]);
  d.anisotropic = m.anisotropic;
  d.sheen = m.sheen;
  d.sheen_tint = mk(m.sheen_tint[0], m.sheen_tint[1], m.sheen_tint[2]);
  d.clearcoat = m.clearcoat;
  d.clearcoat_gloss = 1.0f - m.clearcoat_roughness;
  d.transmission = m.transmission;
  d.eta = m.eta;
  return d;
}

APT_HD float cc_roughness(const Disney& b) { return lerpf(0.1f, 0.001f, b.clearcoat_gloss); }

APT_HD v3 mix3(v3 a, v3 b, float t) { return a * (1.0f - t) + b * t; }

APT_HD v3 specular_fresnel(const Disney& b, v3 o, v3 h) {  // :250-256
  v3 v = b.spec_tint * (b.specular * 0.08f);
  v = mix3(v, b.color, b.metallic);
  float f = schlick_fresnel(fabsf(dot(o, h)));
  return v * (1.0f - f) + mk(f, f, f);
}

APT_HD v3 clearcoat_fresnel(const Disney& b, v3 o, v3 h) {  // :258-261
  float f = lerpf(0.04f, 1.0f, schlick_fresnel(fabsf(dot(o, h)))) * 0.25f * b.clearcoat;
  return mk(f, f, f);
}

APT_HD v3 eval_reflection(v3 col, v3 wo, v3 wi, v3 m, Alpha2 a, float f) {  // :49-58
  float denom = fabsf(4.0f * wo.z * wi.z);
  if (denom == 0.0f) return mk(0.0f, 0.0f, 0.0f);
  return col * (f * ggx_d(a, m) * ggx_g(a, wi, wo) / denom);
}

APT_HD v3 eval_refraction(float eta, v3 col, v3 wo, v3 wi, v3 m, Alpha2 a, float t) {  // :60-80 (adjoint = false)
  if (wo.z == 0.0f || wi.z == 0.0f) return mk(0.0f, 0.0f, 0.0f);
  float cih = dot(m, wi), coh = dot(m, wo);
  float dots = (cih * coh) / (wi.z * wo.z);
  float sd = coh + eta * cih;
  if (fabsf(sd) < 1e-6f) return mk(0.0f, 0.0f, 0.0f);
  float mult = fabsf(dots) * t * ggx_d(a, m) * ggx_g(a, wi, wo) / (sd * sd);
  mult *= eta * eta;
  return col * mult;
}

APT_HD float reflection_jacobian(float coh) { return coh == 0.0f ? 0.0f : 1.0f / (4.0f * fabsf(coh)); }

APT_HD float refraction_jacobian(v3 wo, v3 wi, v3 m, float eta) {  // :90-99
  float cih = dot(m, wi), coh = dot(m, wo);
  float sd = coh + eta * cih;
  if (fabsf(sd) < 1e-6f) return 0.0f;
  return fabsf(cih) * sqr(eta / sd);
}

APT_HD v3 refracted_direction(v3 wo, v3 m, float cwm, float ct, float rcp_eta) {  // :105-113
  v3 wi = cwm > 0.0f ? m * (rcp_eta * cwm - ct) - wo * rcp_eta : m * (rcp_eta * cwm + ct) - wo * rcp_eta;
  return wi * ((3.0f - dot(wi, wi)) * 0.5f);
}

APT_HD v3 half_vec(v3 a, v3 b) {
  v3 h = normalize(a + b);
  return h.z < 0.0f ? -h : h;
}

// evaluate_diffuse / evaluate_sheen (:347-378): value and lobe pdf.
APT_HD float eval_diffuse(const Disney& b, v3 n, v3 wo, v3 wi, v3 m, v3& value) {
  float con = dot(n, wo), cin = dot(n, wi), cih = dot(wi, m);
  float fl = schlick_fresnel(cin), fv = schlick_fresnel(con);
  float fd = 0.0f;
  if (b.subsurface != 1.0f) {
    float fd90 = 0.5f + 2.0f * cih * cih * b.roughness;
    fd = lerpf(1.0f, fd90, fl) * lerpf(1.0f, fd90, fv);
  }
  if (b.subsurface > 0.0f) {
    float f90 = cih * cih * b.roughness;
    float fss = lerpf(1.0f, f90, fl) * lerpf(1.0f, f90, fv);
    float ss = 1.25f * (fss * (1.0f / (fabsf(con) + fabsf(cin)) - 0.5f) + 0.5f);
    fd = lerpf(fd, ss, b.subsurface);
  }
  value = b.color * (fd * INV_PI * (1.0f - b.metallic));
  return fabsf(cin) * INV_PI;
}

APT_HD float eval_sheen(const Disney& b, v3 wo, v3 wi, v3 m, v3& value) {
  float fh = schlick_fresnel(dot(wi, m));
  value = b.sheen_tint * (fh * b.sheen * (1.0f - b.metallic));
  return 1.0f / (2.0f * PI);
}

// evaluate_mf (:314-345): microfacet lobe value (already divided by
// |4 wo.z wi.z|) and its pdf.
APT_HD float eval_mf(const Disney& b, float ax, float ay, v3 wo, v3 wi, v3 m, bool cc, v3& value) {
  value = mk(0.0f, 0.0f, 0.0f);
  if (wo.z == 0.0f || wi.z == 0.0f) return 0.0f;
  float coh = dot(wo, m);
  if (coh == 0.0f) return 0.0f;
  if (cc) {
    value = clearcoat_fresnel(b, wo, m) * (gtr1_d(ax, m) * gtr1_g(ax, wi, wo) / fabsf(4.0f * wo.z * wi.z));
    return gtr1_pdf(ax, m) / fabsf(4.0f * coh);
  }
  Alpha2 a{ax, ay};
  value = specular_fresnel(b, wo, m) * (ggx_d(a, m) * ggx_g(a, wi, wo) / fabsf(4.0f * wo.z * wi.z));
  return ggx_pdf(a, wo, m) / fabsf(4.0f * coh);
}

struct LobeWeights {
  float d, s, sp, cc;
};

APT_HD bool lobe_weights(const Disney& b, LobeWeights& w) {  // :460-471
  w.d = lerpf(b.luminance, 0.0f, b.metallic);
  w.s = lerpf(b.sheen, 0.0f, b.metallic);
  w.sp = lerpf(b.specular, 1.0f, b.metallic);
  w.cc = b.clearcoat * 0.25f;
  float sum = w.d + w.s + w.sp + w.cc;
  if (sum == 0.0f) return false;
  float inv = 1.0f / sum;
  w.d *= inv;
  w.s *= inv;
  w.sp *= inv;
  w.cc *= inv;
  return true;
}

struct ShadeFrame {
  Frame f;     // front-facing shading normal basis
  Frame cc;    // clearcoat basis: the unmapped normal, or the clearcoat normal map's
};

// DisneyBsdf::evaluate (:567-682).  Returns f; pdf receives the density with
// which sample() would produce wi.  Deviations that make evaluate() and
// sample() agree, which the MIS weights need: the reference multiplies the
// specular and clearcoat lobe pdfs by their lobe weight twice (:658-675),
// here once, as in sample(); and the transmission branch uses sample()'s
// relative index eta = n_i/n_t (back face: 1/eta, :401-405) with the Fresnel
// term, refraction half vector and jacobian of sample() (:418, :436-438),
// where the reference tests wol.z of an always front-facing frame (:576-580)
// and uses the reciprocal index (:596, :602-609).
APT_HD v3 disney_eval(const Disney& b, const ShadeFrame& sf, v3 wo, v3 wi, bool back_face,
                      float& pdf) {
  pdf = 0.0f;
  if (b.transmission > 0.5f) {
    v3 wol = to_local(sf.f, wo), wil = to_local(sf.f, wi);
    const float eta = back_face ? 1.0f / b.eta : b.eta;  // n_i / n_t
    if (eta == 1.0f) return mk(0.0f, 0.0f, 0.0f);
    const float ieta = 1.0f / eta;
    Alpha2 a = alpha_from_roughness(b.roughness, b.anisotropic);
    v3 val, m;
    float jac, ct;
    if (wil.z * wol.z >= 0.0f) {
      m = half_vec(wol, wil);
      float cwm = dot(wol, m);
      float f = fresnel_reflectance(cwm, eta, ct);
      val = eval_reflection(b.color, wol, wil, m, a, f);
      pdf = f;  // choose_reflection_probability(1, 1, f)
      jac = reflection_jacobian(cwm);
    } else {
      m = half_vec(wol, wil * ieta);
      float cwm = dot(wol, m);
      float f = fresnel_reflectance(cwm, eta, ct);
      val = eval_refraction(ieta, b.color, wol, wil, m, a, 1.0f - f);
      pdf = 1.0f - f;
      jac = refraction_jacobian(wol, wil, m, ieta);
    }
    pdf *= jac * ggx_pdf(a, wol, m);
    return val;
  }
  LobeWeights w;
  if (!lobe_weights(b, w)) return mk(0.0f, 0.0f, 0.0f);
  v3 value = mk(0.0f, 0.0f, 0.0f);
  if (w.d + w.s > 0.0f) {
    v3 m = normalize(wi + wo);
    v3 c;
    if (w.d > 0.0f) {
      pdf += w.d * eval_diffuse(b, sf.f.n, wo, wi, m, c);
      value += c;
    }
    if (w.s > 0.0f) {
      pdf += w.s * eval_sheen(b, wo, wi, m, c);
      value += c;
    }
  }
  if (w.sp > 0.0f) {
    v3 wol = to_local(sf.f, wo), wil = to_local(sf.f, wi);
    v3 m = normalize(wol + wil);
    Alpha2 a = alpha_from_roughness(b.roughness, b.anisotropic);
    v3 c;
    float p = eval_mf(b, a.x, a.y, wol, wil, m, false, c);
    if (p > 0.0f) {
      pdf += w.sp * p;
      value += c;
    }
  }
  if (w.cc > 0.0f) {
    v3 wol = to_local(sf.cc, wo), wil = to_local(sf.cc, wi);
    v3 m = normalize(wol + wil);
    float al = cc_roughness(b);
    v3 c;
    float p = eval_mf(b, al, al, wol, wil, m, true, c);
    if (p > 0.0f) {
      pdf += w.cc * p;
      value += c;
    }
  }
  return value;
}

struct BsdfSample {
  v3 f;
  v3 wi;
  float pdf;  // 0 = no sample
  bool specular;
};

// DisneyBsdf::sample (:380-565).  dist is the hit distance (Beer absorption
// through back faces).
APT_HD BsdfSample disney_sample(const Disney& b, const ShadeFrame& sf, v3 wo, float dist,
                                bool back_face, float r0, float r1, float r2) {
  BsdfSample s;
  s.f = mk(0.0f, 0.0f, 0.0f);
  s.wi = mk(0.0f, 0.0f, 1.0f);
  s.pdf = 0.0f;
  s.specular = false;
  if (r0 < b.transmission) {
    s.specular = true;
    float r3 = r0 / b.transmission;
    v3 wol = to_local(sf.f, wo);
    float eta = back_face ? 1.0f / b.eta : b.eta;
    if (eta == 1.0f) return s;
    Alpha2 a = alpha_from_roughness(b.roughness, b.anisotropic);
    v3 m = ggx_sample(a, wol, r1, r3);
    float rcp_eta = 1.0f / eta;
    float cwm = clampf(dot(wol, m), -1.0f, 1.0f);
    float ct;
    float f = fresnel_reflectance(cwm, eta, ct);
    float jac, pdf;
    v3 wil, val;
    if (r2 < f) {
      wil = reflect(-wol, m);
      if (wil.z * wol.z <= 0.0f) return s;
      val = eval_reflection(b.color, wol, wil, m, a, f);
      pdf = f;
      jac = reflection_jacobian(cwm);
    } else {
      wil = refracted_direction(wol, m, cwm, ct, eta);
      if (wil.z * wol.z > 0.0f) return s;
      val = eval_refraction(rcp_eta, b.color, wol, wil, m, a, 1.0f - f);
      pdf = 1.0f - f;
      jac = refraction_jacobian(wol, wil, m, rcp_eta);
    }
    pdf *= jac * ggx_pdf(a, wol, m);
    if (!(pdf > 1e-6f)) return s;
    s.wi = to_world(sf.f, wil);
    if (back_face) {
      val = val * mk(expf(-b.transmittance.x * dist * 2.0f), expf(-b.transmittance.y * dist * 2.0f),
                     expf(-b.transmittance.z * dist * 2.0f));
    }
    s.f = val;
    s.pdf = pdf;
    return s;
  }

  float r3 = (r0 - b.transmission) / (1.0f - b.transmission);
  LobeWeights w;
  if (!lobe_weights(b, w)) return s;
  const float cdf_x = w.d, cdf_y = w.d + w.s, cdf_z = w.d + w.s + w.sp;
  float prob = 0.0f;
  v3 value = mk(0.0f, 0.0f, 0.0f);
  v3 wi;
  bool used_d = false, used_s = false, used_sp = false, used_cc = false;
  if (r3 < cdf_y) {
    float rr = r3 / cdf_y;
    float t1 = TWO_PI * rr, t2 = sqrtf(1.0f - r1);
    wi = to_world(sf.f, mk(cosf(t1) * t2, sinf(t1) * t2, sqrtf(r1)));  // diffuse_reflection_cos_weighted
    v3 m = normalize(wi + wo);
    if (r3 < cdf_x) {
      prob = w.d * eval_diffuse(b, sf.f.n, wo, wi, m, value);
      used_d = true;
    } else {
      prob = w.s * eval_sheen(b, wo, wi, m, value);
      used_s = true;
    }
  } else {
    bool cc = !(r3 < cdf_z);
    const Frame& fr = cc ? sf.cc : sf.f;
    v3 wol = to_local(fr, wo);
    if (wol.z == 0.0f) return s;
    float ax, ay;
    if (!cc) {
      Alpha2 a = alpha_from_roughness(b.roughness, b.anisotropic);
      ax = a.x;
      ay = a.y;
    } else {
      ax = ay = cc_roughness(b);
    }
    float rr = cc ? (r3 - cdf_z) / (1.0f - cdf_z) : (r3 - cdf_y) / (cdf_z - cdf_y);
    v3 m = cc ? gtr1_sample(ax, rr, r1) : ggx_sample(Alpha2{ax, ay}, wol, rr, r1);
    v3 wil = reflect(-wol, m);
    if (wil.z == 0.0f) return s;
    float coh = dot(wol, m);
    float cpdf = (cc ? gtr1_pdf(ax, m) : ggx_pdf(Alpha2{ax, ay}, wol, m)) / fabsf(4.0f * coh);
    if (cpdf < 1e-6f) return s;
    if (cc) {
      value = clearcoat_fresnel(b, wol, m) * (gtr1_d(ax, m) * gtr1_g(ax, wil, wol));
    } else {
      Alpha2 a{ax, ay};
      value = specular_fresnel(b, wol, m) * (ggx_d(a, m) * ggx_g(a, wil, wol));
    }
    value *= 1.0f / fabsf(4.0f * wol.z * wil.z);
    prob = (cc ? w.cc : w.sp) * cpdf;
    used_cc = cc;
    used_sp = !cc;
    wi = to_world(fr, wil);
  }
  v3 c;
  if (!used_d && !used_s && (w.d + w.s) > 0.0f) {
    v3 m = normalize(wi + wo);
    if (w.d > 0.0f) {
      prob += w.d * eval_diffuse(b, sf.f.n, wo, wi, m, c);
      value += c;
    }
    if (w.s > 0.0f) {
      prob += w.s * eval_sheen(b, wo, wi, m, c);
      value += c;
    }
  } else if (used_d || used_s) {
    v3 m = normalize(wi + wo);
    if (used_d && w.s > 0.0f) {
      prob += w.s * eval_sheen(b, wo, wi, m, c);
      value += c;
    }
    if (used_s && w.d > 0.0f) {
      prob += w.d * eval_diffuse(b, sf.f.n, wo, wi, m, c);
      value += c;
    }
  }
  if (!used_sp && w.sp > 0.0f) {
    v3 wol = to_local(sf.f, wo), wil = to_local(sf.f, wi);
    v3 m = normalize(wol + wil);
    Alpha2 a = alpha_from_roughness(b.roughness, b.anisotropic);
    prob += w.sp * eval_mf(b, a.x, a.y, wol, wil, m, false, c);
    value += c;
  }
  if (!used_cc && w.cc > 0.0f) {
    v3 wol = to_local(sf.cc, wo), wil = to_local(sf.cc, wi);
    v3 m = normalize(wol + wil);
    float al = cc_roughness(b);
    prob += w.cc * eval_mf(b, al, al, wol, wil, m, true, c);
    value += c;
  }
  if (!(prob > 1e-6f)) return s;
  s.f = value;
  s.wi = wi;
  s.pdf = prob;
  return s;
}

// ---------------------------------------------------------------------------
// Lambertian BRDF (DiffuseBxdf, reflectance/diffuse.rs:24-81)
// ---------------------------------------------------------------------------
APT_HD v3 lambert_eval(v3 albedo, const Frame& f, v3 wo, v3 wi, float& pdf) {
  float co = dot(f.n, wo), ci = dot(f.n, wi);
  if (co * ci <= 0.0f) {
    pdf = 0.0f;
    return mk(0.0f, 0.0f, 0.0f);
  }
  pdf = fabsf(ci) * INV_PI;
  return albedo * INV_PI;
}

APT_HD BsdfSample lambert_sample(v3 albedo, const Frame& f, v3 wo, float r0, float r1) {
  BsdfSample s;
  v3 l = cosine_hemisphere_concentric(r0, r1);
  if (dot(f.n, wo) < 0.0f) l.z = -l.z;
  s.wi = to_world(f, l);
  s.pdf = fabsf(l.z) * INV_PI;
  s.f = albedo * INV_PI;
  s.specular = false;
  if (!(s.pdf > 0.0f)) s.pdf = 0.0f;
  return s;
}

// ---------------------------------------------------------------------------
// Sky and sun (shared/sky_bindings.wgsl:25-78), with a gradient standing in for
// the HDR panorama.
// ---------------------------------------------------------------------------
APT_HD v3 sky_gradient(const apt_sky& k, v3 d) {
  if (d.y >= 0.0f) {
    float t = sqrtf(fminf(d.y, 1.0f));
    return mk(lerpf(k.horizon[0], k.zenith[0], t), lerpf(k.horizon[1], k.zenith[1], t),
              lerpf(k.horizon[2], k.zenith[2], t));
  }
  return mk(k.ground[0], k.ground[1], k.ground[2]);
}

APT_HD v3 sun_to(const apt_sky& k) { return mk(-k.sun_direction[0], -k.sun_direction[1], -k.sun_direction[2]); }
APT_HD float sun_half_angle(const apt_sky& k) { return k.sun_size * 0.1f; }

// Sky::sun_intensity (sky_bindings.wgsl:25-37); the sun's radiance is this
// times sun_color, both when it is sampled and when a path hits it.
APT_HD float sun_intensity(const apt_sky& k, v3 d) {
  if (!(k.sun_enabled > 0.0f)) return 0.0f;
  const float CUTOFF = PI / 1.95f;
  v3 l = sun_to(k);
  float zc = clampf(l.y, -1.0f, 1.0f);
  float it = k.sun_intensity * fmaxf(0.0f, 1.0f - expf(-((CUTOFF - acosf(zc)) / 1.4f)));
  return dot(d, l) > cosf(sun_half_angle(k)) ? it : 0.0f;
}

APT_HD float sun_solid_angle(const apt_sky& k) { return TWO_PI * (1.0f - cosf(sun_half_angle(k))); }

}  // namespace apt
