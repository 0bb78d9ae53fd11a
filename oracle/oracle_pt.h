// ORACLE - test infrastructure only (see oracle_core.cpp).
//
// Shading model shared by the oracle's path tracers (oracle_render.cpp: the
// plain integrator; oracle_restir.cpp: the ReSTIR integrator): ABI struct
// mirrors, vector maths, the RNG draw, the Disney BSDF
// (shared/material/disney_bsdf.wgsl, ggxmdf.wgsl), sky and sun
// (shared/sky_bindings.wgsl), the scene view and the traversal entry points
// of oracle_core.cpp.
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

// ---- textured materials (shared/material/material_pool_bindings.wgsl:15-154) -----
// A texture as the library receives it (apt_texture): width x height RGBA8
// texels, row major, R in the low byte.
struct Texture {
  uint32_t width, height;
  const uint32_t* rgba8;
};
// Per-vertex attributes of a textured scene (each NULL when absent), in world
// space as the library places them: uv n*6 floats, tan and bit n*9 floats (the
// unit tangent and bitangent of each vertex), and the scene's textures.
struct Attribs {
  const float* uv;
  const float* tan;
  const float* bit;
  const Texture* tex;
  uint32_t n_tex, pad;
  const float* sky_rgb;  // sky panorama: sky_w x sky_h RGB floats, row 0 at the zenith (NULL: gradient)
  uint32_t sky_w, sky_h;
};
constexpr uint32_t NO_TEX = 0xFFFFFFFFu;

// textureSampleLevel(tex, sampler, uv, 0.0) (material_pool_bindings.wgsl:15-17)
// with the reference's sampler (linear filtering, repeat addressing,
// src/scene_resources/material_pool.rs:68-74) on an RGBA8 unorm texture: the
// sample point in texel space is (u W - 1/2, v H - 1/2); the four texels around
// it, their indices wrapped modulo the size, are weighted by the fractional
// parts.  A texel channel c reads as c / 255.
template <typename R>
void tex_sample(const Texture& t, R u, R v, R out[4]) {
  R x = u * (R)t.width - R(0.5), y = v * (R)t.height - R(0.5);
  if (!std::isfinite(x)) x = R(-0.5);
  if (!std::isfinite(y)) y = R(-0.5);
  const R fx = std::floor(x), fy = std::floor(y);
  const R ax = x - fx, ay = y - fy;
  auto wrap = [](double i, uint32_t n) {
    double m = std::fmod(i, (double)n);
    if (m < 0) m += n;
    return std::min((uint32_t)m, n - 1);
  };
  const uint32_t x0 = wrap(fx, t.width), x1 = wrap((double)fx + 1.0, t.width);
  const uint32_t y0 = wrap(fy, t.height), y1 = wrap((double)fy + 1.0, t.height);
  const uint32_t q[4] = {t.rgba8[(size_t)y0 * t.width + x0], t.rgba8[(size_t)y0 * t.width + x1],
                         t.rgba8[(size_t)y1 * t.width + x0], t.rgba8[(size_t)y1 * t.width + x1]};
  for (int c = 0; c < 4; ++c) {
    R k[4];
    for (int j = 0; j < 4; ++j) k[j] = (R)((q[j] >> (8 * c)) & 0xFFu) / R(255);
    out[c] = (k[0] * (R(1) - ax) + k[1] * ax) * (R(1) - ay) + (k[2] * (R(1) - ax) + k[3] * ax) * ay;
  }
}

// The same filtering on an RGB float image (the sky panorama).
template <typename R>
void pano_sample(const float* rgb, uint32_t w, uint32_t h, R u, R v, R out[3]) {
  R x = u * (R)w - R(0.5), y = v * (R)h - R(0.5);
  if (!std::isfinite(x)) x = R(-0.5);
  if (!std::isfinite(y)) y = R(-0.5);
  const R fx = std::floor(x), fy = std::floor(y);
  const R ax = x - fx, ay = y - fy;
  auto wrap = [](double i, uint32_t n) {
    double m = std::fmod(i, (double)n);
    if (m < 0) m += n;
    return std::min((uint32_t)m, n - 1);
  };
  const uint32_t x0 = wrap(fx, w), x1 = wrap((double)fx + 1.0, w), y0 = wrap(fy, h), y1 = wrap((double)fy + 1.0, h);
  const float* q[4] = {rgb + 3 * ((size_t)y0 * w + x0), rgb + 3 * ((size_t)y0 * w + x1), rgb + 3 * ((size_t)y1 * w + x0),
                       rgb + 3 * ((size_t)y1 * w + x1)};
  for (int c = 0; c < 3; ++c)
    out[c] = ((R)q[0][c] * (R(1) - ax) + (R)q[1][c] * ax) * (R(1) - ay) + ((R)q[2][c] * (R(1) - ax) + (R)q[3][c] * ax) * ay;
}

// srgb_to_linear (assets/shaders/color.wgsl:9-15), one channel.
template <typename R> R srgb_lin(R c) {
  return c < R(0.04045) ? c / R(12.92) : std::pow((c + R(0.055)) / R(1.055), R(2.4));
}

// The material at a surface point (Material::from_material_descriptor,
// material_pool_bindings.wgsl:126-154): each factor times its texture where the
// slot holds one and the factor is nonzero (:20-92; the colour's test is on
// vec4(color, 1), which always passes); colour and sheen tint through
// srgb_to_linear; luminance = the colour texture's alpha.
template <typename R>
struct MatEval {
  V<R> color, emission, sheen_tint;
  R luminance, metallic, roughness, transmission, sheen, clearcoat, clearcoat_roughness;
};

template <typename R>
MatEval<R> eval_material(const Attribs* at, const Mat& m, R u, R v) {
  MatEval<R> e;
  e.color = f3<R>(m.color);
  e.emission = f3<R>(m.emission);
  e.sheen_tint = f3<R>(m.sheen_tint);
  e.luminance = R(1);
  e.metallic = m.metallic;
  e.roughness = m.roughness;
  e.transmission = m.transmission;
  e.sheen = m.sheen;
  e.clearcoat = m.clearcoat;
  e.clearcoat_roughness = m.clearcoat_roughness;
  if (!at || !at->tex) return e;
  R c[4];
  auto tex = [&](uint32_t id) {
    if (id == NO_TEX || id >= at->n_tex) return false;
    tex_sample<R>(at->tex[id], u, v, c);
    return true;
  };
  if (tex(m.color_texture)) {
    e.color = V<R>(e.color.x * srgb_lin(c[0]), e.color.y * srgb_lin(c[1]), e.color.z * srgb_lin(c[2]));
    e.luminance = c[3];
  }
  if (dotp(e.emission, e.emission) > R(0) && tex(m.emission_texture))
    e.emission = V<R>(e.emission.x * c[0], e.emission.y * c[1], e.emission.z * c[2]);
  if ((e.metallic > R(0) || e.roughness > R(0)) && tex(m.mr_tex)) {
    e.metallic *= c[2];
    e.roughness *= c[1];
  }
  if (e.clearcoat > R(0) && tex(m.clearcoat_texture)) e.clearcoat *= c[0];
  if (e.clearcoat_roughness > R(0) && tex(m.cr_tex)) e.clearcoat_roughness *= c[1];
  if (e.transmission > R(0) && tex(m.transmission_texture)) e.transmission *= c[0];
  if (e.sheen > R(0) && tex(m.sheen_texture)) e.sheen *= c[0];
  if (dotp(e.sheen_tint, e.sheen_tint) > R(0) && tex(m.sheen_tint_texture))
    e.sheen_tint = V<R>(e.sheen_tint.x * srgb_lin(c[0]), e.sheen_tint.y * srgb_lin(c[1]),
                        e.sheen_tint.z * srgb_lin(c[2]));
  return e;
}

// DisneyBsdf::from_material (disney_bsdf.wgsl:226-244) of an evaluated material.
template <typename R>
Disney<R> from_eval(const Mat& m, const MatEval<R>& e) {
  Disney<R> d = from_mat<R>(m);
  d.color = e.color;
  d.metallic = e.metallic;
  d.luminance = e.luminance;
  d.specular = std::max((R)m.specular, R(1) - e.roughness);
  d.roughness = e.roughness;
  d.sheen = e.sheen;
  d.sheen_tint = e.sheen_tint;
  d.clearcoat = e.clearcoat;
  d.clearcoat_gloss = R(1) - e.clearcoat_roughness;
  d.transmission = e.transmission;
  return d;
}

// Texture coordinates of a hit (vertex-interpolated TEXCOORD_0,
// helpers/inline_path_tracer.wgsl:82); bu, bv weight vertices 1 and 2.
template <typename R>
void hit_uv(const Attribs* at, uint32_t prim, R bu, R bv, R& u, R& v) {
  u = v = R(0);
  if (!at || !at->uv) return;
  const float* t = at->uv + 6 * (size_t)prim;
  const R b0 = R(1) - bu - bv;
  u = (R)t[0] * b0 + (R)t[2] * bu + (R)t[4] * bv;
  v = (R)t[1] * b0 + (R)t[3] * bu + (R)t[5] * bv;
}

// The shading and clearcoat normals of a hit before the face flip
// (helpers/inline_path_tracer.wgsl:103-165): the tangent frame interpolated
// from the vertices (VertexPoolBindings::load_tbn, vertex_pool_bindings.wgsl:48-58)
// and normalised; MaterialDescriptor::apply_normal_mapping (:98-110) bends the
// interpolated normal n by the normal texture, blended back towards n by
// normal_scale < 1; the clearcoat lobe takes n without a clearcoat normal
// texture, the mapped normal when both textures are one, else its own map
// (:112-123, no scale).  Without tangents nothing is mapped.
template <typename R>
void mapped_normals(const Attribs* at, const Mat& m, uint32_t prim, R bu, R bv, R u, R v, const V<R>& n,
                    V<R>& nmap, V<R>& ncc) {
  nmap = ncc = n;
  if (!at || !at->tex || !at->tan) return;
  const bool has_n = m.normal_texture != NO_TEX && m.normal_texture < at->n_tex && m.normal_scale > 0.0f;
  const bool has_c = m.clearcoat_normal_texture != NO_TEX && m.clearcoat_normal_texture < at->n_tex;
  if (!has_n && !has_c) return;
  const R b0 = R(1) - bu - bv;
  const float* tt = at->tan + 9 * (size_t)prim;
  const float* tb = at->bit + 9 * (size_t)prim;
  const V<R> T = unit(f3<R>(tt) * b0 + f3<R>(tt + 3) * bu + f3<R>(tt + 6) * bv);
  const V<R> B = unit(f3<R>(tb) * b0 + f3<R>(tb + 3) * bu + f3<R>(tb + 6) * bv);
  R c[4];
  auto bend = [&](uint32_t id) {
    tex_sample<R>(at->tex[id], u, v, c);
    return unit(T * (c[0] * R(2) - R(1)) + B * (c[1] * R(2) - R(1)) + n * (c[2] * R(2) - R(1)));
  };
  if (has_n) {
    nmap = bend(m.normal_texture);
    const R s = m.normal_scale;
    if (s < R(1)) nmap = unit(n * (R(1) - s) + nmap * s);
  }
  if (has_c) ncc = m.clearcoat_normal_texture == m.normal_texture ? nmap : bend(m.clearcoat_normal_texture);
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
V<R> disney_f(const Disney<R>& b, const Frame<R>& fr, const Frame<R>& fc, const V<R>& wo, const V<R>& wi, bool back,
             R& pdf) {
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
    V<R> ol = fc.local(wo), il = fc.local(wi);
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
Smp<R> disney_sample(const Disney<R>& b, const Frame<R>& fr, const Frame<R>& fc, const V<R>& wo, R dist, bool back,
                     R r0, R r1, R r2) {
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
    const Frame<R>& fl = cc ? fc : fr;  // the clearcoat lobe lives in its own frame
    V<R> ol = fl.local(wo);
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
    wi = fl.world(il);
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
    V<R> ol = fc.local(wo), il = fc.local(wi);
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

// Sky::sky without the sun (shared/sky_bindings.wgsl:68-70): the panorama at
// unit_vector_to_panorama_coords(d) = ((atan2(d.z, d.x) + pi) / 2 pi,
// acos(d.y) / pi) (assets/shaders/math.wgsl:101-106) when the scene has one,
// else the analytic gradient.
template <typename R>
V<R> sky_L(const Sky& k, const Attribs* at, const V<R>& d) {
  if (!at || !at->sky_rgb) return sky_rad<R>(k, d);
  const R u = (std::atan2(d.z, d.x) + C<R>::pi) / (R(2) * C<R>::pi);
  const R v = std::acos(cl(d.y, R(-1), R(1))) / C<R>::pi;
  R c[3];
  pano_sample<R>(at->sky_rgb, at->sky_w, at->sky_h, u, v, c);
  return V<R>(c[0], c[1], c[2]);
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
void oracle_set_tally(uint64_t* t4);
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
  const Attribs* at = nullptr;  // uvs, tangent frames, textures (NULL: untextured)
};

struct Cfg {
  uint32_t width, height, row_start, rows, spp, max_bounces, rr_start, seed, frame_idx;
  uint32_t jitter;  // 0: pixel centres (the ReSTIR integrator's camera rays)
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

}  // namespace ocr
