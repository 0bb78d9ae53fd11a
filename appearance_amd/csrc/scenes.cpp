// Procedural scenes and ray sets, generated on the host from a seed.
//
// These stand in for the reference's glTF assets (crates/appearance-model) on
// the configurations the path is measured on.  Every random draw comes from
// the same integer hash chain as the renderer (pcg_hash of a per-item
// hash_combine seed), so a (name, seed) pair always yields the same bytes.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "apt_internal.h"


namespace apt {
namespace {

struct Rng {
  uint32_t s;
  explicit Rng(uint32_t seed, uint32_t stream) : s(hash_combine(seed, stream)) {}
  float uni() { return rng_next(s); }
  float uni_open() {  // (0, 1]
    float u = uni();
    return u > 0.0f ? u : 1.0f / 4294967296.0f;
  }
  float normal() {  // Box-Muller
    float u1 = uni_open(), u2 = uni();
    return std::sqrt(-2.0f * std::log(u1)) * std::cos(TWO_PI * u2);
  }
  v3 sphere() {
    float z = 1.0f - 2.0f * uni();
    float r = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
    float phi = TWO_PI * uni();
    return mk(r * std::cos(phi), r * std::sin(phi), z);
  }
};

void push_tri(apt_scene_data& s, v3 a, v3 b, v3 c, uint32_t mat) {
  float t[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
  s.positions.insert(s.positions.end(), t, t + 9);
  s.mat_ids.push_back(mat);
}

void push_quad(apt_scene_data& s, v3 a, v3 b, v3 c, v3 d, uint32_t mat) {
  push_tri(s, a, b, c, mat);
  push_tri(s, a, c, d, mat);
}

apt_material lambert(float r, float g, float b) {
  apt_material m;
  apt_material_default(&m);
  m.color[0] = r;
  m.color[1] = g;
  m.color[2] = b;
  m.bsdf_model = APT_BSDF_LAMBERT;
  return m;
}

// Column-major camera transform: rotation columns (right, up, back) and the
// translation; the camera looks down its local -Z like the reference's
// Transform (crates/appearance-transform/src/lib.rs:6-8).
void set_camera(apt_camera& c, v3 eye, v3 target, float fov_deg) {
  v3 back = normalize(eye - target);
  v3 up0 = mk(0.0f, 1.0f, 0.0f);
  v3 right = normalize(cross(up0, back));
  v3 up = cross(back, right);
  float m[16] = {right.x, right.y, right.z, 0.0f, up.x, up.y, up.z, 0.0f,
                 back.x,  back.y,  back.z,  0.0f, eye.x, eye.y, eye.z, 1.0f};
  std::memcpy(c.transform, m, sizeof(m));
  c.fov_deg = fov_deg;
  c.near_plane = 0.01f;
  c.far_plane = 1000.0f;
  c._padding = 0;
}

void sky_off(apt_sky& k) { std::memset(&k, 0, sizeof(k)); }

// Box of half extents h centred at c, rotated by ang about +y: 12 triangles.
void push_box(apt_scene_data& s, v3 c, v3 h, float ang, uint32_t mat) {
  float cs = std::cos(ang), sn = std::sin(ang);
  auto P = [&](float sx, float sy, float sz) {
    float x = sx * h.x, z = sz * h.z;
    return mk(c.x + cs * x + sn * z, c.y + sy * h.y, c.z - sn * x + cs * z);
  };
  v3 p000 = P(-1, -1, -1), p100 = P(1, -1, -1), p010 = P(-1, 1, -1), p110 = P(1, 1, -1);
  v3 p001 = P(-1, -1, 1), p101 = P(1, -1, 1), p011 = P(-1, 1, 1), p111 = P(1, 1, 1);
  push_quad(s, p001, p101, p111, p011, mat);  // +z
  push_quad(s, p100, p000, p010, p110, mat);  // -z
  push_quad(s, p101, p100, p110, p111, mat);  // +x
  push_quad(s, p000, p001, p011, p010, mat);  // -x
  push_quad(s, p011, p111, p110, p010, mat);  // +y
  push_quad(s, p000, p100, p101, p001, mat);  // -y
}

// Cornell box: walls x in [-1,1], y in [0,2], z in [-1,1], open towards +z.
void gen_cornell(apt_scene_data& s) {
  s.materials.push_back(lambert(0.73f, 0.73f, 0.73f));  // 0 white
  s.materials.push_back(lambert(0.65f, 0.05f, 0.05f));  // 1 red (left)
  s.materials.push_back(lambert(0.12f, 0.45f, 0.15f));  // 2 green (right)
  apt_material light = lambert(0.0f, 0.0f, 0.0f);
  light.emission[0] = 17.0f;
  light.emission[1] = 12.0f;
  light.emission[2] = 4.0f;
  s.materials.push_back(light);  // 3 light
  const float L = -1.0f, R = 1.0f, B = 0.0f, T = 2.0f, K = -1.0f, F = 1.0f;
  push_quad(s, mk(L, B, F), mk(R, B, F), mk(R, B, K), mk(L, B, K), 0);  // floor
  push_quad(s, mk(L, T, K), mk(R, T, K), mk(R, T, F), mk(L, T, F), 0);  // ceiling
  push_quad(s, mk(L, B, K), mk(R, B, K), mk(R, T, K), mk(L, T, K), 0);  // back
  push_quad(s, mk(L, B, F), mk(L, B, K), mk(L, T, K), mk(L, T, F), 1);  // left
  push_quad(s, mk(R, B, K), mk(R, B, F), mk(R, T, F), mk(R, T, K), 2);  // right
  const float deg = PI / 180.0f;
  push_box(s, mk(0.33f, 0.3f, 0.35f), mk(0.3f, 0.3f, 0.3f), -18.0f * deg, 0);   // short box
  push_box(s, mk(-0.33f, 0.6f, -0.3f), mk(0.3f, 0.6f, 0.3f), 18.0f * deg, 0);   // tall box
  const float ly = 1.99f, lh = 0.25f;
  push_quad(s, mk(-lh, ly, -lh), mk(lh, ly, -lh), mk(lh, ly, lh), mk(-lh, ly, lh), 3);  // light, faces -y
  set_camera(s.camera, mk(0.0f, 1.0f, 3.9f), mk(0.0f, 1.0f, 0.0f), 40.0f);
  sky_off(s.sky);
}

// 1,048,576 triangles: centroids uniform in [-10,10]^3, two edges of
// uniform-sphere direction and LogNormal(-2, 0.5) length.
void gen_bvh_stress(apt_scene_data& s, uint32_t seed, uint32_t n) {
  s.positions.resize((size_t)n * 9);
  s.mat_ids.assign(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    Rng r(seed, i);
    v3 c = mk(r.uni() * 20.0f - 10.0f, r.uni() * 20.0f - 10.0f, r.uni() * 20.0f - 10.0f);
    v3 e1 = r.sphere() * std::exp(-2.0f + 0.5f * r.normal());
    v3 e2 = r.sphere() * std::exp(-2.0f + 0.5f * r.normal());
    v3 v0 = c - (e1 + e2) * (1.0f / 3.0f);
    v3 v1 = v0 + e1, v2 = v0 + e2;
    float t[9] = {v0.x, v0.y, v0.z, v1.x, v1.y, v1.z, v2.x, v2.y, v2.z};
    std::memcpy(&s.positions[(size_t)i * 9], t, sizeof(t));
  }
  s.materials.push_back(lambert(0.5f, 0.5f, 0.5f));
  set_camera(s.camera, mk(0.0f, 0.0f, -25.0f), mk(0.0f, 0.0f, 0.0f), 40.0f);
  sky_off(s.sky);
}

void default_sky(apt_sky& k) {
  std::memset(&k, 0, sizeof(k));
  v3 sun = normalize(mk(-0.4f, -1.0f, -0.3f));  // direction the light travels
  k.sun_direction[0] = sun.x;
  k.sun_direction[1] = sun.y;
  k.sun_direction[2] = sun.z;
  k.sun_size = 0.5f;  // the reference's sun disk spans size*0.1 radians
  k.sun_color[0] = 1.0f;
  k.sun_color[1] = 0.95f;
  k.sun_color[2] = 0.85f;
  k.sun_intensity = 100.0f;  // SunInfo default intensity (sky.rs)
  k.sun_enabled = 1.0f;
  k.zenith[0] = 0.25f;
  k.zenith[1] = 0.45f;
  k.zenith[2] = 0.9f;
  k.horizon[0] = 0.9f;
  k.horizon[1] = 0.9f;
  k.horizon[2] = 1.0f;
  k.ground[0] = 0.2f;
  k.ground[1] = 0.18f;
  k.ground[2] = 0.15f;
}

// 256 UV spheres (32 x 32 x 2 triangles) on a 16 x 16 grid plus a ground quad.
void gen_disney(apt_scene_data& s, uint32_t seed) {
  const int N = 16, SEG = 32;
  const float spacing = 1.2f;
  apt_material ground;
  apt_material_default(&ground);
  ground.color[0] = ground.color[1] = ground.color[2] = 0.5f;
  ground.roughness = 0.8f;
  s.materials.push_back(ground);
  for (int gz = 0; gz < N; ++gz) {
    for (int gx = 0; gx < N; ++gx) {
      const uint32_t id = (uint32_t)(gz * N + gx);
      Rng r(seed, id);
      float rad = 0.3f + 0.2f * r.uni();
      float slack = 0.5f * spacing - rad;  // keeps neighbours apart
      float cx = ((float)gx - 0.5f * (N - 1)) * spacing + (2.0f * r.uni() - 1.0f) * slack;
      float cz = ((float)gz - 0.5f * (N - 1)) * spacing + (2.0f * r.uni() - 1.0f) * slack;
      v3 c = mk(cx, rad, cz);
      apt_material m;
      apt_material_default(&m);
      m.color[0] = r.uni();
      m.color[1] = r.uni();
      m.color[2] = r.uni();
      m.metallic = r.uni();
      m.roughness = 0.05f + 0.95f * r.uni();
      m.specular = r.uni();
      m.sheen = r.uni();
      m.clearcoat = r.uni();
      m.transmission = r.uni() < 0.2f ? 1.0f : 0.0f;
      m.eta = 1.0f / 1.5f;
      m.sheen_tint[0] = m.sheen_tint[1] = m.sheen_tint[2] = 1.0f;
      m.specular_tint[0] = m.specular_tint[1] = m.specular_tint[2] = 1.0f;
      const uint32_t mat = (uint32_t)s.materials.size();
      s.materials.push_back(m);
      auto P = [&](int i, int j) {  // i: longitude, j: latitude
        float th = PI * (float)j / (float)SEG;
        float ph = TWO_PI * (float)i / (float)SEG;
        return mk(std::sin(th) * std::cos(ph), std::cos(th), std::sin(th) * std::sin(ph));
      };
      for (int j = 0; j < SEG; ++j) {
        for (int i = 0; i < SEG; ++i) {
          v3 n00 = P(i, j), n10 = P(i + 1, j), n01 = P(i, j + 1), n11 = P(i + 1, j + 1);
          v3 q[4] = {n00, n01, n11, n10};
          const int tri[2][3] = {{0, 1, 2}, {0, 2, 3}};
          for (int t = 0; t < 2; ++t) {
            v3 a = q[tri[t][0]], b = q[tri[t][1]], cc = q[tri[t][2]];
            push_tri(s, c + a * rad, c + b * rad, c + cc * rad, mat);
            float nn[9] = {a.x, a.y, a.z, b.x, b.y, b.z, cc.x, cc.y, cc.z};
            s.normals.insert(s.normals.end(), nn, nn + 9);
          }
        }
      }
    }
  }
  const float G = 0.5f * N * spacing + 2.0f;
  push_quad(s, mk(-G, 0.0f, G), mk(G, 0.0f, G), mk(G, 0.0f, -G), mk(-G, 0.0f, -G), 0);
  float up[9] = {0, 1, 0, 0, 1, 0, 0, 1, 0};
  for (int t = 0; t < 2; ++t) s.normals.insert(s.normals.end(), up, up + 9);
  set_camera(s.camera, mk(0.0f, 7.5f, 14.0f), mk(0.0f, 0.0f, 0.5f), 40.0f);
  default_sky(s.sky);
}

// "disney" plus 4096 small emissive triangles above the spheres.
void gen_restir(apt_scene_data& s, uint32_t seed) {
  gen_disney(s, seed);
  const uint32_t first_mat = (uint32_t)s.materials.size();
  for (uint32_t i = 0; i < 4096; ++i) {
    Rng r(seed ^ 0x5bd1e995u, i);
    v3 c = mk((2.0f * r.uni() - 1.0f) * 9.0f, 1.5f + 1.5f * r.uni(), (2.0f * r.uni() - 1.0f) * 9.0f);
    v3 e1 = r.sphere() * 0.08f, e2 = r.sphere() * 0.08f;
    apt_material m = lambert(0.0f, 0.0f, 0.0f);
    m.bsdf_model = APT_BSDF_DISNEY;
    m.emission[0] = 1.0f + 49.0f * r.uni();
    m.emission[1] = 1.0f + 49.0f * r.uni();
    m.emission[2] = 1.0f + 49.0f * r.uni();
    s.materials.push_back(m);
    push_tri(s, c, c + e1, c + e2, first_mat + i);
    v3 n = normalize(cross(e1, e2));
    float nn[9] = {n.x, n.y, n.z, n.x, n.y, n.z, n.x, n.y, n.z};
    s.normals.insert(s.normals.end(), nn, nn + 9);
  }
}

}  // namespace
}  // namespace apt

using namespace apt;

extern "C" {

void apt_material_default(apt_material* m) {
  if (!m) return;
  std::memset(m, 0, sizeof(*m));
  m->color[0] = m->color[1] = m->color[2] = 1.0f;
  m->color_texture = APT_INVALID_TEXTURE;
  m->metallic = 0.0f;
  m->roughness = 1.0f;
  m->metallic_roughness_texture = APT_INVALID_TEXTURE;
  m->normal_scale = 1.0f;
  m->normal_texture = APT_INVALID_TEXTURE;
  m->emission_texture = APT_INVALID_TEXTURE;
  m->eta = 1.0f / 1.5f;
  m->sheen_texture = APT_INVALID_TEXTURE;
  m->clearcoat_texture = APT_INVALID_TEXTURE;
  m->clearcoat_roughness_texture = APT_INVALID_TEXTURE;
  m->sheen_tint_texture = APT_INVALID_TEXTURE;
  m->clearcoat_normal_texture = APT_INVALID_TEXTURE;
  m->transmission_texture = APT_INVALID_TEXTURE;
  m->bsdf_model = APT_BSDF_DISNEY;
}

apt_status apt_scene_generate(const char* name, uint32_t seed, apt_scene_data** out) {
  clear_error();
  APT_REQUIRE(name && out, APT_ERR_INVALID_ARG, "apt_scene_generate: null argument");
  std::string nm(name);
  apt_scene_data* s = new apt_scene_data();
  if (nm == "cornell") {
    gen_cornell(*s);
  } else if (nm == "bvh_stress") {
    gen_bvh_stress(*s, seed, 1u << 20);
  } else if (nm.rfind("bvh_stress:", 0) == 0) {
    unsigned long k = std::strtoul(nm.c_str() + 11, nullptr, 10);
    if (k == 0 || k > (1ul << 26)) {
      delete s;
      set_error("apt_scene_generate: bad triangle count in '%s'", name);
      return APT_ERR_INVALID_ARG;
    }
    gen_bvh_stress(*s, seed, (uint32_t)k);
  } else if (nm == "disney") {
    gen_disney(*s, seed);
  } else if (nm == "restir") {
    gen_restir(*s, seed);
  } else {
    delete s;
    set_error("apt_scene_generate: unknown scene '%s'", name);
    return APT_ERR_NOT_FOUND;
  }
  *out = s;
  return APT_OK;
}

apt_status apt_scene_data_view(const apt_scene_data* s, apt_mesh_desc* mesh, apt_camera* camera,
                               apt_sky* sky) {
  APT_REQUIRE(s, APT_ERR_INVALID_ARG, "apt_scene_data_view: null scene");
  if (mesh) {
    mesh->positions = s->positions.data();
    mesh->normals = s->normals.empty() ? nullptr : s->normals.data();
    mesh->material_ids = s->mat_ids.data();
    mesh->n_tris = (uint32_t)(s->positions.size() / 9);
    mesh->n_materials = (uint32_t)s->materials.size();
    mesh->materials = s->materials.data();
  }
  if (camera) *camera = s->camera;
  if (sky) *sky = s->sky;
  return APT_OK;
}

apt_status apt_scene_data_attribs(const apt_scene_data* s, apt_mesh_attribs* out) {
  APT_REQUIRE(s && out, APT_ERR_INVALID_ARG, "apt_scene_data_attribs: null argument");
  std::memset(out, 0, sizeof(*out));
  out->uvs = s->uvs.empty() ? nullptr : s->uvs.data();
  out->tangents = s->tangents.empty() ? nullptr : s->tangents.data();
  out->bitangents = s->bitangents.empty() ? nullptr : s->bitangents.data();
  std::vector<apt_texture>& v = s->texture_views;
  v.resize(s->textures.size());
  for (size_t k = 0; k < v.size(); ++k) v[k] = apt_texture{s->textures[k].width, s->textures[k].height,
                                                           s->textures[k].rgba.data()};
  out->textures = v.empty() ? nullptr : v.data();
  out->n_textures = (uint32_t)v.size();
  return APT_OK;
}

void apt_scene_data_destroy(apt_scene_data* s) { delete s; }

apt_status apt_rays_bvh_stress(uint32_t seed, uint32_t side, uint32_t n_incoherent, apt_ray* out) {
  clear_error();
  APT_REQUIRE(out, APT_ERR_INVALID_ARG, "apt_rays_bvh_stress: out is NULL");
  const float tan_half = std::tan(0.5f * 40.0f * PI / 180.0f);
  const uint64_t nc = (uint64_t)side * side;
  for (uint32_t y = 0; y < side; ++y) {
    for (uint32_t x = 0; x < side; ++x) {
      float ux = ((float)x + 0.5f) / (float)side * 2.0f - 1.0f;
      float uy = -(((float)y + 0.5f) / (float)side * 2.0f - 1.0f);
      v3 d = normalize(mk(ux * tan_half, uy * tan_half, 1.0f));
      apt_ray& r = out[(uint64_t)y * side + x];
      r.org[0] = 0.0f;
      r.org[1] = 0.0f;
      r.org[2] = -25.0f;
      r.tmin = 0.0f;
      r.dir[0] = d.x;
      r.dir[1] = d.y;
      r.dir[2] = d.z;
      r.tmax = 1e30f;
    }
  }
  for (uint32_t i = 0; i < n_incoherent; ++i) {
    Rng g(seed, i);
    v3 o = mk(g.uni() * 20.0f - 10.0f, g.uni() * 20.0f - 10.0f, g.uni() * 20.0f - 10.0f);
    v3 d = g.sphere();
    apt_ray& r = out[nc + i];
    r.org[0] = o.x;
    r.org[1] = o.y;
    r.org[2] = o.z;
    r.tmin = 0.0f;
    r.dir[0] = d.x;
    r.dir[1] = d.y;
    r.dir[2] = d.z;
    r.tmax = 1e30f;
  }
  return APT_OK;
}

}  // extern "C"
