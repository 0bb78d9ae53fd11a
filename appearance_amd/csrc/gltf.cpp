// glTF 2.0 binary (.glb) and JSON (.gltf with external or embedded buffers)
// ingestion into a triangle soup with materials: what the reference's
// asset loader produces for a Model (crates/appearance-model/src/asset.rs:19-380),
// flattened through the node hierarchy so SpawnModel can place it.
//
// Per primitive (triangles only, indices required, as asset.rs:140-160):
// POSITION, NORMAL (generated per vertex from the area-independent sum of
// face normals when absent, mesh.rs:58-79), TEXCOORD_0, TANGENT (generated
// per vertex from positions, normals and uvs when absent, mesh.rs:82-160,
// once some material has a normal map).  Materials follow asset.rs:199-339:
// base colour, metallic, roughness, emission times
// KHR_materials_emissive_strength, KHR_materials_volume absorption
// (1 - attenuation colour) / attenuation distance, KHR_materials_transmission,
// eta = 1 / KHR_materials_ior (1.5), KHR_materials_specular,
// KHR_materials_clearcoat, KHR_materials_sheen (its roughness into `sheen`,
// its colour into `sheen_tint`), alpha cutoff 0.5, and their textures (base
// colour, metallic-roughness, normal with its scale, emission, transmission,
// clearcoat, clearcoat roughness, clearcoat normal, sheen roughness, sheen
// colour), each image decoded once (PNG; other formats leave the slot empty).
// Node transforms are matrix or translation * rotation * scale, composed from
// the default scene's roots down; normals, tangents and bitangents (cross(n,
// t) * w in the mesh's space) go through the inverse transpose and are
// renormalised.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "apt_internal.h"

namespace apt {
namespace {

// ---- a small JSON reader ------------------------------------------------------
struct Json {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  double num = 0.0;
  bool b = false;
  std::string str;
  std::vector<Json> arr;
  std::map<std::string, Json> obj;

  const Json* get(const char* k) const {
    if (kind != OBJ) return nullptr;
    auto it = obj.find(k);
    return it == obj.end() ? nullptr : &it->second;
  }
  double number(const char* k, double dflt) const {
    const Json* j = get(k);
    return j && j->kind == NUM ? j->num : dflt;
  }
  const Json* at(size_t i) const { return kind == ARR && i < arr.size() ? &arr[i] : nullptr; }
};

struct JsonParser {
  const char* p;
  const char* end;
  std::string err;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  bool fail(const char* m) {
    if (err.empty()) err = m;
    return false;
  }
  bool lit(const char* s) {
    const size_t n = std::strlen(s);
    if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return false;
    p += n;
    return true;
  }
  bool string(std::string& out) {
    if (p >= end || *p != '"') return fail("expected a string");
    ++p;
    while (p < end && *p != '"') {
      if (*p == '\\') {
        if (++p >= end) return fail("bad escape");
        switch (*p) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {  // names and URIs: keep the code point as '?' when not ASCII
            if (end - p < 5) return fail("bad \\u escape");
            const unsigned v = (unsigned)std::strtoul(std::string(p + 1, p + 5).c_str(), nullptr, 16);
            out += v < 128 ? (char)v : '?';
            p += 4;
            break;
          }
          default: out += *p;
        }
        ++p;
      } else {
        out += *p++;
      }
    }
    if (p >= end) return fail("unterminated string");
    ++p;
    return true;
  }
  bool value(Json& v, int depth) {
    if (depth > 64) return fail("JSON nested too deeply");
    ws();
    if (p >= end) return fail("unexpected end of JSON");
    const char c = *p;
    if (c == '{') {
      v.kind = Json::OBJ;
      ++p;
      ws();
      if (p < end && *p == '}') {
        ++p;
        return true;
      }
      while (true) {
        ws();
        std::string k;
        if (!string(k)) return false;
        ws();
        if (p >= end || *p != ':') return fail("expected ':'");
        ++p;
        if (!value(v.obj[k], depth + 1)) return false;
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == '}') {
          ++p;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      v.kind = Json::ARR;
      ++p;
      ws();
      if (p < end && *p == ']') {
        ++p;
        return true;
      }
      while (true) {
        v.arr.emplace_back();
        if (!value(v.arr.back(), depth + 1)) return false;
        ws();
        if (p < end && *p == ',') {
          ++p;
          continue;
        }
        if (p < end && *p == ']') {
          ++p;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      v.kind = Json::STR;
      return string(v.str);
    }
    if (lit("true")) {
      v.kind = Json::BOOL;
      v.b = true;
      return true;
    }
    if (lit("false")) {
      v.kind = Json::BOOL;
      return true;
    }
    if (lit("null")) return true;
    char* q = nullptr;
    const std::string tok(p, (size_t)std::min<ptrdiff_t>(end - p, 64));
    v.num = std::strtod(tok.c_str(), &q);
    if (q == tok.c_str()) return fail("bad JSON value");
    v.kind = Json::NUM;
    p += q - tok.c_str();
    return true;
  }
};

// ---- matrices (column-major 4x4, glTF's convention) ---------------------------
struct M4 {
  double m[16];
};

M4 identity() {
  M4 r{};
  r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
  return r;
}

M4 mul(const M4& a, const M4& b) {
  M4 r{};
  for (int c = 0; c < 4; ++c)
    for (int rr = 0; rr < 4; ++rr) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += a.m[k * 4 + rr] * b.m[c * 4 + k];
      r.m[c * 4 + rr] = s;
    }
  return r;
}

M4 node_local(const Json& n) {
  if (const Json* mat = n.get("matrix")) {
    M4 r = identity();
    for (int i = 0; i < 16; ++i)
      if (const Json* e = mat->at(i)) r.m[i] = e->num;
    return r;
  }
  double t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
  if (const Json* a = n.get("translation"))
    for (int i = 0; i < 3; ++i) t[i] = a->at(i) ? a->at(i)->num : 0.0;
  if (const Json* a = n.get("rotation"))
    for (int i = 0; i < 4; ++i) q[i] = a->at(i) ? a->at(i)->num : q[i];
  if (const Json* a = n.get("scale"))
    for (int i = 0; i < 3; ++i) s[i] = a->at(i) ? a->at(i)->num : 1.0;
  const double x = q[0], y = q[1], z = q[2], w = q[3];
  const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w),   // column 0
                       2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w),   // column 1
                       2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)};  // column 2
  M4 r = identity();
  for (int c = 0; c < 3; ++c)
    for (int rr = 0; rr < 3; ++rr) r.m[c * 4 + rr] = R[c * 3 + rr] * s[c];
  r.m[12] = t[0];
  r.m[13] = t[1];
  r.m[14] = t[2];
  return r;
}

// ---- loader ---------------------------------------------------------------------
struct Loader {
  Json doc;
  std::vector<std::vector<uint8_t>> buffers;
  std::string dir;  // for external buffer URIs
  std::string err;
  apt_scene_data* out = nullptr;
  std::map<size_t, uint32_t> image_tex;  // image index -> scene texture (APT_INVALID_TEXTURE: not decoded)
  bool need_tangents = false;            // some material has a normal or clearcoat normal map
  bool any_uv = false;

  // Bytes of image `idx`: a buffer view, a data URI or an external file.
  bool image_bytes(size_t idx, std::vector<uint8_t>& bytes) {
    const Json* im = doc.get("images") ? doc.get("images")->at(idx) : nullptr;
    if (!im) return false;
    if (const Json* bvj = im->get("bufferView")) {
      const Json* bv = doc.get("bufferViews") ? doc.get("bufferViews")->at((size_t)bvj->num) : nullptr;
      if (!bv) return false;
      const size_t buf = (size_t)bv->number("buffer", 0), off = (size_t)bv->number("byteOffset", 0);
      const size_t len = (size_t)bv->number("byteLength", 0);
      if (buf >= buffers.size() || off + len > buffers[buf].size()) return false;
      bytes.assign(buffers[buf].begin() + (long)off, buffers[buf].begin() + (long)(off + len));
      return true;
    }
    const Json* uri = im->get("uri");
    if (!uri) return false;
    const std::string& u = uri->str;
    if (u.rfind("data:", 0) == 0) {
      const size_t comma = u.find(',');
      if (comma == std::string::npos || u.find(";base64", 0) > comma) return false;
      bytes = base64(u.substr(comma + 1));
      return true;
    }
    std::ifstream f(dir + u, std::ios::binary);
    if (!f) return false;
    bytes.assign((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return true;
  }

  // The scene texture of a textureInfo object ({"index": texture}), decoding
  // its image on first use; APT_INVALID_TEXTURE when absent or undecodable.
  uint32_t texture(const Json* info) {
    if (!info) return APT_INVALID_TEXTURE;
    const Json* ti = info->get("index");
    const Json* tex = (ti && doc.get("textures")) ? doc.get("textures")->at((size_t)ti->num) : nullptr;
    const Json* src = tex ? tex->get("source") : nullptr;
    if (!src) return APT_INVALID_TEXTURE;
    const size_t img = (size_t)src->num;
    auto it = image_tex.find(img);
    if (it != image_tex.end()) return it->second;
    uint32_t id = APT_INVALID_TEXTURE;
    std::vector<uint8_t> bytes;
    apt_scene_texture t;
    std::string e;
    if (image_bytes(img, bytes) && (png_decode(bytes.data(), bytes.size(), t.width, t.height, t.rgba, e) ||
                                    jpeg_decode(bytes.data(), bytes.size(), t.width, t.height, t.rgba, e))) {
      id = (uint32_t)out->textures.size();
      out->textures.push_back(std::move(t));
    }
    image_tex[img] = id;
    return id;
  }

  bool fail(const std::string& m) {
    if (err.empty()) err = m;
    return false;
  }

  bool load_buffers(const std::vector<uint8_t>* glb_bin) {
    const Json* bufs = doc.get("buffers");
    if (!bufs) return true;
    for (size_t i = 0; i < bufs->arr.size(); ++i) {
      const Json& b = bufs->arr[i];
      const Json* uri = b.get("uri");
      if (!uri) {
        if (i != 0 || !glb_bin) return fail("buffer without a URI outside a .glb");
        buffers.push_back(*glb_bin);
        continue;
      }
      const std::string& u = uri->str;
      const std::string pre = "data:";
      if (u.rfind(pre, 0) == 0) {
        const size_t comma = u.find(',');
        if (comma == std::string::npos || u.find(";base64", 0) > comma) return fail("unsupported data URI");
        buffers.push_back(base64(u.substr(comma + 1)));
        continue;
      }
      std::ifstream f(dir + u, std::ios::binary);
      if (!f) return fail("cannot open buffer " + dir + u);
      buffers.emplace_back((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    }
    return true;
  }

  static std::vector<uint8_t> base64(const std::string& s) {
    std::vector<uint8_t> o;
    uint32_t acc = 0;
    int bits = 0;
    for (char c : s) {
      int v;
      if (c >= 'A' && c <= 'Z') v = c - 'A';
      else if (c >= 'a' && c <= 'z') v = c - 'a' + 26;
      else if (c >= '0' && c <= '9') v = c - '0' + 52;
      else if (c == '+') v = 62;
      else if (c == '/') v = 63;
      else continue;
      acc = (acc << 6) | (uint32_t)v;
      bits += 6;
      if (bits >= 8) {
        bits -= 8;
        o.push_back((uint8_t)(acc >> bits));
      }
    }
    return o;
  }

  // Reads accessor `idx` as `comps` floats per element (ints converted; the
  // normalized flag scales them to [0, 1] / [-1, 1]).
  bool read_floats(int idx, int comps, std::vector<float>& v) {
    const Json* acc = doc.get("accessors") ? doc.get("accessors")->at((size_t)idx) : nullptr;
    if (!acc) return fail("missing accessor");
    if (acc->get("sparse")) return fail("sparse accessors are not supported");
    const size_t count = (size_t)acc->number("count", 0);
    const int ctype = (int)acc->number("componentType", 0);
    const bool norm = acc->get("normalized") && acc->get("normalized")->b;
    const int csize = ctype == 5126 || ctype == 5125 ? 4 : ctype == 5123 || ctype == 5122 ? 2 : 1;
    v.assign(count * comps, 0.0f);
    const Json* bvj = acc->get("bufferView");
    if (!bvj) return true;  // all zeros
    const Json* bv = doc.get("bufferViews") ? doc.get("bufferViews")->at((size_t)bvj->num) : nullptr;
    if (!bv) return fail("missing buffer view");
    const size_t buf = (size_t)bv->number("buffer", 0);
    if (buf >= buffers.size()) return fail("missing buffer");
    const size_t off = (size_t)bv->number("byteOffset", 0) + (size_t)acc->number("byteOffset", 0);
    const size_t stride = bv->get("byteStride") ? (size_t)bv->number("byteStride", 0) : (size_t)csize * comps;
    const std::vector<uint8_t>& B = buffers[buf];
    if (count && off + (count - 1) * stride + (size_t)csize * comps > B.size()) return fail("accessor out of range");
    for (size_t i = 0; i < count; ++i)
      for (int c = 0; c < comps; ++c) {
        const uint8_t* p = B.data() + off + i * stride + (size_t)c * csize;
        float x = 0.0f;
        switch (ctype) {
          case 5126: std::memcpy(&x, p, 4); break;
          case 5125: { uint32_t u; std::memcpy(&u, p, 4); x = (float)u; break; }
          case 5123: { uint16_t u; std::memcpy(&u, p, 2); x = norm ? u / 65535.0f : (float)u; break; }
          case 5122: { int16_t u; std::memcpy(&u, p, 2); x = norm ? std::max(u / 32767.0f, -1.0f) : (float)u; break; }
          case 5121: x = norm ? p[0] / 255.0f : (float)p[0]; break;
          case 5120: { int8_t u = (int8_t)p[0]; x = norm ? std::max(u / 127.0f, -1.0f) : (float)u; break; }
          default: return fail("unsupported component type");
        }
        v[i * comps + c] = x;
      }
    return true;
  }

  bool read_indices(int idx, std::vector<uint32_t>& out_idx) {
    const Json* acc = doc.get("accessors") ? doc.get("accessors")->at((size_t)idx) : nullptr;
    if (!acc) return fail("missing index accessor");
    const int ctype = (int)acc->number("componentType", 0);
    if (ctype != 5125 && ctype != 5123 && ctype != 5121) return fail("bad index component type");
    std::vector<float> f;  // exact for indices below 2^24
    const size_t count = (size_t)acc->number("count", 0);
    if (count >= (1u << 24) && ctype == 5125) return fail("index buffers above 2^24 entries are not supported");
    if (!read_floats(idx, 1, f)) return false;
    out_idx.resize(count);
    for (size_t i = 0; i < count; ++i) out_idx[i] = (uint32_t)f[i];
    return true;
  }

  apt_material material(const Json* m) {
    apt_material r;
    apt_material_default(&r);
    r.roughness = 1.0f;  // glTF's roughnessFactor default
    r.specular_tint[0] = r.specular_tint[1] = r.specular_tint[2] = 1.0f;  // Material::default (material.rs)
    r.alpha_cutoff = 0.5f;
    if (!m) {  // glTF's default material: metallic 1, roughness 1
      r.metallic = 1.0f;
      return r;
    }
    if (const Json* pbr = m->get("pbrMetallicRoughness")) {
      if (const Json* c = pbr->get("baseColorFactor"))
        for (int i = 0; i < 3; ++i) r.color[i] = (float)(c->at(i) ? c->at(i)->num : 1.0);
      r.metallic = (float)pbr->number("metallicFactor", 1.0);
      r.roughness = (float)pbr->number("roughnessFactor", 1.0);
      r.color_texture = texture(pbr->get("baseColorTexture"));
      r.metallic_roughness_texture = texture(pbr->get("metallicRoughnessTexture"));
    } else {
      r.metallic = 1.0f;  // the glTF default of an absent pbrMetallicRoughness
    }
    float strength = 1.0f;
    const Json* ext = m->get("extensions");
    if (ext)
      if (const Json* es = ext->get("KHR_materials_emissive_strength"))
        strength = (float)es->number("emissiveStrength", 1.0);
    if (const Json* e = m->get("emissiveFactor"))
      for (int i = 0; i < 3; ++i) r.emission[i] = (float)(e->at(i) ? e->at(i)->num : 0.0) * strength;
    r.emission_texture = texture(m->get("emissiveTexture"));
    if (const Json* nt = m->get("normalTexture")) {
      r.normal_texture = texture(nt);
      r.normal_scale = (float)nt->number("scale", 1.0);
    }
    double ior = 1.5;
    if (ext) {
      if (const Json* v = ext->get("KHR_materials_volume")) {
        const double dist = v->number("attenuationDistance", INFINITY);
        const Json* col = v->get("attenuationColor");
        for (int i = 0; i < 3; ++i) {
          const double c = col && col->at(i) ? col->at(i)->num : 1.0;
          r.absorption[i] = (float)((1.0 - c) / dist);
        }
      }
      if (const Json* t = ext->get("KHR_materials_transmission")) {
        r.transmission = (float)t->number("transmissionFactor", 0.0);
        r.transmission_texture = texture(t->get("transmissionTexture"));
      }
      if (const Json* i = ext->get("KHR_materials_ior")) ior = i->number("ior", 1.5);
      if (const Json* sp = ext->get("KHR_materials_specular")) {
        r.specular = (float)sp->number("specularFactor", 1.0);
        if (const Json* c = sp->get("specularColorFactor"))
          for (int k = 0; k < 3; ++k) r.specular_tint[k] = (float)(c->at(k) ? c->at(k)->num : 1.0);
        else
          r.specular_tint[0] = r.specular_tint[1] = r.specular_tint[2] = 1.0f;
      }
      if (const Json* cc = ext->get("KHR_materials_clearcoat")) {
        r.clearcoat = (float)cc->number("clearcoatFactor", 0.0);
        r.clearcoat_roughness = (float)cc->number("clearcoatRoughnessFactor", 0.0);
        r.clearcoat_texture = texture(cc->get("clearcoatTexture"));
        r.clearcoat_roughness_texture = texture(cc->get("clearcoatRoughnessTexture"));
        r.clearcoat_normal_texture = texture(cc->get("clearcoatNormalTexture"));
      }
      if (const Json* sh = ext->get("KHR_materials_sheen")) {
        r.sheen = (float)sh->number("sheenRoughnessFactor", 0.0);
        if (const Json* c = sh->get("sheenColorFactor"))
          for (int k = 0; k < 3; ++k) r.sheen_tint[k] = (float)(c->at(k) ? c->at(k)->num : 0.0);
        r.sheen_texture = texture(sh->get("sheenRoughnessTexture"));
        r.sheen_tint_texture = texture(sh->get("sheenColorTexture"));
      }
    }
    r.eta = (float)(1.0 / ior);
    r.alpha_cutoff = (float)m->number("alphaCutoff", 0.5);
    return r;
  }

  bool mesh(const Json& me, const M4& world) {
    const Json* prims = me.get("primitives");
    if (!prims) return true;
    // normals through the inverse transpose of the 3x3 part
    const double* a = world.m;
    const double c00 = a[5] * a[10] - a[9] * a[6], c01 = a[9] * a[2] - a[1] * a[10], c02 = a[1] * a[6] - a[5] * a[2];
    const double c10 = a[8] * a[6] - a[4] * a[10], c11 = a[0] * a[10] - a[8] * a[2], c12 = a[4] * a[2] - a[0] * a[6];
    const double c20 = a[4] * a[9] - a[8] * a[5], c21 = a[8] * a[1] - a[0] * a[9], c22 = a[0] * a[5] - a[4] * a[1];
    // rows of inverse-transpose = cofactors (up to the determinant, which renormalisation removes; its sign kept)
    const double det = a[0] * c00 + a[4] * c01 + a[8] * c02;
    const double sg = det < 0.0 ? -1.0 : 1.0;
    for (const Json& p : prims->arr) {
      if ((int)p.number("mode", 4) != 4) return fail("only triangle primitives are supported");
      const Json* attrs = p.get("attributes");
      const Json* pos = attrs ? attrs->get("POSITION") : nullptr;
      if (!pos) return fail("a primitive without POSITION");
      const Json* ind = p.get("indices");
      if (!ind) return fail("a primitive without indices (the reference requires them)");
      std::vector<float> P, N, UV, TG;
      std::vector<uint32_t> I;
      if (!read_floats((int)pos->num, 3, P) || !read_indices((int)ind->num, I)) return false;
      if (const Json* n = attrs->get("NORMAL"))
        if (!read_floats((int)n->num, 3, N)) return false;
      if (const Json* t = attrs->get("TEXCOORD_0"))
        if (!read_floats((int)t->num, 2, UV)) return false;
      if (const Json* t = attrs->get("TANGENT"))
        if (need_tangents && !read_floats((int)t->num, 4, TG)) return false;
      const size_t nv = P.size() / 3;
      for (uint32_t i : I)
        if (i >= nv) return fail("index out of range");
      if (!UV.empty() && UV.size() != nv * 2) return fail("TEXCOORD_0 count differs from POSITION");
      if (!TG.empty() && TG.size() != nv * 4) return fail("TANGENT count differs from POSITION");
      any_uv = any_uv || !UV.empty();
      if (N.empty()) {  // mesh.rs:58-79: sum of unit face normals per vertex
        std::vector<double> acc(nv * 3, 0.0);
        for (size_t t = 0; t + 2 < I.size(); t += 3) {
          const float* p0 = &P[3 * I[t]];
          const float* p1 = &P[3 * I[t + 1]];
          const float* p2 = &P[3 * I[t + 2]];
          const double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
          const double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
          double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
          const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
          if (l > 0.0)
            for (int k = 0; k < 3; ++k)
              for (int j = 0; j < 3; ++j) acc[3 * I[t + j] + k] += n[k] / l;
        }
        N.resize(nv * 3);
        for (size_t v = 0; v < nv; ++v) {
          const double l = std::sqrt(acc[3 * v] * acc[3 * v] + acc[3 * v + 1] * acc[3 * v + 1] +
                                     acc[3 * v + 2] * acc[3 * v + 2]);
          for (int k = 0; k < 3; ++k) N[3 * v + k] = l > 0.0 ? (float)(acc[3 * v + k] / l) : 0.0f;
        }
      }
      if (need_tangents && TG.empty()) {
        std::vector<float> uv0 = UV;
        if (uv0.empty()) uv0.assign(nv * 2, 0.0f);
        generate_tangents(P, N, uv0, I, TG);
      }
      const uint32_t mat = (uint32_t)p.number("material", 0);
      if (mat >= out->materials.size()) return fail("material index out of range");
      // a vector through the inverse transpose, renormalised
      const auto itx = [&](const double* n, std::vector<float>& dst) {
        const double w[3] = {sg * (c00 * n[0] + c01 * n[1] + c02 * n[2]), sg * (c10 * n[0] + c11 * n[1] + c12 * n[2]),
                             sg * (c20 * n[0] + c21 * n[1] + c22 * n[2])};
        const double l = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        for (int r = 0; r < 3; ++r) dst.push_back(l > 0.0 ? (float)(w[r] / l) : 0.0f);
      };
      for (size_t t = 0; t + 2 < I.size(); t += 3) {
        for (int j = 0; j < 3; ++j) {
          const uint32_t vi = I[t + j];
          const float* v = &P[3 * vi];
          for (int r = 0; r < 3; ++r)
            out->positions.push_back(
                (float)(a[r] * v[0] + a[4 + r] * v[1] + a[8 + r] * v[2] + a[12 + r]));
          const float* n = &N[3 * vi];
          const double nd[3] = {n[0], n[1], n[2]};
          itx(nd, out->normals);
          out->uvs.push_back(UV.empty() ? 0.0f : UV[2 * vi]);
          out->uvs.push_back(UV.empty() ? 0.0f : UV[2 * vi + 1]);
          if (need_tangents) {
            const float* tg = &TG[4 * vi];
            const double td[3] = {tg[0], tg[1], tg[2]};
            // bitangent in the mesh's space (vertex_pool_bindings.wgsl:43-46)
            const double bd[3] = {(nd[1] * td[2] - nd[2] * td[1]) * tg[3], (nd[2] * td[0] - nd[0] * td[2]) * tg[3],
                                  (nd[0] * td[1] - nd[1] * td[0]) * tg[3]};
            itx(td, out->tangents);
            out->tangents.push_back(tg[3]);
            itx(bd, out->bitangents);
          }
        }
        out->mat_ids.push_back(mat);
      }
    }
    return true;
  }

  bool node(size_t idx, const M4& parent, int depth) {
    if (depth > 256) return fail("node hierarchy too deep (a cycle?)");
    const Json* nodes = doc.get("nodes");
    const Json* n = nodes ? nodes->at(idx) : nullptr;
    if (!n) return fail("missing node");
    const M4 world = mul(parent, node_local(*n));
    if (const Json* mi = n->get("mesh")) {
      const Json* me = doc.get("meshes") ? doc.get("meshes")->at((size_t)mi->num) : nullptr;
      if (!me) return fail("missing mesh");
      if (!mesh(*me, world)) return false;
    }
    if (const Json* ch = n->get("children"))
      for (const Json& c : ch->arr)
        if (!node((size_t)c.num, world, depth + 1)) return false;
    return true;
  }

  bool run() {
    const Json* mats = doc.get("materials");
    if (mats)
      for (const Json& m : mats->arr) out->materials.push_back(material(&m));
    if (out->materials.empty()) out->materials.push_back(material(nullptr));
    for (const apt_material& m : out->materials)
      need_tangents = need_tangents || m.normal_texture != APT_INVALID_TEXTURE ||
                      m.clearcoat_normal_texture != APT_INVALID_TEXTURE;
    const Json* scenes = doc.get("scenes");
    const size_t si = (size_t)doc.number("scene", 0);
    const Json* sc = scenes ? scenes->at(si) : nullptr;
    if (!sc) return true;  // no default scene: no geometry (asset.rs:34)
    if (const Json* roots = sc->get("nodes"))
      for (const Json& r : roots->arr)
        if (!node((size_t)r.num, identity(), 0)) return false;
    if (out->positions.size() / 9 >= (1u << 28)) return fail("more than 2^28 triangles");
    if (!any_uv) out->uvs.clear();
    return true;
  }
};

float3 f3at(const std::vector<float>& v, size_t i) { return make_float3(v[3 * i], v[3 * i + 1], v[3 * i + 2]); }
float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
float3 scale3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

bool read_file(const char* path, std::vector<uint8_t>& data) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  data.assign((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  return true;
}

}  // namespace

void generate_tangents(const std::vector<float>& pos, const std::vector<float>& nrm, const std::vector<float>& uv,
                       const std::vector<uint32_t>& idx, std::vector<float>& out) {
  const size_t nv = pos.size() / 3;
  std::vector<float3> tan1(nv, make_float3(0.0f, 0.0f, 0.0f)), tan2(nv, make_float3(0.0f, 0.0f, 0.0f));
  for (size_t i = 0; i + 2 < idx.size(); i += 3) {
    const uint32_t i1 = idx[i], i2 = idx[i + 1], i3 = idx[i + 2];
    const float3 v1 = f3at(pos, i1), v2 = f3at(pos, i2), v3 = f3at(pos, i3);
    const float x1 = v2.x - v1.x, x2 = v3.x - v1.x, y1 = v2.y - v1.y, y2 = v3.y - v1.y;
    const float z1 = v2.z - v1.z, z2 = v3.z - v1.z;
    const float s1 = uv[2 * i2] - uv[2 * i1], s2 = uv[2 * i3] - uv[2 * i1];
    const float t1 = uv[2 * i2 + 1] - uv[2 * i1 + 1], t2 = uv[2 * i3 + 1] - uv[2 * i1 + 1];
    const float rdiv = s1 * t2 - s2 * t1;
    const float r = rdiv == 0.0f ? 0.0f : 1.0f / rdiv;
    const float3 sdir = make_float3((t2 * x1 - t1 * x2) * r, (t2 * y1 - t1 * y2) * r, (t2 * z1 - t1 * z2) * r);
    const float3 tdir = make_float3((s1 * x2 - s2 * x1) * r, (s1 * y2 - s2 * y1) * r, (s1 * z2 - s2 * z1) * r);
    for (uint32_t k : {i1, i2, i3}) {
      tan1[k] = make_float3(tan1[k].x + sdir.x, tan1[k].y + sdir.y, tan1[k].z + sdir.z);
      tan2[k] = make_float3(tan2[k].x + tdir.x, tan2[k].y + tdir.y, tan2[k].z + tdir.z);
    }
  }
  out.assign(nv * 4, 0.0f);
  for (size_t i = 0; i < nv; ++i) {
    const float3 n = f3at(nrm, i), t = tan1[i];
    float3 xyz = sub3(t, scale3(n, dot3(n, t)));
    const float l = std::sqrt(dot3(xyz, xyz));
    if (l > 0.0f && std::isfinite(l)) {
      xyz = scale3(xyz, 1.0f / l);
    } else {
      // degenerate uvs: any unit vector orthogonal to the normal (the
      // reference's normalize of a zero vector gives NaN here)
      const float3 a = std::fabs(n.x) > 0.9f ? make_float3(0.0f, 1.0f, 0.0f) : make_float3(1.0f, 0.0f, 0.0f);
      xyz = cross3(n, a);
      const float la = std::sqrt(dot3(xyz, xyz));
      xyz = la > 0.0f ? scale3(xyz, 1.0f / la) : make_float3(1.0f, 0.0f, 0.0f);
    }
    const float w = dot3(cross3(n, t), tan2[i]) < 0.0f ? -1.0f : 1.0f;
    out[4 * i] = xyz.x;
    out[4 * i + 1] = xyz.y;
    out[4 * i + 2] = xyz.z;
    out[4 * i + 3] = w;
  }
}

apt_status gltf_load(const char* path, apt_scene_data** out) {
  std::vector<uint8_t> data;
  APT_REQUIRE(read_file(path, data), APT_ERR_NOT_FOUND, "glTF: cannot read '%s'", path);
  Loader L;
  const std::string sp(path);
  const size_t slash = sp.find_last_of('/');
  L.dir = slash == std::string::npos ? "" : sp.substr(0, slash + 1);
  const uint8_t* json = data.data();
  size_t json_len = data.size();
  std::vector<uint8_t> bin;
  bool glb = data.size() >= 12 && std::memcmp(data.data(), "glTF", 4) == 0;
  if (glb) {
    uint32_t version, length;
    std::memcpy(&version, data.data() + 4, 4);
    std::memcpy(&length, data.data() + 8, 4);
    APT_REQUIRE(version == 2 && length <= data.size() && data.size() >= 20, APT_ERR_UNSUPPORTED,
                "glTF: '%s' is not a glTF 2.0 binary", path);
    size_t off = 12;
    json = nullptr;
    while (off + 8 <= length) {
      uint32_t clen, ctype;
      std::memcpy(&clen, data.data() + off, 4);
      std::memcpy(&ctype, data.data() + off + 4, 4);
      APT_REQUIRE(off + 8 + clen <= length, APT_ERR_UNSUPPORTED, "glTF: truncated chunk in '%s'", path);
      if (ctype == 0x4E4F534Au && !json) {  // "JSON"
        json = data.data() + off + 8;
        json_len = clen;
      } else if (ctype == 0x004E4942u && bin.empty()) {  // "BIN\0"
        bin.assign(data.data() + off + 8, data.data() + off + 8 + clen);
      }
      off += 8 + clen;
    }
    APT_REQUIRE(json != nullptr, APT_ERR_UNSUPPORTED, "glTF: '%s' has no JSON chunk", path);
  }
  JsonParser P{reinterpret_cast<const char*>(json), reinterpret_cast<const char*>(json) + json_len, {}};
  APT_REQUIRE(P.value(L.doc, 0), APT_ERR_UNSUPPORTED, "glTF: '%s': %s", path, P.err.c_str());
  std::unique_ptr<apt_scene_data> s(new apt_scene_data());
  L.out = s.get();
  if (!L.load_buffers(glb ? &bin : nullptr) || !L.run()) {
    set_error("glTF: '%s': %s", path, L.err.c_str());
    return APT_ERR_UNSUPPORTED;
  }
  // a camera at +z looking down -z, as the procedural scenes' default
  const float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 5, 1};
  std::memcpy(s->camera.transform, id, sizeof(id));
  s->camera.fov_deg = 60.0f;
  s->camera.near_plane = 0.01f;
  s->camera.far_plane = 1000.0f;
  *out = s.release();
  return APT_OK;
}

}  // namespace apt

extern "C" apt_status apt_scene_load_gltf(const char* path, apt_scene_data** out) {
  apt::clear_error();
  APT_REQUIRE(path && out, APT_ERR_INVALID_ARG, "apt_scene_load_gltf: null argument");
  return apt::gltf_load(path, out);
}
