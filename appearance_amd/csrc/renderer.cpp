// Renderer object of the C ABI: the host side of PathTracerGpu
// (crates/appearance-path-tracer-gpu/src/lib.rs:250-615) and of the
// NodeRenderer job interface (crates/appearance-render-loop/src/node.rs:14-25).
//
// Host responsibilities: world actions (camera, models spawned from procedural
// or glTF assets as instances of per-asset BLASes, placed by a TLAS rebuilt on
// every action), scene upload (BVH or TLAS, triangle attributes, emitter list), frame
// buffers for the requested slab, the sample/bounce schedule of the wavefront
// integrator, progressive accumulation and the blocking readback.
#include <chrono>
#include <cstddef>
#include <map>
#include <memory>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "apt_internal.h"
#include "integrator.h"
#include "jpeg.h"
#include "restir.h"
#include "scene_gpu.h"
#include "tlas.h"

namespace {

using namespace apt;

// A loaded model asset: its triangles resident in HBM in object space and its
// BLAS in the renderer's two-level scene.  Spawns of the same asset path share
// it, as the reference shares one SceneModel per asset path
// (scene_resources/mod.rs:115-133); it lives until Clear.
struct Asset {
  apt_scene_data* data = nullptr;
  uint32_t n_tris = 0, n_mats = 0;
  bool emissive = false;
  float* d_xyz = nullptr;
  float* d_nrm = nullptr;
  uint32_t* d_mat = nullptr;
  float2* d_uv = nullptr;      // per vertex, or null
  float4* d_tan = nullptr;     // per vertex (xyz, handedness), or null
  float4* d_bit = nullptr;     // per vertex, or null
  uint32_t* d_texels = nullptr;       // the asset's textures, one after another
  std::vector<TexDesc> tex;           // their descriptors (into d_texels)
  uint32_t blas = INVALID;  // BLAS index in the renderer's two-level scene
  ~Asset() {
    apt_scene_data_destroy(data);
    for (void* p : {(void*)d_xyz, (void*)d_nrm, (void*)d_mat, (void*)d_uv, (void*)d_tan, (void*)d_bit,
                    (void*)d_texels})
      if (p) (void)hipFree(p);
  }
};

struct Model {
  uint8_t uuid[16];
  float xform[16];  // column-major
  std::shared_ptr<Asset> asset;
};

// Copies an asset's triangles to the device once (checking its material ids).
apt_status asset_resident(Asset& a) {
  if (a.d_xyz) return APT_OK;
  apt_mesh_desc md;
  apt_scene_data_view(a.data, &md, nullptr, nullptr);
  const size_t n = md.n_tris;
  std::vector<uint32_t> ids(n, 0u);
  for (size_t t = 0; t < n; ++t) {
    const uint32_t id = md.material_ids ? md.material_ids[t] : 0u;
    APT_REQUIRE(id < md.n_materials, APT_ERR_INVALID_ARG, "model: triangle %zu uses material %u of %u", t, id,
                md.n_materials);
    ids[t] = id;
  }
  const auto put = [](auto*& dst, const auto* src, size_t count) -> apt_status {
    hipError_t e = hipMalloc(&dst, std::max<size_t>(1, count) * sizeof(*src));
    if (e == hipSuccess && count) e = hipMemcpy(dst, src, count * sizeof(*src), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      set_error("model upload: %s", hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? APT_ERR_OOM : APT_ERR_DEVICE;
    }
    return APT_OK;
  };
  apt_status st = put(a.d_xyz, md.positions, 9 * n);
  if (st == APT_OK && md.normals) st = put(a.d_nrm, md.normals, 9 * n);
  if (st == APT_OK) st = put(a.d_mat, ids.data(), n);
  // uvs, tangent frames and textures (glTF assets)
  const apt_scene_data& d = *a.data;
  if (st == APT_OK && !d.uvs.empty()) st = put(a.d_uv, reinterpret_cast<const float2*>(d.uvs.data()), 3 * n);
  if (st == APT_OK && !d.tangents.empty())
    st = put(a.d_tan, reinterpret_cast<const float4*>(d.tangents.data()), 3 * n);
  if (st == APT_OK && !d.bitangents.empty()) {
    std::vector<float4> b(3 * n);
    for (size_t j = 0; j < 3 * n; ++j)
      b[j] = make_float4(d.bitangents[3 * j], d.bitangents[3 * j + 1], d.bitangents[3 * j + 2], 0.0f);
    st = put(a.d_bit, b.data(), 3 * n);
  }
  if (st == APT_OK && !d.textures.empty()) {
    std::vector<uint32_t> texels;
    std::vector<size_t> off;
    for (const apt_scene_texture& t : d.textures) {
      off.push_back(texels.size());
      texels.insert(texels.end(), t.rgba.begin(), t.rgba.end());
    }
    st = put(a.d_texels, texels.data(), texels.size());
    a.tex.clear();
    for (size_t k = 0; st == APT_OK && k < d.textures.size(); ++k)
      a.tex.push_back(TexDesc{a.d_texels + off[k], d.textures[k].width, d.textures[k].height});
  }
  if (st != APT_OK) {
    for (void* p : {(void*)a.d_xyz, (void*)a.d_nrm, (void*)a.d_mat, (void*)a.d_uv, (void*)a.d_tan, (void*)a.d_bit,
                    (void*)a.d_texels})
      if (p) (void)hipFree(p);
    a.d_xyz = a.d_nrm = nullptr;
    a.d_mat = nullptr;
    a.d_uv = nullptr;
    a.d_tan = a.d_bit = nullptr;
    a.d_texels = nullptr;
    a.tex.clear();
  }
  return st;
}

// A material's texture indices offset by the model's first texture in the
// scene's texture table.
void offset_textures(apt_material& m, uint32_t base) {
  for (uint32_t* t : {&m.color_texture, &m.metallic_roughness_texture, &m.normal_texture, &m.emission_texture,
                      &m.sheen_texture, &m.clearcoat_texture, &m.clearcoat_roughness_texture,
                      &m.sheen_tint_texture, &m.clearcoat_normal_texture, &m.transmission_texture})
    if (*t != APT_INVALID_TEXTURE) *t += base;
}

// Page-locked host memory, so the frame's readback runs at the link's rate
// with no staging copy.
struct PinnedBytes {
  uint8_t* p = nullptr;
  size_t n = 0;
  apt_status ensure(size_t count) {
    if (count <= n && p) return APT_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipHostMalloc((void**)&p, count ? count : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipHostMalloc(%zu B) failed: %s", count, hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? APT_ERR_OOM : APT_ERR_DEVICE;
    }
    n = count;
    return APT_OK;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  apt_status ensure(size_t count) {
    if (count <= n && p) return APT_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc(%zu B) failed: %s", count * sizeof(T), hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? APT_ERR_OOM : APT_ERR_DEVICE;
    }
    n = count;
    return APT_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Per-pixel state of the ReSTIR integrator for one slab.
struct RestirState {
  DevBuf<float4> ray_o, ray_d, thr, radiance, vtx_p, vtx_n, vtx_t, vtx_m, demod[2], work, col[2], hist;
  DevBuf<float4> rq_o, rq_d, rq_hit;
  DevBuf<uint32_t> rng, count, counters, rq_ctr;
  DevBuf<DiRes> di, di_tmp, di_hist[2];
  DevBuf<GiRes> gi, gi_tmp, gi_hist[2];
  DevBuf<GTexel> gbuf[2];
  DevBuf<float2> mom[2];
  DevBuf<uint2> guide;
  DevBuf<float> di_self;
  size_t npix = 0;

  template <typename F>
  void each(F&& f) {
    f(ray_o); f(ray_d); f(thr); f(radiance); f(vtx_p); f(vtx_n); f(vtx_t); f(vtx_m); f(demod[0]); f(demod[1]);
    f(work); f(col[0]); f(col[1]); f(hist); f(rng); f(count); f(di); f(di_tmp); f(di_hist[0]);
    f(di_hist[1]); f(gi); f(gi_tmp); f(gi_hist[0]); f(gi_hist[1]); f(gbuf[0]); f(gbuf[1]); f(mom[0]); f(mom[1]);
    f(guide); f(di_self); f(rq_o); f(rq_d); f(rq_hit);
  }
  // Sizes every buffer for n pixels and clears them (no history).
  apt_status reset(size_t n, hipStream_t s) {
    apt_status st = APT_OK;
    each([&](auto& b) {
      if (st == APT_OK) st = b.ensure(n);
      if (st == APT_OK && hipMemsetAsync(b.p, 0, n * sizeof(*b.p), s) != hipSuccess) {
        set_error("restir: clearing state failed");
        st = APT_ERR_DEVICE;
      }
    });
    if (st == APT_OK) st = counters.ensure(2 * RS_CTR_SLOTS);
    if (st == APT_OK) st = rq_ctr.ensure(RQ_WORDS);
    npix = st == APT_OK ? n : 0;
    return st;
  }
  void release() {
    each([](auto& b) { b.release(); });
    counters.release();
    rq_ctr.release();
    npix = 0;
  }
  RestirBuffers view() {
    RestirBuffers b{};
    b.ray_o = ray_o.p; b.ray_d = ray_d.p; b.thr = thr.p; b.rng = rng.p; b.radiance = radiance.p;
    b.vtx_p = vtx_p.p; b.vtx_n = vtx_n.p; b.vtx_t = vtx_t.p; b.vtx_m = vtx_m.p;
    b.di = di.p; b.di_tmp = di_tmp.p; b.di_hist[0] = di_hist[0].p; b.di_hist[1] = di_hist[1].p;
    b.di_self = di_self.p;
    b.gi = gi.p; b.gi_tmp = gi_tmp.p; b.gi_hist[0] = gi_hist[0].p; b.gi_hist[1] = gi_hist[1].p;
    b.gbuf[0] = gbuf[0].p; b.gbuf[1] = gbuf[1].p;
    b.demod[0] = demod[0].p; b.demod[1] = demod[1].p; b.work = work.p;
    b.svgf_col[0] = col[0].p; b.svgf_col[1] = col[1].p; b.svgf_mom[0] = mom[0].p; b.svgf_mom[1] = mom[1].p;
    b.svgf_guide = guide.p; b.svgf_count = count.p; b.svgf_hist = hist.p;
    b.counters = counters.p;
    b.rq_o = rq_o.p; b.rq_d = rq_d.p; b.rq_hit = rq_hit.p; b.rq_ctr = rq_ctr.p;
    return b;
  }
};

CamView cam_view(const SceneView& sc) {
  return CamView{sc.cam_pos, sc.cam_right, sc.cam_up, sc.cam_back, sc.cam_tan_half, sc.cam_aspect};
}

// Normal transform: inverse transpose of the upper 3x3.
bool normal_matrix(const float* m, float* n) {
  const double a = m[0], b = m[4], c = m[8], d = m[1], e = m[5], f = m[9], g = m[2], h = m[6], i = m[10];
  const double A = e * i - f * h, B = -(d * i - f * g), Cc = d * h - e * g;
  const double D = -(b * i - c * h), E = a * i - c * g, F = -(a * h - b * g);
  const double G = b * f - c * e, H = -(a * f - c * d), I = a * e - b * d;
  const double det = a * A + b * B + c * Cc;
  if (std::fabs(det) < 1e-30) return false;
  // rows of the inverse transpose = cofactor rows / det
  n[0] = (float)(A / det); n[1] = (float)(B / det); n[2] = (float)(Cc / det);
  n[3] = (float)(D / det); n[4] = (float)(E / det); n[5] = (float)(F / det);
  n[6] = (float)(G / det); n[7] = (float)(H / det); n[8] = (float)(I / det);
  return true;
}

}  // namespace

struct apt_renderer {
  apt_config cfg{};
  // camera (CameraUpdate); the reference's default camera until one arrives
  apt_camera cam{};
  uint64_t cam_version = 1, accum_cam_version = 0;
  // scene: spawned models (instances of assets in a two-level scene) or an
  // explicit triangle soup (apt_scene_set, one BVH)
  std::vector<Model> models;
  std::map<std::string, std::shared_ptr<Asset>> assets;
  apt_tlas* tlas = nullptr;
  bool emitters_dirty = true;
  std::vector<float> pos, nrm, uvs, tangents, bitangents;
  std::vector<apt_scene_texture> textures;
  std::vector<uint32_t> mat_ids;
  std::vector<apt_material> mats;
  apt_sky sky{};
  // sky panorama (apt_set_sky_panorama): RGB texels as float4, or none
  DevBuf<float4> d_sky_pano;
  uint32_t pano_w = 0, pano_h = 0;
  bool scene_dirty = true, explicit_scene = false, has_normals = false;
  uint64_t scene_version = 1, accum_scene_version = 0;
  // device scene
  SceneGpu gpu;
  // frame
  uint32_t fw = 0, fh = 0, frow0 = 0, frows = 0;
  DevBuf<float4> d_frame_sum, d_accum_tm, d_accum_lin;
  DevBuf<uint32_t> d_rgba;
  uint32_t accum_frames = 0;
  std::vector<float4> h_lin;
  // paths
  uint32_t path_cap = 0;
  DevBuf<float4> p_o, p_d, p_t, p_l, p_h, p_so, p_sd, p_sc;
  DevBuf<uint32_t> p_rng, p_q0, p_q1, p_qs, p_cnt;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // start/end events around each extend, shade and shadow launch of a frame
  // (6 per bounce per pass), for the per-stage device times
  std::vector<hipEvent_t> stage_ev;
  apt_render_stats stats{};
  uint32_t frame_idx = 0;
  PinnedBytes cb_pixels;  // the callback's borrowed RGBA8 slab
  // apt_start_render_jpeg: resolve the RGBA8 frame into d_rgba even with no
  // host output, and compress its blocks there
  bool rgba_on_device = false;
  JpegEncoder jpeg;
  // ReSTIR integrator: per-pixel state, frame parity and the previous camera
  RestirState rs;
  uint32_t rs_cur = 0;
  bool rs_history = false;
  CamView rs_prev_cam{};
  uint64_t rs_scene_version = 0;
  uint32_t rs_w = 0, rs_h = 0, rs_row0 = 0;

  ~apt_renderer() {
    models.clear();
    assets.clear();
    scene_gpu_free(gpu);
    apt_tlas_destroy(tlas);
    d_frame_sum.release(); d_accum_tm.release(); d_accum_lin.release(); d_rgba.release();
    d_sky_pano.release();
    p_o.release(); p_d.release(); p_t.release(); p_l.release(); p_h.release(); p_so.release();
    p_sd.release(); p_sc.release(); p_rng.release(); p_q0.release(); p_q1.release();
    p_qs.release(); p_cnt.release();
    rs.release();
    cb_pixels.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t e : stage_ev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

apt_status validate_config(const apt_config& c) {
  APT_REQUIRE(c.max_bounces >= 1 && c.max_bounces <= 64, APT_ERR_INVALID_ARG,
              "config: max_bounces %u outside 1..64", c.max_bounces);
  APT_REQUIRE(c.sample_count >= 1 && c.sample_count <= 65536, APT_ERR_INVALID_ARG,
              "config: sample_count %u outside 1..65536", c.sample_count);
  APT_REQUIRE(c.bvh_width == 2 || c.bvh_width == 4 || c.bvh_width == 8, APT_ERR_INVALID_ARG,
              "config: bvh_width %u is not 2, 4 or 8", c.bvh_width);
  APT_REQUIRE(c.lds_stack_depth == 8 || c.lds_stack_depth == 12 || c.lds_stack_depth == 16,
              APT_ERR_INVALID_ARG, "config: lds_stack_depth %u is not 8, 12 or 16", c.lds_stack_depth);
  APT_REQUIRE(c.integrator == APT_INTEGRATOR_PLAIN || c.integrator == APT_INTEGRATOR_RESTIR,
              APT_ERR_INVALID_ARG, "config: unknown integrator %u", c.integrator);
  if (c.integrator == APT_INTEGRATOR_PLAIN) {
    APT_REQUIRE(!c.restir_di && !c.restir_gi, APT_ERR_INVALID_ARG,
                "config: restir_di/restir_gi need integrator APT_INTEGRATOR_RESTIR");
    APT_REQUIRE(!c.svgf && !c.taa && !c.firefly_filter, APT_ERR_INVALID_ARG,
                "config: svgf, taa and firefly_filter need the G-buffer of APT_INTEGRATOR_RESTIR");
    return APT_OK;
  }
  APT_REQUIRE(c.di_candidates <= 64, APT_ERR_INVALID_ARG, "config: di_candidates %u above 64", c.di_candidates);
  APT_REQUIRE(c.di_spatial_neighbours <= 16, APT_ERR_INVALID_ARG, "config: di_spatial_neighbours %u above 16",
              c.di_spatial_neighbours);
  APT_REQUIRE(c.di_spatial_radius >= 0.0f && c.di_spatial_radius <= 4096.0f, APT_ERR_INVALID_ARG,
              "config: di_spatial_radius %g outside [0, 4096]", (double)c.di_spatial_radius);
  APT_REQUIRE(c.svgf_iterations <= 8, APT_ERR_INVALID_ARG, "config: svgf_iterations %u above 8", c.svgf_iterations);
  APT_REQUIRE(c.svgf_history >= 1, APT_ERR_INVALID_ARG, "config: svgf_history must be at least 1");
  APT_REQUIRE(c.gi_spatial_passes <= 4, APT_ERR_INVALID_ARG, "config: gi_spatial_passes %u above 4",
              c.gi_spatial_passes);
  return APT_OK;
}

// Row-major 3x4 of a column-major 4x4 (the instance layout).
void xform_rows(const float* m, float* out) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 4; ++c) out[4 * r + c] = m[4 * c + r];
}

// The spawned models as a two-level scene: each asset's BLAS is built once in
// object space (the device builder) and copied into the renderer's scene; a
// world action then rebuilds the TLAS over the instances and re-places the
// world-space attributes the shading reads (positions, normals, material ids)
// from the resident object-space triangles.  The emitter list is rebuilt only
// when the set of models changes or an emissive model moves.
apt_status upload_instances(apt_renderer* r) {
  const apt_config& c = r->cfg;
  hipStream_t s = r->stream;
  apt_status st;
  if (r->tlas && r->tlas->width != c.bvh_width) {
    r->gpu.tlas = nullptr;
    apt_tlas_destroy(r->tlas);
    r->tlas = nullptr;
    for (auto& kv : r->assets) kv.second->blas = INVALID;
  }
  if (!r->tlas) {
    apt_bvh_opts o{};
    o.width = c.bvh_width;
    o.lds_stack = c.lds_stack_depth;
    if ((st = apt_tlas_create(&o, &r->tlas)) != APT_OK) return st;
  }
  r->tlas->lds_stack = c.lds_stack_depth;
  size_t total = 0;
  bool all_normals = !r->models.empty(), any_uv = false, any_tan = false;
  for (Model& m : r->models) {
    Asset& a = *m.asset;
    if ((st = asset_resident(a)) != APT_OK) return st;
    if (a.blas == INVALID) {
      apt_bvh* b = nullptr;
      const apt_bvh_opts o{c.bvh_width, 4, 32, APT_BUILDER_GPU, c.lds_stack_depth, 0};
      st = apt_bvh_build_device(a.d_xyz, a.n_tris, &o, &b);
      if (st == APT_OK) st = check_trace_stack(b->stats, "model");
      if (st == APT_OK) st = apt_tlas_add_blas(r->tlas, b, &a.blas);
      apt_bvh_destroy(b);
      if (st != APT_OK) return st;
    }
    all_normals = all_normals && a.d_nrm != nullptr;
    any_uv = any_uv || a.d_uv != nullptr;
    any_tan = any_tan || a.d_tan != nullptr;
    total += a.n_tris;
  }
  APT_REQUIRE(total < (1u << 28), APT_ERR_INVALID_ARG, "scene: %zu triangles exceed 2^28", total);
  APT_REQUIRE(r->models.size() <= TLAS_MAX_INSTANCES, APT_ERR_INVALID_ARG, "scene: %zu models exceed %u",
              r->models.size(), TLAS_MAX_INSTANCES);
  r->gpu.tlas = nullptr;
  if (total == 0) return APT_OK;
  std::vector<apt_instance> inst(r->models.size());
  std::vector<ModelXform> place(r->models.size());
  r->mats.clear();
  // texture table: each asset's textures once, in the order of first use
  std::vector<TexDesc> tex;
  std::map<const Asset*, uint32_t> tex_base;
  for (const Model& m : r->models)
    if (!m.asset->tex.empty() && tex_base.emplace(m.asset.get(), (uint32_t)tex.size()).second)
      tex.insert(tex.end(), m.asset->tex.begin(), m.asset->tex.end());
  uint32_t base = 0;
  for (size_t k = 0; k < r->models.size(); ++k) {
    const Model& m = r->models[k];
    std::memset(&inst[k], 0, sizeof(apt_instance));
    xform_rows(m.xform, inst[k].transform);
    inst[k].blas = m.asset->blas;
    inst[k].custom_index = base;
    ModelXform& x = place[k];
    std::memcpy(x.m, m.xform, sizeof(x.m));
    x.nm_ok = normal_matrix(m.xform, x.nm) ? 1u : 0u;
    x.mat_base = (uint32_t)r->mats.size();
    apt_mesh_desc md;
    apt_scene_data_view(m.asset->data, &md, nullptr, nullptr);
    const auto tb = tex_base.find(m.asset.get());
    for (uint32_t k = 0; k < md.n_materials; ++k) {
      apt_material mt = md.materials[k];
      if (tb != tex_base.end()) offset_textures(mt, tb->second);
      r->mats.push_back(mt);
    }
    base += m.asset->n_tris;
  }
  if ((st = apt_tlas_set_instances(r->tlas, inst.data(), (uint32_t)inst.size())) != APT_OK) return st;
  // tangent frames need every model's normals
  any_tan = any_tan && all_normals;
  if ((st = scene_reserve(r->gpu, (uint32_t)total, all_normals, any_uv, any_tan, r->mats.data(),
                          (uint32_t)r->mats.size(), tex.data(), (uint32_t)tex.size(), s)) != APT_OK)
    return st;
  base = 0;
  for (size_t k = 0; k < r->models.size(); ++k) {
    const Asset& a = *r->models[k].asset;
    ModelAttribs at;
    at.uv = a.d_uv;
    at.tan = a.d_tan;
    at.bit = a.d_bit;
    if ((st = scene_place_model(a.d_xyz, a.d_nrm, a.d_mat, a.n_tris, place[k], at, r->gpu, base, s)) != APT_OK)
      return st;
    base += a.n_tris;
  }
  if (r->emitters_dirty) {
    if ((st = scene_emitters(r->gpu, s)) != APT_OK) return st;
    r->emitters_dirty = false;
  } else {
    APT_HIP(hipStreamSynchronize(s));
  }
  r->gpu.sky = r->sky;
  r->gpu.tlas = r->tlas;
  return APT_OK;
}

// The scene as the kernels see it: the device scene, the camera and the sky
// panorama.
SceneView render_view(const apt_renderer* r, uint32_t width, uint32_t height) {
  SceneView sc = scene_gpu_view(r->gpu, r->cam, width, height);
  if (r->pano_w) {
    sc.sky_tex = r->d_sky_pano.p;
    sc.sky_w = r->pano_w;
    sc.sky_h = r->pano_h;
  }
  return sc;
}

apt_status upload_scene(apt_renderer* r) {
  if (!r->scene_dirty) return APT_OK;
  const auto t0 = std::chrono::steady_clock::now();
  if (!r->explicit_scene) {
    if (r->gpu.bvh) scene_gpu_free(r->gpu);  // an explicit scene before
    apt_status st = upload_instances(r);
    if (st != APT_OK) return st;
  } else {
    scene_gpu_free(r->gpu);
    const uint32_t n = (uint32_t)r->mat_ids.size();
    if (n > 0) {
      apt_status st = scene_gpu_upload(r->pos.data(), r->has_normals ? r->nrm.data() : nullptr,
                                       r->mat_ids.data(), n, r->mats.data(), (uint32_t)r->mats.size(),
                                       &r->sky, r->cfg.bvh_width, r->cfg.lds_stack_depth, r->gpu);
      if (st != APT_OK) return st;
      std::vector<apt_texture> tv;
      for (const apt_scene_texture& t : r->textures) tv.push_back(apt_texture{t.width, t.height, t.rgba.data()});
      apt_mesh_attribs at{};
      at.uvs = r->uvs.empty() ? nullptr : r->uvs.data();
      at.tangents = r->tangents.empty() ? nullptr : r->tangents.data();
      at.bitangents = r->bitangents.empty() ? nullptr : r->bitangents.data();
      at.textures = tv.empty() ? nullptr : tv.data();
      at.n_textures = (uint32_t)tv.size();
      st = scene_gpu_attribs(r->gpu, r->pos.data(), r->has_normals ? r->nrm.data() : nullptr, n, at);
      if (st != APT_OK) return st;
    }
  }
  r->scene_dirty = false;
  r->scene_version++;
  r->stats.ms_scene = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return APT_OK;
}

apt_status ensure_paths(apt_renderer* r, uint32_t n) {
  if (n <= r->path_cap) return APT_OK;
  apt_status st;
  for (DevBuf<float4>* b : {&r->p_o, &r->p_d, &r->p_t, &r->p_l, &r->p_h, &r->p_so, &r->p_sd, &r->p_sc})
    if ((st = b->ensure(n)) != APT_OK) return st;
  for (DevBuf<uint32_t>* b : {&r->p_rng, &r->p_q0, &r->p_q1, &r->p_qs})
    if ((st = b->ensure(n)) != APT_OK) return st;
  if ((st = r->p_cnt.ensure(CNT_COUNT)) != APT_OK) return st;
  r->path_cap = n;
  return APT_OK;
}

// Copies the frame out (RGBA8 block layout and/or the FP32 mean) and waits.
apt_status finish_frame(apt_renderer* r, size_t npix, uint8_t* out_rgba, float* out_accum) {
  hipStream_t s = r->stream;
  if (out_rgba) APT_HIP(hipMemcpyAsync(out_rgba, r->d_rgba.p, npix * 4, hipMemcpyDeviceToHost, s));
  if (out_accum) {
    r->h_lin.resize(npix);
    APT_HIP(hipMemcpyAsync(r->h_lin.data(), r->d_accum_lin.p, npix * sizeof(float4), hipMemcpyDeviceToHost, s));
  }
  APT_HIP(hipStreamSynchronize(s));
  if (out_accum) {
    for (size_t i = 0; i < npix; ++i) {
      const float4 v = r->h_lin[i];
      const float inv = v.w > 0.0f ? 1.0f / v.w : 0.0f;
      out_accum[3 * i] = v.x * inv;
      out_accum[3 * i + 1] = v.y * inv;
      out_accum[3 * i + 2] = v.z * inv;
    }
  }
  return APT_OK;
}

// The ReSTIR frame (PathTracerGpu::render, src/lib.rs:354-576): samples at
// pixel centres with reuse against the previous frame's reservoirs and
// G-buffer, the post chain, then the shared resolve.
apt_status render_restir(apt_renderer* r, uint32_t width, uint32_t height, uint32_t row_start,
                         uint32_t rows, uint8_t* out_rgba, float* out_accum) {
  hipStream_t s = r->stream;
  const size_t npix = (size_t)width * rows;
  apt_status st;
  if (r->rs.npix != npix || r->rs_scene_version != r->scene_version || r->rs_w != width || r->rs_h != height ||
      r->rs_row0 != row_start) {
    if ((st = r->rs.reset(npix, s)) != APT_OK) return st;
    r->rs_history = false;
    r->rs_cur = 0;
    r->rs_scene_version = r->scene_version;
    r->rs_w = width;
    r->rs_h = height;
    r->rs_row0 = row_start;
  }
  APT_HIP(hipMemsetAsync(r->rs.radiance.p, 0, npix * sizeof(float4), s));
  APT_HIP(hipMemsetAsync(r->rs.counters.p, 0, 2 * RS_CTR_SLOTS * sizeof(uint32_t), s));
  const SceneView sc = render_view(r, width, height);
  const CamView cur = cam_view(sc);
  const apt_config& c = r->cfg;
  RestirParams p{};
  p.width = width;
  p.rows = rows;
  p.row_start = row_start;
  p.height = height;
  p.max_bounces = c.max_bounces;
  p.di_candidates = c.di_candidates;
  p.di_mcap = c.di_mcap;
  p.neighbours = c.di_spatial_neighbours;
  p.radius = c.di_spatial_radius;
  p.gi_mcap = c.gi_mcap;
  p.gi_spatial_passes = c.gi_spatial_passes;
  p.restir_di = c.restir_di ? 1u : 0u;
  p.restir_gi = c.restir_gi ? 1u : 0u;
  p.cur = r->rs_cur;
  p.svgf_history = c.svgf_history;
  p.atrous_passes = c.svgf_iterations;
  p.first_frame = r->rs_history ? 0u : 1u;
  p.prev_cam = r->rs_history ? r->rs_prev_cam : cur;
  const RestirBuffers b = r->rs.view();
  const uint32_t spp = c.sample_count;
  APT_HIP(hipEventRecord(r->ev0, s));
  for (uint32_t k = 0; k < spp; ++k) {
    p.seed = r->frame_idx * spp + k + c.seed;  // lib.rs:374
    if ((st = restir_sample(p, sc, b, s)) != APT_OK) return st;
  }
  if ((st = restir_post(p, b, c.firefly_filter != 0, c.svgf != 0, c.taa != 0, r->d_frame_sum.p, s)) != APT_OK)
    return st;
  FrameView fv{width, height, row_start, rows, spp, r->frame_idx, c.seed, c.max_bounces, 2};
  if ((st = resolve_frame(fv, r->d_frame_sum.p, r->d_accum_tm.p, r->d_accum_lin.p, r->accum_frames,
                          out_rgba || r->rgba_on_device ? r->d_rgba.p : nullptr, s)) != APT_OK)
    return st;
  APT_HIP(hipEventRecord(r->ev1, s));
  uint32_t part[2 * RS_CTR_SLOTS];
  APT_HIP(hipMemcpyAsync(part, r->rs.counters.p, sizeof(part), hipMemcpyDeviceToHost, s));
  if ((st = finish_frame(r, npix, out_rgba, out_accum)) != APT_OK) return st;
  float ms = 0.0f;
  APT_HIP(hipEventElapsedTime(&ms, r->ev0, r->ev1));
  const double ms_scene = r->stats.ms_scene;  // kept until the next rebuild
  r->stats = apt_render_stats{};
  r->stats.ms_scene = ms_scene;
  r->stats.paths = (uint64_t)npix * spp;
  uint64_t cnt[2] = {0, 0};
  for (uint32_t k = 0; k < 2 * RS_CTR_SLOTS; ++k) cnt[k / RS_CTR_SLOTS] += part[k];
  r->stats.extension_rays = cnt[0];
  r->stats.shadow_rays = cnt[1];
  r->stats.segments = cnt[0];
  r->stats.ms_total = ms;
  r->accum_frames++;
  r->frame_idx++;
  r->rs_cur ^= 1u;
  r->rs_prev_cam = cur;
  r->rs_history = true;
  return APT_OK;
}

apt_status render_impl(apt_renderer* r, uint32_t width, uint32_t height, uint32_t row_start,
                       uint32_t row_end, uint8_t* out_rgba, size_t out_len, float* out_accum,
                       size_t accum_len) {
  APT_REQUIRE(width > 0 && height > 0 && row_start < row_end && row_end <= height,
              APT_ERR_INVALID_ARG, "render: bad frame %ux%u rows [%u,%u)", width, height, row_start, row_end);
  const uint32_t rows = row_end - row_start;
  APT_REQUIRE((!out_rgba && !r->rgba_on_device) || (width % 64 == 0 && rows % 64 == 0), APT_ERR_INVALID_ARG,
              "render: block-layout output needs width (%u) and slab height (%u) to be multiples "
              "of 64 (host.rs:476)", width, rows);
  const size_t npix = (size_t)width * rows;
  APT_REQUIRE(npix <= (1u << 26), APT_ERR_INVALID_ARG, "render: %zu pixels exceed 2^26", npix);
  APT_REQUIRE(!out_rgba || out_len >= npix * 4, APT_ERR_INVALID_ARG,
              "render: out_rgba holds %zu bytes, needs %zu", out_len, npix * 4);
  APT_REQUIRE(!out_accum || accum_len >= npix * 3, APT_ERR_INVALID_ARG,
              "render: out_accum_rgb holds %zu floats, needs %zu", accum_len, npix * 3);
  apt_status st = upload_scene(r);
  if (st != APT_OK) return st;
  APT_REQUIRE(r->gpu.bvh != nullptr || r->gpu.tlas != nullptr, APT_ERR_INVALID_ARG, "render: the scene is empty");

  hipStream_t s = r->stream;
  const bool resized = r->fw != width || r->fh != height || r->frow0 != row_start || r->frows != rows;
  if (resized) {
    if ((st = r->d_frame_sum.ensure(npix)) != APT_OK) return st;
    if ((st = r->d_accum_tm.ensure(npix)) != APT_OK) return st;
    if ((st = r->d_accum_lin.ensure(npix)) != APT_OK) return st;
    if ((st = r->d_rgba.ensure(npix)) != APT_OK) return st;
    r->fw = width;
    r->fh = height;
    r->frow0 = row_start;
    r->frows = rows;
  }
  // Accumulation restarts on any change of view, scene or slab, or when
  // accumulation is off (lib.rs:349-352).
  if (resized || !r->cfg.accum_frames || r->accum_cam_version != r->cam_version ||
      r->accum_scene_version != r->scene_version) {
    APT_HIP(hipMemsetAsync(r->d_accum_tm.p, 0, npix * sizeof(float4), s));
    APT_HIP(hipMemsetAsync(r->d_accum_lin.p, 0, npix * sizeof(float4), s));
    r->accum_frames = 0;
    r->accum_cam_version = r->cam_version;
    r->accum_scene_version = r->scene_version;
  }
  APT_HIP(hipMemsetAsync(r->d_frame_sum.p, 0, npix * sizeof(float4), s));
  if (r->cfg.integrator == APT_INTEGRATOR_RESTIR)
    return render_restir(r, width, height, row_start, rows, out_rgba, out_accum);

  const uint32_t spp = r->cfg.sample_count;
  // Paths in flight per pass: with 2^24 (2.4 GB of path state) a 1080p frame
  // runs up to 8 samples per pass, so the late bounces' short queues still
  // hold enough rays to fill the device.
  const uint64_t budget = r->cfg.queue_chunk ? r->cfg.queue_chunk : (1u << 24);
  const uint32_t per_pass = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(spp, budget / npix));
  if ((st = ensure_paths(r, (uint32_t)(per_pass * npix))) != APT_OK) return st;
  APT_HIP(hipMemsetAsync(r->p_cnt.p, 0, CNT_COUNT * sizeof(uint32_t), s));

  FrameView fv{width, height, row_start, rows, spp, r->frame_idx, r->cfg.seed, r->cfg.max_bounces,
               r->cfg.rr_start_bounce};
  SceneView sc = render_view(r, width, height);
  PathView ps{};
  ps.o_pdf = r->p_o.p; ps.d_flags = r->p_d.p; ps.thr = r->p_t.p; ps.rad = r->p_l.p;
  ps.hit = r->p_h.p; ps.sh_o = r->p_so.p; ps.sh_d = r->p_sd.p; ps.sh_c = r->p_sc.p;
  ps.rng = r->p_rng.p; ps.queue[0] = r->p_q0.p; ps.queue[1] = r->p_q1.p;
  ps.shadow_queue = r->p_qs.p; ps.counters = r->p_cnt.p;

  // Stage events for every launch of the frame, up to a bounded count (a
  // frame of many passes is timed as a whole only).
  const uint32_t passes = (spp + per_pass - 1) / per_pass;
  const size_t n_ev = (size_t)passes * 6 * r->cfg.max_bounces;
  const bool staged = n_ev <= 1536;
  while (staged && r->stage_ev.size() < n_ev) {
    hipEvent_t e;
    APT_HIP(hipEventCreate(&e));
    r->stage_ev.push_back(e);
  }
  APT_HIP(hipEventRecord(r->ev0, s));
  for (uint32_t base = 0, pass = 0; base < spp; base += per_pass, ++pass) {
    const uint32_t cnt = std::min(per_pass, spp - base);
    hipEvent_t* ev = staged ? r->stage_ev.data() + (size_t)pass * 6 * r->cfg.max_bounces : nullptr;
    st = integrate_pass(fv, sc, ps, (uint32_t)(cnt * npix), base, cnt, r->d_frame_sum.p, s, ev);
    if (st != APT_OK) return st;
  }
  st = resolve_frame(fv, r->d_frame_sum.p, r->d_accum_tm.p, r->d_accum_lin.p, r->accum_frames,
                     out_rgba || r->rgba_on_device ? r->d_rgba.p : nullptr, s);
  if (st != APT_OK) return st;
  APT_HIP(hipEventRecord(r->ev1, s));
  uint32_t cnt[CNT_COUNT];
  APT_HIP(hipMemcpyAsync(cnt, r->p_cnt.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
  if ((st = finish_frame(r, npix, out_rgba, out_accum)) != APT_OK) return st;
  float ms = 0.0f;
  APT_HIP(hipEventElapsedTime(&ms, r->ev0, r->ev1));
  const double ms_scene = r->stats.ms_scene;  // kept until the next rebuild
  r->stats = apt_render_stats{};
  r->stats.ms_scene = ms_scene;
  r->stats.paths = (uint64_t)npix * spp;
  r->stats.extension_rays = cnt[CNT_STAT_EXT];
  r->stats.shadow_rays = cnt[CNT_STAT_SHADOW];
  r->stats.segments = cnt[CNT_STAT_EXT];
  r->stats.ms_total = ms;
  if (staged) {
    double stage_ms[3] = {0.0, 0.0, 0.0};
    for (size_t k = 0; k < n_ev; k += 2) {
      float e = 0.0f;
      APT_HIP(hipEventElapsedTime(&e, r->stage_ev[k], r->stage_ev[k + 1]));
      stage_ms[(k / 2) % 3] += e;
    }
    r->stats.ms_extend = stage_ms[0];
    r->stats.ms_shade = stage_ms[1];
    r->stats.ms_shadow = stage_ms[2];
    r->stats.ms_other = std::max(0.0, ms - stage_ms[0] - stage_ms[1] - stage_ms[2]);
  }
  r->accum_frames++;
  r->frame_idx++;
  return APT_OK;
}

bool parse_procedural(const char* path, std::string& name, uint32_t& seed) {
  const std::string p(path);
  const std::string pre = "procedural://";
  if (p.rfind(pre, 0) != 0) return false;
  std::string rest = p.substr(pre.size());
  seed = 0;
  const size_t q = rest.find('?');
  if (q != std::string::npos) {
    const std::string args = rest.substr(q + 1);
    rest = rest.substr(0, q);
    const size_t k = args.find("seed=");
    if (k != std::string::npos) seed = (uint32_t)std::strtoul(args.c_str() + k + 5, nullptr, 10);
  }
  name = rest;
  return !name.empty();
}

}  // namespace

extern "C" {

void apt_config_default(apt_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->max_bounces = 2;        // PathTracerGpuConfig::default (lib.rs:230-245)
  c->sample_count = 1;
  c->accum_frames = 0;
  c->rr_start_bounce = 3;    // plain integrator; trace.wgsl:221: roulette once bounce > 1
  c->integrator = APT_INTEGRATOR_RESTIR;
  c->restir_di = 1;
  c->restir_gi = 1;
  c->svgf = 0;
  c->firefly_filter = 1;
  c->taa = 0;
  c->di_candidates = 4;      // NUM_AREA_SAMPLES (nee.wgsl:125)
  c->di_mcap = 20;           // restir_di_temporal.wgsl:159
  c->di_spatial_neighbours = 3;  // NUM_SAMPLES (restir_di_spatial.wgsl:15)
  c->di_spatial_radius = 0.0f;   // 5% of (width + height) / 2 (restir_di_spatial.wgsl:96)
  c->gi_mcap = 30;           // restir_gi_temporal.wgsl:100
  c->svgf_iterations = 5;    // lib.rs:523
  c->svgf_history = 24;      // lib.rs:522
  c->bvh_width = 4;
  c->queue_chunk = 0;
  c->lds_stack_depth = 8;
  c->seed = 0;
  c->frame_idx = 0;
  c->gi_spatial_passes = 0;  // spatial_pass_count: 0 (lib.rs:441)
}

apt_status apt_create(const apt_config* cfg, apt_renderer** out) {
  clear_error();
  APT_REQUIRE(out, APT_ERR_INVALID_ARG, "apt_create: out is NULL");
  apt_config c;
  if (cfg) c = *cfg;
  else apt_config_default(&c);
  apt_status st = validate_config(c);
  if (st != APT_OK) return st;
  apt_renderer* r = new apt_renderer();
  r->cfg = c;
  r->frame_idx = c.frame_idx;
  // Default camera: at the origin looking down -Z, 60 degree fov.
  const float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  std::memcpy(r->cam.transform, id, sizeof(id));
  r->cam.fov_deg = 60.0f;
  r->cam.near_plane = 0.01f;
  r->cam.far_plane = 1000.0f;
  hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&r->ev0);
  if (e == hipSuccess) e = hipEventCreate(&r->ev1);
  if (e != hipSuccess) {
    set_error("apt_create: %s", hipGetErrorString(e));
    delete r;
    return APT_ERR_DEVICE;
  }
  *out = r;
  return APT_OK;
}

void apt_destroy(apt_renderer* r) { delete r; }

apt_status apt_set_config(apt_renderer* r, const apt_config* cfg) {
  clear_error();
  APT_REQUIRE(r && cfg, APT_ERR_INVALID_ARG, "apt_set_config: null argument");
  apt_status st = validate_config(*cfg);
  if (st != APT_OK) return st;
  if (cfg->bvh_width != r->cfg.bvh_width) r->scene_dirty = true;
  if (r->gpu.bvh) r->gpu.bvh->lds_stack = cfg->lds_stack_depth;
  if (r->tlas) r->tlas->lds_stack = cfg->lds_stack_depth;
  if (cfg->sample_count != r->cfg.sample_count || cfg->max_bounces != r->cfg.max_bounces ||
      cfg->rr_start_bounce != r->cfg.rr_start_bounce || cfg->seed != r->cfg.seed ||
      std::memcmp(&cfg->integrator, &r->cfg.integrator,
                  offsetof(apt_config, bvh_width) - offsetof(apt_config, integrator)) != 0)
    r->cam_version++;  // restart accumulation
  if (cfg->integrator != r->cfg.integrator) r->rs.npix = 0;  // no ReSTIR history across a switch
  r->cfg = *cfg;
  return APT_OK;
}

apt_status apt_world_action(apt_renderer* r, uint32_t ty, const void* bytes, size_t len) {
  clear_error();
  APT_REQUIRE(r && (bytes || len == 0), APT_ERR_INVALID_ARG, "apt_world_action: null argument");
  const uint8_t* b = static_cast<const uint8_t*>(bytes);
  switch (ty) {
    case 0: {  // CameraUpdate
      APT_REQUIRE(len == 80, APT_ERR_INVALID_ARG, "CameraUpdate carries 80 bytes, got %zu", len);
      std::memcpy(&r->cam, b, 80);
      r->cam_version++;
      return APT_OK;
    }
    case 1: {  // SpawnModel: Mat4, Uuid, 256-byte NUL-terminated asset path
      APT_REQUIRE(len == 336, APT_ERR_INVALID_ARG, "SpawnModel carries 336 bytes, got %zu", len);
      char path[257];
      std::memcpy(path, b + 80, 256);
      path[256] = 0;
      std::shared_ptr<Asset>& slot = r->assets[path];
      if (!slot) {
        auto a = std::make_shared<Asset>();
        std::string name;
        uint32_t seed;
        apt_status st = parse_procedural(path, name, seed) ? apt_scene_generate(name.c_str(), seed, &a->data)
                                                           : gltf_load(path, &a->data);
        if (st != APT_OK) {
          r->assets.erase(path);
          return st;
        }
        apt_mesh_desc md;
        apt_scene_data_view(a->data, &md, nullptr, nullptr);
        a->n_tris = md.n_tris;
        a->n_mats = md.n_materials;
        for (uint32_t k = 0; k < md.n_materials; ++k) {
          const float* e = md.materials[k].emission;
          a->emissive = a->emissive || e[0] > 0.0f || e[1] > 0.0f || e[2] > 0.0f;
        }
        slot = a;
      }
      Model m;
      std::memcpy(m.xform, b, 64);
      std::memcpy(m.uuid, b + 64, 16);
      m.asset = slot;
      apt_sky sky;
      apt_scene_data_view(m.asset->data, nullptr, nullptr, &sky);
      if (sky.sun_enabled > 0.0f || sky.zenith[0] > 0.0f || sky.zenith[1] > 0.0f || sky.zenith[2] > 0.0f)
        r->sky = sky;
      r->models.push_back(m);
      r->explicit_scene = false;
      r->scene_dirty = true;
      r->emitters_dirty = true;
      return APT_OK;
    }
    case 2: {  // TransformModel
      APT_REQUIRE(len == 80, APT_ERR_INVALID_ARG, "TransformModel carries 80 bytes, got %zu", len);
      for (Model& m : r->models) {
        if (std::memcmp(m.uuid, b + 64, 16) == 0) {
          std::memcpy(m.xform, b, 64);
          r->scene_dirty = true;
          if (m.asset->emissive) r->emitters_dirty = true;
          return APT_OK;
        }
      }
      set_error("TransformModel: unknown entity");
      return APT_ERR_NOT_FOUND;
    }
    case 3: {  // DestroyModel
      APT_REQUIRE(len == 16, APT_ERR_INVALID_ARG, "DestroyModel carries 16 bytes, got %zu", len);
      for (size_t i = 0; i < r->models.size(); ++i) {
        if (std::memcmp(r->models[i].uuid, b, 16) == 0) {
          r->models.erase(r->models.begin() + (long)i);
          r->scene_dirty = true;
          r->emitters_dirty = true;
          return APT_OK;
        }
      }
      set_error("DestroyModel: unknown entity");
      return APT_ERR_NOT_FOUND;
    }
    case 4: {  // Clear: the models and their assets (the reference drops its SceneModels, mod.rs:151-153)
      APT_REQUIRE(len == 4, APT_ERR_INVALID_ARG, "Clear carries 4 bytes, got %zu", len);
      r->models.clear();
      r->assets.clear();
      r->gpu.tlas = nullptr;
      apt_tlas_destroy(r->tlas);
      r->tlas = nullptr;
      r->explicit_scene = false;
      r->scene_dirty = true;
      r->emitters_dirty = true;
      return APT_OK;
    }
    default:
      set_error("apt_world_action: unknown action type %u", ty);
      return APT_ERR_INVALID_ARG;
  }
}

apt_status apt_scene_set(apt_renderer* r, const apt_mesh_desc* mesh, const apt_sky* sky) {
  return apt_scene_set_attribs(r, mesh, nullptr, sky);
}

apt_status apt_scene_set_attribs(apt_renderer* r, const apt_mesh_desc* mesh, const apt_mesh_attribs* attribs,
                                 const apt_sky* sky) {
  clear_error();
  APT_REQUIRE(r && mesh, APT_ERR_INVALID_ARG, "apt_scene_set: null argument");
  APT_REQUIRE(mesh->n_tris == 0 || mesh->positions, APT_ERR_INVALID_ARG, "apt_scene_set: no positions");
  APT_REQUIRE(mesh->n_materials > 0 && mesh->materials, APT_ERR_INVALID_ARG,
              "apt_scene_set: at least one material is required");
  const size_t n = mesh->n_tris;
  r->pos.assign(mesh->positions, mesh->positions + 9 * n);
  r->has_normals = mesh->normals != nullptr;
  if (r->has_normals) r->nrm.assign(mesh->normals, mesh->normals + 9 * n);
  else r->nrm.clear();
  if (mesh->material_ids) r->mat_ids.assign(mesh->material_ids, mesh->material_ids + n);
  else r->mat_ids.assign(n, 0u);
  r->mats.assign(mesh->materials, mesh->materials + mesh->n_materials);
  apt_mesh_attribs at{};
  if (attribs) at = *attribs;
  APT_REQUIRE(at.n_textures == 0 || at.textures, APT_ERR_INVALID_ARG, "apt_scene_set: n_textures without textures");
  for (const apt_material& m : r->mats)
    for (uint32_t t : {m.color_texture, m.metallic_roughness_texture, m.normal_texture, m.emission_texture,
                       m.sheen_texture, m.clearcoat_texture, m.clearcoat_roughness_texture, m.sheen_tint_texture,
                       m.clearcoat_normal_texture, m.transmission_texture})
      APT_REQUIRE(t == APT_INVALID_TEXTURE || t < at.n_textures, APT_ERR_INVALID_ARG,
                  "apt_scene_set: a material uses texture %u of %u", t, at.n_textures);
  r->uvs.assign(at.uvs ? at.uvs : nullptr, at.uvs ? at.uvs + 6 * n : nullptr);
  r->tangents.assign(at.tangents ? at.tangents : nullptr, at.tangents ? at.tangents + 12 * n : nullptr);
  r->bitangents.assign(at.bitangents ? at.bitangents : nullptr, at.bitangents ? at.bitangents + 9 * n : nullptr);
  r->textures.clear();
  for (uint32_t k = 0; k < at.n_textures; ++k) {
    const apt_texture& t = at.textures[k];
    APT_REQUIRE(t.width > 0 && t.height > 0 && t.rgba8, APT_ERR_INVALID_ARG, "apt_scene_set: texture %u is empty", k);
    apt_scene_texture c;
    c.width = t.width;
    c.height = t.height;
    c.rgba.assign(t.rgba8, t.rgba8 + (size_t)t.width * t.height);
    r->textures.push_back(std::move(c));
  }
  if (sky) r->sky = *sky;
  else std::memset(&r->sky, 0, sizeof(r->sky));
  r->explicit_scene = true;
  r->scene_dirty = true;
  return APT_OK;
}

apt_status apt_set_sky_panorama(apt_renderer* r, uint32_t width, uint32_t height, const float* rgb) {
  clear_error();
  APT_REQUIRE(r, APT_ERR_INVALID_ARG, "apt_set_sky_panorama: renderer is NULL");
  if (!rgb || width == 0 || height == 0) {
    r->d_sky_pano.release();
    r->pano_w = r->pano_h = 0;
    r->scene_version++;
    return APT_OK;
  }
  APT_REQUIRE(width <= APT_SKY_PANORAMA_MAX && height <= APT_SKY_PANORAMA_MAX, APT_ERR_INVALID_ARG,
              "apt_set_sky_panorama: %ux%u exceeds %u texels a side", width, height, APT_SKY_PANORAMA_MAX);
  const size_t n = (size_t)width * height;
  std::vector<float4> tex(n);
  for (size_t i = 0; i < n; ++i) {
    const float c[3] = {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
    for (float v : c)
      APT_REQUIRE(std::isfinite(v) && v >= 0.0f, APT_ERR_INVALID_ARG,
                  "apt_set_sky_panorama: texel %zu is negative or not finite", i);
    tex[i] = make_float4(c[0], c[1], c[2], 0.0f);
  }
  apt_status st = r->d_sky_pano.ensure(n);
  if (st != APT_OK) return st;
  APT_HIP(hipMemcpy(r->d_sky_pano.p, tex.data(), n * sizeof(float4), hipMemcpyHostToDevice));
  r->pano_w = width;
  r->pano_h = height;
  r->scene_version++;  // a new sky: accumulation and reservoir history restart
  return APT_OK;
}

apt_status apt_render(apt_renderer* r, uint32_t width, uint32_t height, uint32_t row_start,
                      uint32_t row_end, uint8_t* out_rgba, size_t out_len, float* out_accum_rgb,
                      size_t accum_len) {
  clear_error();
  APT_REQUIRE(r, APT_ERR_INVALID_ARG, "apt_render: renderer is NULL");
  return render_impl(r, width, height, row_start, row_end, out_rgba, out_len, out_accum_rgb, accum_len);
}

apt_status apt_render_cb(apt_renderer* r, uint32_t width, uint32_t height, uint32_t row_start,
                         uint32_t row_end, apt_result_cb cb, void* user) {
  clear_error();
  APT_REQUIRE(r && cb, APT_ERR_INVALID_ARG, "apt_render_cb: null argument");
  APT_REQUIRE(row_end > row_start, APT_ERR_INVALID_ARG, "apt_render_cb: empty slab");
  const size_t len = (size_t)width * (row_end - row_start) * 4;
  apt_status st = r->cb_pixels.ensure(len);
  if (st != APT_OK) return st;
  st = render_impl(r, width, height, row_start, row_end, r->cb_pixels.p, len, nullptr, 0);
  if (st != APT_OK) return st;
  cb(r->cb_pixels.p, len, user);
  return APT_OK;
}

namespace {

// StartRenderData as it arrives (host.rs:85-93), checked to be a slab of
// whole 64x64 blocks (RENDER_BLOCK_SIZE, host.rs:17).
apt_status parse_start(const void* data, size_t len, const char* who, apt_start_render_data& d) {
  APT_REQUIRE(len == sizeof(apt_start_render_data), APT_ERR_INVALID_ARG,
              "%s: StartRenderData carries %zu bytes, got %zu", who, sizeof(apt_start_render_data), len);
  std::memcpy(&d, data, sizeof(d));
  APT_REQUIRE(d.row_end > d.row_start && d.row_end <= d.height && d.width % 64 == 0 &&
                  (d.row_end - d.row_start) % 64 == 0,
              APT_ERR_INVALID_ARG, "%s: %ux%u rows [%u,%u) is not a slab of whole 64x64 blocks", who, d.width,
              d.height, d.row_start, d.row_end);
  return APT_OK;
}

}  // namespace

apt_status apt_start_render_jpeg(apt_renderer* r, const void* start_render_data, size_t len, int quality,
                                 apt_block_cb cb, void* user) {
  clear_error();
  APT_REQUIRE(r && start_render_data && cb, APT_ERR_INVALID_ARG, "apt_start_render_jpeg: null argument");
  APT_REQUIRE(quality >= 1 && quality <= 100, APT_ERR_INVALID_ARG, "apt_start_render_jpeg: quality %d outside 1..100",
              quality);
  apt_start_render_data d;
  apt_status st = parse_start(start_render_data, len, "apt_start_render_jpeg", d);
  if (st != APT_OK) return st;
  constexpr uint32_t B = 64;
  const uint32_t nx = d.width / B, ny = (d.row_end - d.row_start) / B;
  APT_REQUIRE((size_t)nx * ny <= JPEG_MAX_BLOCKS, APT_ERR_INVALID_ARG,
              "apt_start_render_jpeg: %u blocks exceed %u", nx * ny, JPEG_MAX_BLOCKS);
  // the frame stays in HBM: resolved to RGBA8 blocks there, compressed there,
  // and only the files cross to the host
  r->rgba_on_device = true;
  st = render_impl(r, d.width, d.height, d.row_start, d.row_end, nullptr, 0, nullptr, 0);
  r->rgba_on_device = false;
  if (st != APT_OK) return st;
  if ((st = r->jpeg.encode(r->d_rgba.p, nx * ny, quality, r->stream)) != APT_OK) return st;
  for (uint32_t by = 0; by < ny; ++by)
    for (uint32_t bx = 0; bx < nx; ++bx) {
      const uint32_t b = by * nx + bx;
      const uint32_t first = r->jpeg.offsets[b];
      cb(by * B + d.row_start, bx, d.frame_idx, r->jpeg.h_files + first, r->jpeg.offsets[b + 1] - first, user);
    }
  return APT_OK;
}

apt_status apt_start_render(apt_renderer* r, const void* start_render_data, size_t len, apt_block_cb cb,
                            void* user) {
  clear_error();
  APT_REQUIRE(r && start_render_data && cb, APT_ERR_INVALID_ARG, "apt_start_render: null argument");
  apt_start_render_data d;
  apt_status st = parse_start(start_render_data, len, "apt_start_render", d);
  if (st != APT_OK) return st;
  constexpr uint32_t B = 64;  // RENDER_BLOCK_SIZE
  const size_t bytes = (size_t)d.width * (d.row_end - d.row_start) * 4;
  st = r->cb_pixels.ensure(bytes);
  if (st != APT_OK) return st;
  st = render_impl(r, d.width, d.height, d.row_start, d.row_end, r->cb_pixels.p, bytes, nullptr, 0);
  if (st != APT_OK) return st;
  // Node::start_render's slicing (node.rs:55-75): block (x, y) of the slab
  // starts at pixel (y * nx + x) * 64 * 64 of the block-major frame.
  const uint32_t nx = d.width / B, ny = (d.row_end - d.row_start) / B;
  for (uint32_t by = 0; by < ny; ++by)
    for (uint32_t bx = 0; bx < nx; ++bx) {
      const size_t first = ((size_t)by * nx + bx) * B * B;
      cb(by * B + d.row_start, bx, d.frame_idx, r->cb_pixels.p + first * 4, (size_t)B * B * 4, user);
    }
  return APT_OK;
}

apt_status apt_set_frame_index(apt_renderer* r, uint32_t frame_idx) {
  clear_error();
  APT_REQUIRE(r, APT_ERR_INVALID_ARG, "apt_set_frame_index: renderer is NULL");
  r->frame_idx = frame_idx;
  r->cam_version++;  // a new sequence restarts accumulation
  return APT_OK;
}

apt_status apt_get_render_stats(const apt_renderer* r, apt_render_stats* out) {
  APT_REQUIRE(r && out, APT_ERR_INVALID_ARG, "apt_get_render_stats: null argument");
  *out = r->stats;
  return APT_OK;
}

}  // extern "C"
