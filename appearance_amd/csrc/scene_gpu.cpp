// Scene upload shared by the renderer and the server.
#include <cmath>
#include <cstring>
#include <vector>

#include "scene_gpu.h"

namespace apt {

namespace {

template <typename T>
apt_status upload(T*& dst, const T* src, size_t n, size_t& bytes) {
  const size_t b = std::max<size_t>(1, n) * sizeof(T);
  hipError_t e = hipMalloc(&dst, b);
  if (e != hipSuccess) {
    dst = nullptr;
    set_error("scene upload: hipMalloc(%zu B) failed: %s", b, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? APT_ERR_OOM : APT_ERR_DEVICE;
  }
  bytes += b;
  if (n) APT_HIP(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return APT_OK;
}

// xyz triples (3 per triangle) -> float4 with w = 0.
__global__ void k_widen3(const float* __restrict__ xyz, float4* __restrict__ out, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[i] = make_float4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], 0.0f);
}

// Uploads n vertices (xyz triples) once and widens them to float4 on the
// device.  Returns the device copy of the triples in *keep_xyz if asked
// (the caller frees it), else frees it.
apt_status upload_widened(float4*& dst, const float* src, size_t n, size_t& bytes, float** keep_xyz) {
  float* d_xyz = nullptr;
  apt_status st = upload(d_xyz, src, 3 * n, bytes);
  if (st != APT_OK) return st;
  bytes -= std::max<size_t>(1, 3 * n) * sizeof(float);  // staging, not resident
  hipError_t e = hipMalloc(&dst, std::max<size_t>(1, n) * sizeof(float4));
  if (e != hipSuccess) {
    dst = nullptr;
    (void)hipFree(d_xyz);
    set_error("scene upload: hipMalloc(%zu B) failed: %s", n * sizeof(float4), hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? APT_ERR_OOM : APT_ERR_DEVICE;
  }
  bytes += std::max<size_t>(1, n) * sizeof(float4);
  if (n) {
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_widen3, dim3(blocks), dim3(256), 0, nullptr, d_xyz, dst, (uint32_t)n);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
      (void)hipFree(d_xyz);
      set_error("scene upload: widening failed: %s", hipGetErrorString(e));
      return APT_ERR_DEVICE;
    }
  }
  if (keep_xyz) *keep_xyz = d_xyz;
  else (void)hipFree(d_xyz);
  return APT_OK;
}

}  // namespace

apt_status scene_gpu_upload(const float* pos, const float* nrm, const uint32_t* mat_ids, uint32_t n,
                            const apt_material* mats, uint32_t n_mats, const apt_sky* sky,
                            uint32_t bvh_width, uint32_t lds_stack, SceneGpu& out) {
  APT_REQUIRE(n > 0 && pos && mats && n_mats > 0, APT_ERR_INVALID_ARG,
              "scene: needs triangles and at least one material");
  for (uint32_t t = 0; t < n; ++t) {
    const uint32_t m = mat_ids ? mat_ids[t] : 0u;
    APT_REQUIRE(m < n_mats, APT_ERR_INVALID_ARG, "scene: triangle %u uses material %u of %u", t, m, n_mats);
  }
  SceneGpu s;
  s.n_tris = n;
  s.n_mats = n_mats;
  if (sky) s.sky = *sky;
  // The device builder gives the host builder's bytes about 13x sooner
  // (1M triangles: 13 ms against 175 ms), which every spawn, move or removal
  // of a model pays.
  apt_bvh_opts o{bvh_width, 4, 32, APT_BUILDER_GPU, lds_stack};
  // The positions cross PCIe once: the builder reads the uploaded triples.
  float* d_xyz = nullptr;
  apt_status st = upload_widened(s.d_pos, pos, (size_t)n * 3, s.bytes, &d_xyz);
  if (st != APT_OK) { scene_gpu_free(s); return st; }
  st = apt_bvh_build_device(d_xyz, n, &o, &s.bvh);
  (void)hipFree(d_xyz);
  if (st != APT_OK) { scene_gpu_free(s); return st; }
  s.bytes += s.bvh->stats.device_bytes;
  if (nrm && (st = upload_widened(s.d_nrm, nrm, (size_t)n * 3, s.bytes, nullptr)) != APT_OK) {
    scene_gpu_free(s);
    return st;
  }
  std::vector<uint32_t> ids(mat_ids ? mat_ids : nullptr, mat_ids ? mat_ids + n : nullptr);
  if (!mat_ids) ids.assign(n, 0u);
  if ((st = upload(s.d_mat, ids.data(), ids.size(), s.bytes)) != APT_OK) { scene_gpu_free(s); return st; }
  if ((st = upload(s.d_mats, mats, n_mats, s.bytes)) != APT_OK) { scene_gpu_free(s); return st; }
  std::vector<uint32_t> em;
  for (uint32_t t = 0; t < n; ++t) {
    const apt_material& m = mats[ids[t]];
    if (!(m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f)) continue;
    const float* p = pos + 9 * (size_t)t;
    const float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]}, e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
    const float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    if (c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > 0.0f) em.push_back(t);  // degenerate emitters emit nothing
  }
  s.n_emissive = (uint32_t)em.size();
  if ((st = upload(s.d_emissive, em.data(), em.size(), s.bytes)) != APT_OK) { scene_gpu_free(s); return st; }
  out = s;
  return APT_OK;
}

void scene_gpu_free(SceneGpu& s) {
  if (s.bvh) apt_bvh_destroy(s.bvh);
  for (void* p : {(void*)s.d_pos, (void*)s.d_nrm, (void*)s.d_mat, (void*)s.d_mats, (void*)s.d_emissive})
    if (p) (void)hipFree(p);
  s = SceneGpu();
}

SceneView scene_gpu_view(const SceneGpu& s, const apt_camera& cam, uint32_t width, uint32_t height) {
  SceneView sc{};
  sc.nodes = s.bvh ? s.bvh->d_nodes : nullptr;
  sc.bvh_tris = s.bvh ? s.bvh->d_tris : nullptr;
  sc.root = s.bvh ? s.bvh->root : 0;
  sc.n_nodes = s.bvh ? s.bvh->n_nodes : 0;
  sc.width = s.bvh ? s.bvh->width : 4;
  sc.lds_stack = s.bvh ? s.bvh->lds_stack : 8;
  sc.refill = s.bvh ? s.bvh->refill : 32;
  sc.tri_pos = s.d_pos;
  sc.tri_nrm = s.d_nrm;
  sc.tri_mat = s.d_mat;
  sc.mats = s.d_mats;
  sc.emissive = s.d_emissive;
  sc.n_emissive = s.n_emissive;
  sc.sky = s.sky;
  // CameraUpdate: rotation and translation of the world transform (scale
  // dropped, as to_scale_rotation_translation does in lib.rs:289-296), a
  // vertical-fov pinhole with aspect = full width / height (lib.rs:333-334).
  const float* m = cam.transform;
  sc.cam_right = normalize(mk(m[0], m[1], m[2]));
  sc.cam_up = normalize(mk(m[4], m[5], m[6]));
  sc.cam_back = normalize(mk(m[8], m[9], m[10]));
  sc.cam_pos = mk(m[12], m[13], m[14]);
  sc.cam_tan_half = std::tan(0.5f * cam.fov_deg * PI / 180.0f);
  sc.cam_aspect = (float)width / (float)height;
  return sc;
}

}  // namespace apt
