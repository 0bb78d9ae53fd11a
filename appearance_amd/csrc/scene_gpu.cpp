// Scene upload shared by the renderer and the server.
#include <cmath>
#include <cstring>
#include <vector>

#include "scene_gpu.h"
#include "tlas.h"

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

// Emitters: the material emits and the edges' cross product is not zero (a
// degenerate emitter emits nothing).  The same float operations as a host
// loop (no contraction), so the list does not depend on where it is made.
__global__ void k_emit_flags(const float* __restrict__ xyz, const uint32_t* __restrict__ mat,
                             const apt_material* __restrict__ mats, uint8_t* __restrict__ flag, uint32_t n) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const apt_material& m = mats[mat[t]];
    bool e = m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f;
    if (e) {
      const float* p = xyz + 9 * (size_t)t;
      const float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]}, e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
      const float c[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      e = c[0] * c[0] + c[1] * c[1] + c[2] * c[2] > 0.0f;
    }
    flag[t] = e ? 1 : 0;
  }
}

// A model's vertices in world space: positions M p (((m0 x + m4 y) + m8 z) +
// m12 per row, column-major M), normals through the inverse transpose of its
// 3x3 part and renormalised, material ids offset by the model's material base.
// A vector through the normal matrix (or kept), renormalised.
__device__ __forceinline__ float4 place_dir(const ModelXform& x, float q0, float q1, float q2) {
  float v[3] = {q0, q1, q2};
  if (x.nm_ok)
    for (int r = 0; r < 3; ++r) v[r] = x.nm[3 * r] * q0 + x.nm[3 * r + 1] * q1 + x.nm[3 * r + 2] * q2;
  const float l = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (l > 0.0f) for (int r = 0; r < 3; ++r) v[r] /= l;
  return make_float4(v[0], v[1], v[2], 0.0f);
}

// A model's vertices in world space: positions M p (((m0 x + m4 y) + m8 z) +
// m12 per row, column-major M), normals, tangents and bitangents through the
// inverse transpose of its 3x3 part and renormalised (as the reference maps
// the interpolated frame, trace.wgsl:143-156), uvs as they are, material ids
// offset by the model's material base.  Arrays the scene does not keep are
// null; a model without an attribute the scene keeps writes zeros.
__global__ void k_place(const float* __restrict__ lxyz, const float* __restrict__ lnrm,
                        const uint32_t* __restrict__ lmat, uint32_t n, ModelXform x, ModelAttribs at,
                        float4* __restrict__ opos, float4* __restrict__ onrm, float2* __restrict__ ouv,
                        float4* __restrict__ otan, float4* __restrict__ obit, uint32_t* __restrict__ omat) {
  const uint32_t nv = 3 * n;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nv; j += gridDim.x * blockDim.x) {
    const float* p = lxyz + 3 * (size_t)j;
    float o[3];
    for (int r = 0; r < 3; ++r) o[r] = x.m[r] * p[0] + x.m[4 + r] * p[1] + x.m[8 + r] * p[2] + x.m[12 + r];
    opos[j] = make_float4(o[0], o[1], o[2], 0.0f);
    if (onrm) {
      const float* q = lnrm + 3 * (size_t)j;
      onrm[j] = place_dir(x, q[0], q[1], q[2]);
    }
    if (ouv) ouv[j] = at.uv ? at.uv[j] : make_float2(0.0f, 0.0f);
    if (otan) {
      if (at.tan && lnrm) {
        const float4 t = at.tan[j];
        const float* q = lnrm + 3 * (size_t)j;
        float4 b;
        if (at.bit) {
          b = at.bit[j];
        } else {  // cross(n, t) * w in object space (vertex_pool_bindings.wgsl:43-46)
          b = make_float4((q[1] * t.z - q[2] * t.y) * t.w, (q[2] * t.x - q[0] * t.z) * t.w,
                          (q[0] * t.y - q[1] * t.x) * t.w, 0.0f);
        }
        otan[j] = place_dir(x, t.x, t.y, t.z);
        obit[j] = place_dir(x, b.x, b.y, b.z);
      } else {
        otan[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        obit[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
    }
    if (j < n) omat[j] = x.mat_base + (lmat ? lmat[j] : 0u);
  }
}

// k_emit_flags on float4 positions (the same float operations).
__global__ void k_emit_flags4(const float4* __restrict__ pos, const uint32_t* __restrict__ mat,
                              const apt_material* __restrict__ mats, uint8_t* __restrict__ flag, uint32_t n) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const apt_material& m = mats[mat[t]];
    bool e = m.emission[0] > 0.0f || m.emission[1] > 0.0f || m.emission[2] > 0.0f;
    if (e) {
      const float4 a = pos[3 * (size_t)t], b = pos[3 * (size_t)t + 1], c = pos[3 * (size_t)t + 2];
      const float e1[3] = {b.x - a.x, b.y - a.y, b.z - a.z}, e2[3] = {c.x - a.x, c.y - a.y, c.z - a.z};
      const float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                           e1[0] * e2[1] - e1[1] * e2[0]};
      e = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2] > 0.0f;
    }
    flag[t] = e ? 1 : 0;
  }
}

uint32_t grid_for(size_t n) { return (uint32_t)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096)); }

apt_status sync_null(const char* what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    set_error("scene upload: %s failed: %s", what, hipGetErrorString(e));
    return APT_ERR_DEVICE;
  }
  return APT_OK;
}

apt_status alloc_dev(void** p, size_t b) {
  hipError_t e = hipMalloc(p, std::max<size_t>(1, b));
  if (e != hipSuccess) {
    *p = nullptr;
    set_error("scene upload: hipMalloc(%zu B) failed: %s", b, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? APT_ERR_OOM : APT_ERR_DEVICE;
  }
  return APT_OK;
}

// n vertices (xyz triples in device memory) -> a new float4 buffer.
apt_status widen(const float* d_xyz, size_t n, float4*& dst, size_t& bytes) {
  apt_status st = alloc_dev((void**)&dst, n * sizeof(float4));
  if (st != APT_OK) return st;
  bytes += std::max<size_t>(1, n * sizeof(float4));
  if (n) hipLaunchKernelGGL(k_widen3, dim3(grid_for(n)), dim3(256), 0, nullptr, d_xyz, dst, (uint32_t)n);
  return sync_null("widening");
}

}  // namespace

apt_status scene_gpu_upload_device(const float* d_xyz, const float* d_nrm, const uint32_t* d_mat_ids,
                                   uint32_t n, const apt_material* mats, uint32_t n_mats,
                                   const apt_sky* sky, uint32_t bvh_width, uint32_t lds_stack,
                                   SceneGpu& out) {
  APT_REQUIRE(n > 0 && d_xyz && d_mat_ids && mats && n_mats > 0, APT_ERR_INVALID_ARG,
              "scene: needs triangles and at least one material");
  SceneGpu s;
  s.n_tris = n;
  s.n_mats = n_mats;
  if (sky) s.sky = *sky;
  // The device builder gives the host builder's bytes about 13x sooner
  // (1M triangles: 13 ms against 175 ms), which every spawn, move or removal
  // of a model pays.
  apt_bvh_opts o{bvh_width, 4, 32, APT_BUILDER_GPU, lds_stack};
  apt_status st = widen(d_xyz, (size_t)n * 3, s.d_pos, s.bytes);
  if (st == APT_OK && (st = apt_bvh_build_device(d_xyz, n, &o, &s.bvh)) == APT_OK)
    s.bytes += s.bvh->stats.device_bytes;
  // the renderer, the server and the autotuner trace every scene through here
  if (st == APT_OK) st = check_trace_stack(s.bvh->stats, "scene");
  if (st == APT_OK && d_nrm) st = widen(d_nrm, (size_t)n * 3, s.d_nrm, s.bytes);
  if (st == APT_OK && (st = alloc_dev((void**)&s.d_mat, (size_t)n * sizeof(uint32_t))) == APT_OK) {
    s.bytes += (size_t)n * sizeof(uint32_t);
    if (hipMemcpy(s.d_mat, d_mat_ids, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToDevice) != hipSuccess)
      st = sync_null("material id copy");
  }
  if (st == APT_OK) st = upload(s.d_mats, mats, n_mats, s.bytes);
  std::vector<uint8_t> flag(n);
  if (st == APT_OK) {
    uint8_t* d_flag = nullptr;
    st = alloc_dev((void**)&d_flag, n);
    if (st == APT_OK) {
      hipLaunchKernelGGL(k_emit_flags, dim3(grid_for(n)), dim3(256), 0, nullptr, d_xyz, s.d_mat, s.d_mats,
                         d_flag, n);
      st = sync_null("emitter flags");
      if (st == APT_OK && hipMemcpy(flag.data(), d_flag, n, hipMemcpyDeviceToHost) != hipSuccess)
        st = sync_null("emitter flag readback");
      (void)hipFree(d_flag);
    }
  }
  if (st != APT_OK) {
    scene_gpu_free(s);
    return st;
  }
  std::vector<uint32_t> em;
  for (uint32_t t = 0; t < n; ++t)
    if (flag[t]) em.push_back(t);
  s.n_emissive = (uint32_t)em.size();
  if ((st = upload(s.d_emissive, em.data(), em.size(), s.bytes)) != APT_OK) { scene_gpu_free(s); return st; }
  out = s;
  return APT_OK;
}

apt_status scene_gpu_upload(const float* pos, const float* nrm, const uint32_t* mat_ids, uint32_t n,
                            const apt_material* mats, uint32_t n_mats, const apt_sky* sky,
                            uint32_t bvh_width, uint32_t lds_stack, SceneGpu& out) {
  APT_REQUIRE(n > 0 && pos && mats && n_mats > 0, APT_ERR_INVALID_ARG,
              "scene: needs triangles and at least one material");
  for (uint32_t t = 0; t < n; ++t) {
    const uint32_t m = mat_ids ? mat_ids[t] : 0u;
    APT_REQUIRE(m < n_mats, APT_ERR_INVALID_ARG, "scene: triangle %u uses material %u of %u", t, m, n_mats);
  }
  // Each array crosses PCIe once; the device copy feeds the build and the
  // attribute buffers.
  std::vector<uint32_t> ids;
  if (!mat_ids) ids.assign(n, 0u);
  float *d_xyz = nullptr, *d_nrm = nullptr;
  uint32_t* d_ids = nullptr;
  size_t staged = 0;
  apt_status st = upload(d_xyz, pos, (size_t)n * 9, staged);
  if (st == APT_OK && nrm) st = upload(d_nrm, nrm, (size_t)n * 9, staged);
  if (st == APT_OK) st = upload(d_ids, mat_ids ? mat_ids : ids.data(), (size_t)n, staged);
  if (st == APT_OK)
    st = scene_gpu_upload_device(d_xyz, d_nrm, d_ids, n, mats, n_mats, sky, bvh_width, lds_stack, out);
  for (void* p : {(void*)d_xyz, (void*)d_nrm, (void*)d_ids})
    if (p) (void)hipFree(p);
  return st;
}

apt_status scene_place_model(const float* d_xyz, const float* d_nrm, const uint32_t* d_mat, uint32_t n,
                             const ModelXform& x, const ModelAttribs& at, SceneGpu& s, uint32_t base,
                             hipStream_t st) {
  if (n == 0) return APT_OK;
  APT_REQUIRE((size_t)base + n <= s.cap_tris, APT_ERR_INVALID_ARG, "scene placement: prims [%u, %u) past %zu",
              base, base + n, s.cap_tris);
  APT_REQUIRE(!s.d_nrm || d_nrm, APT_ERR_INVALID_ARG, "scene placement: the model has no normals");
  const size_t v0 = 3 * (size_t)base;
  hipLaunchKernelGGL(k_place, dim3(grid_for((size_t)n * 3)), dim3(256), 0, st, d_xyz, d_nrm, d_mat, n, x, at,
                     s.d_pos + v0, s.d_nrm ? s.d_nrm + v0 : nullptr, s.d_uv ? s.d_uv + v0 : nullptr,
                     s.d_tan ? s.d_tan + v0 : nullptr, s.d_tan ? s.d_bit + v0 : nullptr, s.d_mat + base);
  APT_HIP(hipGetLastError());
  return APT_OK;
}

apt_status scene_reserve(SceneGpu& s, uint32_t n, bool with_normals, bool with_uv, bool with_tangents,
                         const apt_material* mats, uint32_t n_mats, const TexDesc* tex, uint32_t n_tex,
                         hipStream_t st) {
  const bool have_uv = s.d_uv != nullptr, have_tan = s.d_tan != nullptr;
  if (n > s.cap_tris || (with_normals && !s.d_nrm) || with_uv != have_uv || with_tangents != have_tan) {
    for (void* p : {(void*)s.d_pos, (void*)s.d_nrm, (void*)s.d_mat, (void*)s.d_uv, (void*)s.d_tan, (void*)s.d_bit})
      if (p) (void)hipFree(p);
    s.d_pos = s.d_nrm = s.d_tan = s.d_bit = nullptr;
    s.d_uv = nullptr;
    s.d_mat = nullptr;
    s.cap_tris = 0;
    const size_t cap = std::max<size_t>(n, 1);
    apt_status r = alloc_dev((void**)&s.d_pos, cap * 3 * sizeof(float4));
    if (r == APT_OK && with_normals) r = alloc_dev((void**)&s.d_nrm, cap * 3 * sizeof(float4));
    if (r == APT_OK && with_uv) r = alloc_dev((void**)&s.d_uv, cap * 3 * sizeof(float2));
    if (r == APT_OK && with_tangents) r = alloc_dev((void**)&s.d_tan, cap * 3 * sizeof(float4));
    if (r == APT_OK && with_tangents) r = alloc_dev((void**)&s.d_bit, cap * 3 * sizeof(float4));
    if (r == APT_OK) r = alloc_dev((void**)&s.d_mat, cap * sizeof(uint32_t));
    if (r != APT_OK) return r;
    s.cap_tris = cap;
  } else if (!with_normals && s.d_nrm) {
    (void)hipFree(s.d_nrm);
    s.d_nrm = nullptr;
  }
  if (n_mats > s.cap_mats) {
    if (s.d_mats) (void)hipFree(s.d_mats);
    s.d_mats = nullptr;
    s.cap_mats = 0;
    apt_status r = alloc_dev((void**)&s.d_mats, (size_t)n_mats * sizeof(apt_material));
    if (r != APT_OK) return r;
    s.cap_mats = n_mats;
  }
  if (n_mats) APT_HIP(hipMemcpyAsync(s.d_mats, mats, (size_t)n_mats * sizeof(apt_material), hipMemcpyHostToDevice, st));
  if (n_tex > s.cap_tex) {
    if (s.d_tex) (void)hipFree(s.d_tex);
    s.d_tex = nullptr;
    s.cap_tex = 0;
    apt_status r = alloc_dev((void**)&s.d_tex, (size_t)n_tex * sizeof(TexDesc));
    if (r != APT_OK) return r;
    s.cap_tex = n_tex;
  }
  if (n_tex) APT_HIP(hipMemcpyAsync(s.d_tex, tex, (size_t)n_tex * sizeof(TexDesc), hipMemcpyHostToDevice, st));
  s.n_tex = n_tex;
  s.n_tris = n;
  s.n_mats = n_mats;
  s.bytes = s.cap_tris * (3 * sizeof(float4) * (1 + (s.d_nrm ? 1 : 0) + (s.d_tan ? 2 : 0)) +
                          (s.d_uv ? 3 * sizeof(float2) : 0) + sizeof(uint32_t)) +
            s.cap_mats * sizeof(apt_material) + s.cap_emissive * sizeof(uint32_t) + s.cap_tex * sizeof(TexDesc);
  return APT_OK;
}

apt_status scene_gpu_attribs(SceneGpu& s, const float* pos, const float* nrm, uint32_t n,
                             const apt_mesh_attribs& a) {
  const size_t nv = 3 * (size_t)n;
  if (a.uvs) {
    std::vector<float2> uv(nv);
    for (size_t j = 0; j < nv; ++j) uv[j] = make_float2(a.uvs[2 * j], a.uvs[2 * j + 1]);
    apt_status r = upload(s.d_uv, uv.data(), nv, s.bytes);
    if (r != APT_OK) return r;
  }
  if (a.tangents) {
    std::vector<float4> t(nv), b(nv);
    for (size_t j = 0; j < nv; ++j) {
      const float* q = a.tangents + 4 * j;
      float n3[3];
      if (nrm) {
        for (int k = 0; k < 3; ++k) n3[k] = nrm[3 * j + k];
      } else {  // the face normal
        const float* p = pos + 9 * (j / 3);
        const float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]}, e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
        n3[0] = e1[1] * e2[2] - e1[2] * e2[1];
        n3[1] = e1[2] * e2[0] - e1[0] * e2[2];
        n3[2] = e1[0] * e2[1] - e1[1] * e2[0];
      }
      float bx, by, bz;
      if (a.bitangents) {
        bx = a.bitangents[3 * j];
        by = a.bitangents[3 * j + 1];
        bz = a.bitangents[3 * j + 2];
      } else {
        bx = (n3[1] * q[2] - n3[2] * q[1]) * q[3];
        by = (n3[2] * q[0] - n3[0] * q[2]) * q[3];
        bz = (n3[0] * q[1] - n3[1] * q[0]) * q[3];
      }
      const float lt = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
      const float lb = std::sqrt(bx * bx + by * by + bz * bz);
      t[j] = lt > 0.0f ? make_float4(q[0] / lt, q[1] / lt, q[2] / lt, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      b[j] = lb > 0.0f ? make_float4(bx / lb, by / lb, bz / lb, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    apt_status r = upload(s.d_tan, t.data(), nv, s.bytes);
    if (r == APT_OK) r = upload(s.d_bit, b.data(), nv, s.bytes);
    if (r != APT_OK) return r;
  }
  if (a.n_textures) {
    APT_REQUIRE(a.textures, APT_ERR_INVALID_ARG, "scene: n_textures without textures");
    size_t total = 0;
    for (uint32_t k = 0; k < a.n_textures; ++k) {
      const apt_texture& t = a.textures[k];
      APT_REQUIRE(t.width > 0 && t.height > 0 && t.rgba8, APT_ERR_INVALID_ARG, "scene: texture %u is empty", k);
      total += (size_t)t.width * t.height;
    }
    std::vector<uint32_t> texels;
    texels.reserve(total);
    std::vector<size_t> off(a.n_textures);
    for (uint32_t k = 0; k < a.n_textures; ++k) {
      const apt_texture& t = a.textures[k];
      off[k] = texels.size();
      texels.insert(texels.end(), t.rgba8, t.rgba8 + (size_t)t.width * t.height);
    }
    apt_status r = upload(s.d_texels, texels.data(), texels.size(), s.bytes);
    if (r != APT_OK) return r;
    std::vector<TexDesc> d(a.n_textures);
    for (uint32_t k = 0; k < a.n_textures; ++k)
      d[k] = TexDesc{s.d_texels + off[k], a.textures[k].width, a.textures[k].height};
    r = upload(s.d_tex, d.data(), d.size(), s.bytes);
    if (r != APT_OK) return r;
    s.n_tex = a.n_textures;
  }
  return APT_OK;
}

apt_status scene_emitters(SceneGpu& s, hipStream_t st) {
  const uint32_t n = s.n_tris;
  std::vector<uint8_t> flag(n);
  if (n) {
    uint8_t* d_flag = nullptr;
    apt_status r = alloc_dev((void**)&d_flag, n);
    if (r != APT_OK) return r;
    hipLaunchKernelGGL(k_emit_flags4, dim3(grid_for(n)), dim3(256), 0, st, s.d_pos, s.d_mat, s.d_mats, d_flag, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(flag.data(), d_flag, n, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_flag);
    if (e != hipSuccess) {
      set_error("scene emitters: %s", hipGetErrorString(e));
      return APT_ERR_DEVICE;
    }
  }
  std::vector<uint32_t> em;
  for (uint32_t t = 0; t < n; ++t)
    if (flag[t]) em.push_back(t);
  if (em.size() > s.cap_emissive || !s.d_emissive) {
    if (s.d_emissive) (void)hipFree(s.d_emissive);
    s.d_emissive = nullptr;
    s.cap_emissive = 0;
    apt_status r = alloc_dev((void**)&s.d_emissive, std::max<size_t>(1, em.size()) * sizeof(uint32_t));
    if (r != APT_OK) return r;
    s.cap_emissive = std::max<size_t>(1, em.size());
  }
  if (!em.empty())
    APT_HIP(hipMemcpyAsync(s.d_emissive, em.data(), em.size() * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  APT_HIP(hipStreamSynchronize(st));
  s.n_emissive = (uint32_t)em.size();
  return APT_OK;
}

void scene_gpu_free(SceneGpu& s) {
  if (s.bvh) apt_bvh_destroy(s.bvh);
  for (void* p : {(void*)s.d_pos, (void*)s.d_nrm, (void*)s.d_mat, (void*)s.d_mats, (void*)s.d_emissive,
                  (void*)s.d_uv, (void*)s.d_tan, (void*)s.d_bit, (void*)s.d_tex, (void*)s.d_texels})
    if (p) (void)hipFree(p);
  s = SceneGpu();
}

SceneView scene_gpu_view(const SceneGpu& s, const apt_camera& cam, uint32_t width, uint32_t height) {
  SceneView sc{};
  if (s.tlas) {
    sc.nodes = s.tlas->d_nodes;
    sc.bvh_tris = s.tlas->d_tris;
    sc.root = s.tlas->root;
    sc.n_nodes = s.tlas->n_tlas_nodes;
    sc.width = s.tlas->width;
    sc.lds_stack = s.tlas->lds_stack;
    sc.refill = s.tlas->refill;
    sc.insts = s.tlas->d_inst;
  } else {
    sc.nodes = s.bvh ? s.bvh->d_nodes : nullptr;
    sc.bvh_tris = s.bvh ? s.bvh->d_tris : nullptr;
    sc.root = s.bvh ? s.bvh->root : 0;
    sc.n_nodes = s.bvh ? s.bvh->n_nodes : 0;
    sc.width = s.bvh ? s.bvh->width : 4;
    sc.lds_stack = s.bvh ? s.bvh->lds_stack : 8;
    sc.refill = s.bvh ? s.bvh->refill : 32;
    sc.insts = nullptr;
  }
  sc.tri_pos = s.d_pos;
  sc.tri_nrm = s.d_nrm;
  sc.tri_uv = s.d_uv;
  sc.tri_tan = s.d_tan;
  sc.tri_bit = s.d_bit;
  sc.tex = s.n_tex ? s.d_tex : nullptr;
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
