// A scene resident in HBM: BVH, triangle attributes, materials, emitter list
// and sky.  Used by the renderer (one scene) and the server (many).
#pragma once

#include "integrator.h"

namespace apt {

struct SceneGpu {
  apt_bvh* bvh = nullptr;
  // A two-level scene (spawned models) traces this instead of bvh; owned by
  // the caller, and the attribute arrays below keep their capacity across
  // world actions (cap_tris, cap_mats, cap_emissive).
  apt_tlas* tlas = nullptr;
  size_t cap_tris = 0, cap_mats = 0, cap_emissive = 0;
  float4* d_pos = nullptr;      // 3 per triangle
  float4* d_nrm = nullptr;      // 3 per triangle or null
  float2* d_uv = nullptr;       // 3 per triangle or null
  float4* d_tan = nullptr;      // 3 per triangle (world tangent) or null
  float4* d_bit = nullptr;      // 3 per triangle (world bitangent) or null
  TexDesc* d_tex = nullptr;     // texture table or null
  uint32_t* d_texels = nullptr; // texels of an explicit scene's textures (spawned models keep their own)
  uint32_t n_tex = 0;
  size_t cap_tex = 0;
  uint32_t* d_mat = nullptr;
  apt_material* d_mats = nullptr;
  uint32_t* d_emissive = nullptr;
  uint32_t n_tris = 0, n_mats = 0, n_emissive = 0;
  apt_sky sky{};
  size_t bytes = 0;
};

// Builds the BVH and uploads the attributes of a triangle soup.
apt_status scene_gpu_upload(const float* pos, const float* nrm, const uint32_t* mat_ids, uint32_t n,
                            const apt_material* mats, uint32_t n_mats, const apt_sky* sky,
                            uint32_t bvh_width, uint32_t lds_stack, SceneGpu& out);
// Same, from triangles already in device memory: xyz triples and optional
// normals (9 floats per triangle) and material ids (checked by the caller);
// the arrays stay the caller's.
apt_status scene_gpu_upload_device(const float* d_xyz, const float* d_nrm, const uint32_t* d_mat_ids,
                                   uint32_t n, const apt_material* mats, uint32_t n_mats,
                                   const apt_sky* sky, uint32_t bvh_width, uint32_t lds_stack,
                                   SceneGpu& out);
void scene_gpu_free(SceneGpu& s);
// Uvs, tangent frames and textures of an explicit scene (apt_scene_set_attribs):
// per vertex the unit tangent and the bitangent (given, or cross(n, t) * w with
// the vertex normal, or the face normal without normals), in world space.
apt_status scene_gpu_attribs(SceneGpu& s, const float* pos, const float* nrm, uint32_t n,
                             const apt_mesh_attribs& a);

// A model's placement: the column-major world transform, the normal matrix
// (rows of the inverse transpose of its 3x3 part; nm_ok = 0 keeps normals as
// they are) and the offset of its materials in the flattened scene.
struct ModelXform {
  float m[16];
  float nm[9];
  uint32_t nm_ok;
  uint32_t mat_base;
};

// A model's optional per-vertex attributes in object space: uvs (float2),
// tangents (xyz, handedness) and bitangents (xyz; null: cross(n, t) * w).
struct ModelAttribs {
  const float2* uv = nullptr;
  const float4* tan = nullptr;
  const float4* bit = nullptr;
};

// Places a model's n resident triangles in world space in a two-level
// scene's attribute arrays from prim `base` on: float4 positions, unit
// normals (when the scene keeps normals), material ids + mat_base.
// Asynchronous on stream st.
apt_status scene_place_model(const float* d_xyz, const float* d_nrm, const uint32_t* d_mat, uint32_t n,
                             const ModelXform& x, const ModelAttribs& at, SceneGpu& s, uint32_t base,
                             hipStream_t st);
// Sizes a two-level scene's attribute arrays for n triangles (normals, uvs,
// tangent frames as asked), n_mats materials and n_tex texture descriptors,
// keeping them when they are big enough, and uploads materials and
// descriptors.
apt_status scene_reserve(SceneGpu& s, uint32_t n, bool with_normals, bool with_uv, bool with_tangents,
                         const apt_material* mats, uint32_t n_mats, const TexDesc* tex, uint32_t n_tex,
                         hipStream_t st);
// Rebuilds the emitter list of a two-level scene from its world positions
// and materials; synchronous on st.
apt_status scene_emitters(SceneGpu& s, hipStream_t st);

// Kernel view of a scene seen through a CameraUpdate camera (rotation and
// translation of the world transform, vertical fov, aspect = width/height).
SceneView scene_gpu_view(const SceneGpu& s, const apt_camera& cam, uint32_t width, uint32_t height);

}  // namespace apt
