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

// A model's placement: the column-major world transform, the normal matrix
// (rows of the inverse transpose of its 3x3 part; nm_ok = 0 keeps normals as
// they are) and the offset of its materials in the flattened scene.
struct ModelXform {
  float m[16];
  float nm[9];
  uint32_t nm_ok;
  uint32_t mat_base;
};

// Places a model's n resident triangles in world space in a two-level
// scene's attribute arrays from prim `base` on: float4 positions, unit
// normals (when the scene keeps normals), material ids + mat_base.
// Asynchronous on stream st.
apt_status scene_place_model(const float* d_xyz, const float* d_nrm, const uint32_t* d_mat, uint32_t n,
                             const ModelXform& x, SceneGpu& s, uint32_t base, hipStream_t st);
// Sizes a two-level scene's attribute arrays for n triangles (normals when
// with_normals) and n_mats materials, keeping them when they are big enough,
// and uploads the materials.
apt_status scene_reserve(SceneGpu& s, uint32_t n, bool with_normals, const apt_material* mats, uint32_t n_mats,
                         hipStream_t st);
// Rebuilds the emitter list of a two-level scene from its world positions
// and materials; synchronous on st.
apt_status scene_emitters(SceneGpu& s, hipStream_t st);

// Kernel view of a scene seen through a CameraUpdate camera (rotation and
// translation of the world transform, vertical fov, aspect = width/height).
SceneView scene_gpu_view(const SceneGpu& s, const apt_camera& cam, uint32_t width, uint32_t height);

}  // namespace apt
