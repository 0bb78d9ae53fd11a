// A scene resident in HBM: BVH, triangle attributes, materials, emitter list
// and sky.  Used by the renderer (one scene) and the server (many).
#pragma once

#include "integrator.h"

namespace apt {

struct SceneGpu {
  apt_bvh* bvh = nullptr;
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

// Writes a model's n triangles in world space (positions, unit normals if
// out_nrm, material ids + mat_base) on the device; synchronous.
apt_status scene_flatten_model(const float* d_xyz, const float* d_nrm, const uint32_t* d_mat, uint32_t n,
                               const ModelXform& x, float* out_xyz, float* out_nrm, uint32_t* out_mat);

// Kernel view of a scene seen through a CameraUpdate camera (rotation and
// translation of the world transform, vertical fov, aspect = width/height).
SceneView scene_gpu_view(const SceneGpu& s, const apt_camera& cam, uint32_t width, uint32_t height);

}  // namespace apt
