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
void scene_gpu_free(SceneGpu& s);

// Kernel view of a scene seen through a CameraUpdate camera (rotation and
// translation of the world transform, vertical fov, aspect = width/height).
SceneView scene_gpu_view(const SceneGpu& s, const apt_camera& cam, uint32_t width, uint32_t height);

}  // namespace apt
