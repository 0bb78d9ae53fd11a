// Device views shared by the renderer (host) and the integrator kernels.
#pragma once

#include "apt_internal.h"

namespace apt {

enum Counter : uint32_t {
  CNT_QUEUE0 = 0,
  CNT_QUEUE1 = 1,
  CNT_SHADOW = 2,
  CNT_STAT_EXT = 3,     // extension rays traced (running total)
  CNT_STAT_SHADOW = 4,  // shadow rays traced (running total)
  CNT_POOL_EXT = 16,                              // ray pool of the extend stage (POOL_WORDS,
                                                  // zeroed by the kernel before it)
  CNT_POOL_SHADOW = CNT_POOL_EXT + POOL_WORDS,    // ray pool of the shadow stage
  CNT_COUNT = CNT_POOL_SHADOW + POOL_WORDS
};

// The frame being rendered: a slab of rows [row_start, row_start + rows) of a
// width x height image.
struct FrameView {
  uint32_t width, height, row_start, rows;
  uint32_t spp;          // samples per pixel of this render call
  uint32_t frame_idx;
  uint32_t seed;
  uint32_t max_bounces;
  uint32_t rr_start;
};

// A texture of a scene: RGBA8 texels (row major, R in the low byte).
struct TexDesc {
  const uint32_t* texels;
  uint32_t width, height;
};

// Scene resident in HBM.  Triangle attributes are in input (prim) order; the
// BVH keeps its own copy of the positions in leaf order.
struct SceneView {
  const void* nodes;            // WNode<width>[] (a two-level scene: its node arena)
  const float4* bvh_tris;       // (a two-level scene: its triangle arena)
  uint32_t root;                // (a two-level scene: the TLAS root)
  uint32_t n_nodes;             // nodes the pool kernels may stage in LDS (a two-level scene: the TLAS's)
  const InstRec* insts;         // instance records of a two-level scene, or null
  uint32_t width;
  uint32_t lds_stack;           // traversal stack entries in LDS (8, 12, 16)
  uint32_t refill;              // persistent traversal refill threshold (65: one ray per thread)
  const float4* tri_pos;        // 3 per triangle
  const float4* tri_nrm;        // 3 per triangle, or null (geometric normals)
  const float2* tri_uv;         // 3 per triangle, or null (uv 0)
  const float4* tri_tan;        // 3 per triangle (tangent), or null (no normal maps)
  const float4* tri_bit;        // 3 per triangle (bitangent), or null
  const TexDesc* tex;           // the scene's textures, or null (no textured material)
  const float4* sky_tex;        // equirectangular sky panorama (RGB, w unused), or null (gradient)
  uint32_t sky_w, sky_h;
  const uint32_t* tri_mat;
  const apt_material* mats;
  const uint32_t* emissive;     // prims with emission
  uint32_t n_emissive;
  apt_sky sky;
  v3 cam_pos, cam_right, cam_up, cam_back;
  float cam_tan_half, cam_aspect;
};

// Whether a scene needs the textured kernel variants: material textures or a
// sky panorama.
__host__ __device__ inline bool scene_textured(const SceneView& s) { return s.tex || s.sky_tex; }

// A kernel's scene view with the textured-scene arrays known to be absent at
// compile time when TEX is false: every textured branch of the shading code
// (tested on these pointers) folds away, so kernels for untextured scenes
// keep the registers and occupancy they have without texture support.
template <bool TEX>
__device__ __forceinline__ SceneView scene_view(const SceneView& s) {
  SceneView v = s;
  if constexpr (!TEX) {
    v.tri_uv = nullptr;
    v.tri_tan = nullptr;
    v.tri_bit = nullptr;
    v.tex = nullptr;
    v.sky_tex = nullptr;
  }
  return v;
}

// Path state, structure of arrays over path slots.
struct PathView {
  float4* o_pdf;    // origin, pdf of the BSDF sample that made the ray (0: camera)
  float4* d_flags;  // direction, flag bits
  float4* thr;      // throughput
  float4* rad;      // radiance gathered so far
  float4* hit;      // t, u, v, prim
  float4* sh_o;     // shadow ray origin, tmax
  float4* sh_d;     // shadow ray direction
  float4* sh_c;     // unoccluded contribution
  uint32_t* rng;
  uint32_t* queue[2];
  uint32_t* shadow_queue;
  uint32_t* counters;  // Counter
};

// stage_ev: NULL, or 6 * fv.max_bounces events recorded around each bounce's
// extend, shade and shadow launches (start, end per stage).
apt_status integrate_pass(const FrameView& fv, const SceneView& sc, const PathView& ps,
                          uint32_t n_paths, uint32_t sample_base, uint32_t s_count,
                          float4* frame_sum, hipStream_t s, hipEvent_t* stage_ev = nullptr);
apt_status resolve_frame(const FrameView& fv, const float4* frame_sum, float4* accum_tm,
                         float4* accum_lin, uint32_t accum_frames, uint32_t* out_rgba,
                         hipStream_t s);

}  // namespace apt
