// Two-level scene object of the C ABI (apt_tlas): BLASes relocated into one
// node arena and one triangle arena, a TLAS over instance boxes at the front
// of the node arena, and the instance records it points at.  Replaces the
// reference's per-mesh BLAS and per-frame TLAS
// (src/scene_resources/scene_model.rs:53-88, scene_resources/mod.rs:159-276).
#pragma once

#include "apt_internal.h"

struct apt_tlas {
  uint32_t width = 4;
  uint32_t lds_stack = 8;
  uint32_t refill = 32;
  // node arena: TLAS_NODE_CAP TLAS slots, then the BLASes (width-wide nodes)
  void* d_nodes = nullptr;
  size_t node_cap = 0, node_used = 0;
  // triangle arena: 3 float4 per triangle, the BLASes' leaf order
  float4* d_tris = nullptr;
  size_t tri_cap = 0, tri_used = 0;
  apt::InstRec* d_inst = nullptr;  // TLAS_MAX_INSTANCES records
  uint32_t* d_counters = nullptr;  // APT_BVH_COUNTERS ray-pool counter sets
  mutable uint32_t counter_next = 0;
  struct Blas {
    uint32_t node_base, tri_base, n_nodes, n_tris, root, max_stack;
    float lo[3], hi[3];  // object-space bounds (the root's decoded child boxes)
  };
  std::vector<Blas> blas;
  uint32_t root = apt::REF_EMPTY;  // TLAS root ref (REF_EMPTY: no instances)
  uint32_t n_tlas_nodes = 0;
  uint32_t n_inst = 0;
  uint32_t max_stack = 0;
  double ms_update = 0.0;
  std::vector<apt::InstRec> h_inst;  // the records of the last update, in leaf order
};

namespace apt {
// Ray queries over a two-level scene (trace.hip).
apt_status launch_trace_tlas(const apt_tlas* t, const apt_ray* d_rays, uint32_t n, apt_hit* d_hits,
                             uint32_t* d_occ, bool any, hipStream_t s);
// Adds a BLAS already resident in device memory (nodes with refs local to
// it, triangles in its leaf order) to the arenas; synchronous on stream s.
apt_status tlas_add_blas_device(apt_tlas* t, const void* d_nodes, uint32_t n_nodes, const float4* d_tris,
                                uint32_t n_tris, uint32_t root, uint32_t max_stack, uint32_t* index,
                                hipStream_t s);
}  // namespace apt
