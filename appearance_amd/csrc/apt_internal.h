// Internal declarations of libapt shared by the host and kernel translation
// units: error reporting, the wide-BVH device layout and the BVH object.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/apt.h"
#include "apt_math.h"

namespace apt {

// Thread-local last-error message behind apt_last_error().
void set_error(const char* fmt, ...);
void clear_error();

#define APT_HIP(call)                                                                 \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      ::apt::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                       __LINE__);                                                     \
      return e_ == hipErrorOutOfMemory ? APT_ERR_OOM : APT_ERR_DEVICE;                \
    }                                                                                 \
  } while (0)

#define APT_REQUIRE(cond, status, ...) \
  do {                                 \
    if (!(cond)) {                     \
      ::apt::set_error(__VA_ARGS__);   \
      return status;                   \
    }                                  \
  } while (0)

// The library's own stream (default priority) used when callers pass NULL.
hipStream_t lib_stream();

// ---------------------------------------------------------------------------
// Wide BVH layout in HBM.
//
// A node of width W quantizes its W child boxes to an 8-bit grid over the
// node's own box: grid origin (ox, oy, oz) = the box's lower corner, grid step
// 2^(e - 127) per axis, and for each child six bytes (lo and hi per axis, in
// steps) rounded outwards, so every decoded box contains the child's padded
// box.  Then W child refs:
//   ref <  0x80000000          internal node index
//   ref == REF_EMPTY           unused slot (never entered)
//   ref &  0x80000000          leaf: first triangle in bits 0..27, count-1 in
//                              bits 28..30 (1..8 triangles)
// 48 / 64 / 96 B for W = 2 / 4 / 8, against 64 / 128 / 256 B for float boxes.
// Triangles are stored in leaf order as three float4 (v0 | prim bits, v1, v2),
// 48 B each, so a leaf is one contiguous run.
// ---------------------------------------------------------------------------
constexpr uint32_t REF_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t REF_LEAF = 0x80000000u;

APT_HD bool ref_is_leaf(uint32_t r) { return (r & REF_LEAF) != 0u; }
APT_HD uint32_t ref_leaf_first(uint32_t r) { return r & 0x0FFFFFFFu; }
APT_HD uint32_t ref_leaf_count(uint32_t r) { return ((r >> 28) & 7u) + 1u; }
APT_HD uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
  return REF_LEAF | ((count - 1u) << 28) | first;
}

enum { QLO_X = 0, QLO_Y = 1, QLO_Z = 2, QHI_X = 3, QHI_Y = 4, QHI_Z = 5 };

template <int W>
struct alignas(16) WNode {
  float ox, oy, oz;       // grid origin
  uint8_t ex, ey, ez;     // grid step 2^(e - 127) per axis, e in 1..254
  uint8_t pad0;
  uint32_t child[W];
  uint8_t q[6][W];        // QLO_X .. QHI_Z rows, one byte per child
};
static_assert(sizeof(WNode<2>) == 48, "WNode<2> size");
static_assert(sizeof(WNode<4>) == 64, "WNode<4> size");
static_assert(sizeof(WNode<8>) == 96, "WNode<8> size");

inline size_t wide_node_bytes(uint32_t width) {
  return width == 2 ? sizeof(WNode<2>) : width == 4 ? sizeof(WNode<4>) : sizeof(WNode<8>);
}

// 2^(e - 127) for a biased exponent e in 1..254 (a normal float, exact).
APT_HD float qstep(uint32_t e) {
  union {
    uint32_t u;
    float f;
  } c;
  c.u = e << 23;
  return c.f;
}

// ---------------------------------------------------------------------------
// Two-level scenes.
//
// Node arena: [TLAS, TLAS_NODE_CAP slots][BLAS 0][BLAS 1]...  Every BLAS is
// relocated into it (internal refs + its node base, leaf first + its
// triangle base), so all BLASes share one node array and one triangle array
// and the TLAS's top levels sit at the front, where the pool kernels stage
// them in LDS.  The TLAS is a wide BVH over instance boxes whose leaves hold
// one instance each: leaf first = the index of the instance record.
//
// Instance record, 64 B, in TLAS leaf order: rows of the world-to-object
// 3x4 matrix, then the BLAS root ref (relocated), the prim base added to
// every BLAS-local prim id, and flags.  A ray enters the instance as
// o' = ((m0 x + m1 y) + m2 z) + m3 per row (d' without m3), no fused
// multiply-adds, or unchanged with INST_IDENTITY.
// ---------------------------------------------------------------------------
struct alignas(16) InstRec {
  float w2o[12];
  uint32_t blas_root;
  uint32_t prim_base;
  uint32_t flags;
  uint32_t pad;
};
static_assert(sizeof(InstRec) == 64, "InstRec size");
constexpr uint32_t INST_IDENTITY = 1u;
constexpr uint32_t TLAS_MAX_INSTANCES = APT_TLAS_MAX_INSTANCES;  // MAX_TLAS_INSTANCES, scene_resources/mod.rs:20
// A wide tree over n one-instance leaves has at most n - 1 internal nodes.
constexpr uint32_t TLAS_NODE_CAP = TLAS_MAX_INSTANCES;

// A point and a direction mapped by a 3x4 matrix held as three float4 rows
// (the evaluation order of an instance record).
APT_HD v3 xform_point(const float4& r0, const float4& r1, const float4& r2, v3 p) {
  return mk(((r0.x * p.x + r0.y * p.y) + r0.z * p.z) + r0.w, ((r1.x * p.x + r1.y * p.y) + r1.z * p.z) + r1.w,
            ((r2.x * p.x + r2.y * p.y) + r2.z * p.z) + r2.w);
}
APT_HD v3 xform_dir(const float4& r0, const float4& r1, const float4& r2, v3 d) {
  return mk((r0.x * d.x + r0.y * d.y) + r0.z * d.z, (r1.x * d.x + r1.y * d.y) + r1.z * d.z,
            (r2.x * d.x + r2.y * d.y) + r2.z * d.z);
}

// Inverse of an object-to-world affine 3x4 (row major) in double, rounded to
// float: the world-to-object rows of an instance record.  False when the 3x3
// part is singular.  *cond receives ||A||_F ||A^-1||_F (>= the 2-norm
// condition number), which sizes the instance box padding.
bool affine_inverse(const float* o2w, float* w2o, double* cond);

// Host-side binary BVH produced by the builders before collapse.
struct Node2 {
  float lo[3], hi[3];
  uint32_t left, right;   // children (internal only)
  uint32_t first, count;  // triangle range in the reordered index list (leaf only)
};

struct BvhHostResult {
  std::vector<Node2> nodes;     // node 0 is the root
  std::vector<uint32_t> order;  // reordered triangle indices
  double sah_cost = 0.0;
};

// Binned-SAH build on the host (bvh_host.cpp).
apt_status build_bvh2_host(const float* tri_xyz, uint32_t n, uint32_t max_leaf, uint32_t bins,
                           BvhHostResult& out);

double bvh2_sah_cost(const std::vector<Node2>& nodes);

// Collapse of a binary tree into the wide layout on the host only: the wide
// nodes (breadth first, wide_node_bytes(width) bytes each), the root ref and
// the statistics (n_nodes, n_leaves, depth, max_stack).  Child boxes are
// padded by `pad`.
apt_status collapse_wide_host(const BvhHostResult& b2, uint32_t width, float pad, std::vector<uint8_t>& nodes,
                              uint32_t& root, apt_bvh_stats& stats);
// Scene-scale padding used for every node box (see box_pad in bvh_host.cpp).
float scene_pad(const float* tri_xyz, uint32_t n);

struct apt_bvh_impl;

}  // namespace apt

// A triangle soup with per-triangle materials (procedural scenes, glTF files).
struct apt_scene_texture {
  uint32_t width = 0, height = 0;
  std::vector<uint32_t> rgba;  // RGBA8, row major
};

struct apt_scene_data {
  std::vector<float> positions;   // n_tris * 9
  std::vector<float> normals;     // n_tris * 9 or empty
  std::vector<uint32_t> mat_ids;  // n_tris
  std::vector<apt_material> materials;
  std::vector<float> uvs;         // n_tris * 6 or empty
  std::vector<float> tangents;    // n_tris * 12 (xyz, handedness) or empty
  std::vector<float> bitangents;  // n_tris * 9 or empty (then cross(n, t) * w)
  std::vector<apt_scene_texture> textures;
  mutable std::vector<apt_texture> texture_views;  // views of `textures` for apt_scene_data_attribs
  apt_camera camera{};
  apt_sky sky{};
};

namespace apt {
// glTF ingestion (gltf.cpp): a .glb / .gltf file into a scene.
apt_status gltf_load(const char* path, apt_scene_data** out);
// PNG image -> RGBA8 (image.cpp); false with a reason in err.
bool png_decode(const uint8_t* data, size_t n, uint32_t& w, uint32_t& h, std::vector<uint32_t>& rgba,
                std::string& err);
// Baseline JPEG image -> RGBA8 (jpeg_decode.cpp); false with a reason in err.
bool jpeg_decode(const uint8_t* data, size_t n, uint32_t& w, uint32_t& h, std::vector<uint32_t>& rgba,
                 std::string& err);
// Per-vertex tangents of an indexed mesh (Lengyel's method, the reference's
// generate_tangents, crates/appearance-model/src/mesh.rs:82-160): 4 floats per
// vertex (unit tangent orthogonalised against the normal, handedness +-1).
void generate_tangents(const std::vector<float>& pos, const std::vector<float>& nrm, const std::vector<float>& uv,
                       const std::vector<uint32_t>& idx, std::vector<float>& out);
}  // namespace apt

// The opaque handle of the C ABI.
struct apt_bvh {
  uint32_t width = 4;
  uint32_t n_tris = 0;
  uint32_t n_nodes = 0;
  uint32_t root = 0;        // root ref (an internal node or a single leaf)
  uint32_t lds_stack = 16;  // traversal stack entries kept in LDS (8, 12, 16)
  uint32_t refill = 32;     // persistent traversal refill threshold (65: one ray per thread)
  void* d_nodes = nullptr;  // WNode<width>[n_nodes]
  float4* d_tris = nullptr; // 3 * n_tris float4 in leaf order
  uint32_t* d_counters = nullptr;     // APT_BVH_COUNTERS ray-pool counters
  mutable uint32_t counter_next = 0;  // next counter slot (atomic)
  apt_bvh_stats stats{};
};
constexpr uint32_t APT_BVH_COUNTERS = 64;  // ray pools in a BVH's ring

// A persistent launch's ray pool: POOL_PARTS counters POOL_STRIDE words apart
// (two ranges of rays per XCD, see pool_take in trace_device.h).
constexpr uint32_t POOL_PARTS = 16;
constexpr uint32_t POOL_STRIDE = 16;
constexpr uint32_t POOL_WORDS = POOL_PARTS * POOL_STRIDE;

namespace apt {
// Collapse a binary BVH into a wide one and upload it with the leaf-order
// triangle records (bvh_host.cpp).
// With upload = false only the wide tree's statistics are filled in (no
// device memory is touched).
apt_status upload_wide_bvh(const float* tri_xyz, uint32_t n, const BvhHostResult& b2, uint32_t width, float pad,
                           apt_bvh* out, bool upload = true);
// The same binary tree built on the device from device-resident triangles
// and collapsed there into `out` (bvh_gpu.hip): byte-identical to
// build_bvh2_host + upload_wide_bvh (the SAH cost statistic is summed in
// another order).  Synchronous on stream s.
apt_status build_wide_device(const float* d_tri_xyz, uint32_t n, uint32_t max_leaf, uint32_t bins,
                             uint32_t width, hipStream_t s, apt_bvh* out);
// Stack entries every traversal kernel holds whatever its LDS share (the
// smallest LDS stack, 8, plus the private overflow array), and the check
// every path that traces a BVH makes first: a deeper tree would lose subtrees
// when the stack is full, so it is refused with APT_ERR_UNSUPPORTED (trace.hip).
uint32_t trace_stack_capacity();
apt_status check_trace_stack(const apt_bvh_stats& st, const char* who);
// Launch the traversal kernels (trace.hip).
apt_status launch_trace_closest(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n,
                                apt_hit* d_hits, hipStream_t s);
apt_status launch_trace_any(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n,
                            uint32_t* d_occ, hipStream_t s);
}  // namespace apt
