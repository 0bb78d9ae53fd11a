/*
 * apt.h - C ABI of libapt, the MI355X (gfx950) path-tracing hot path.
 *
 * Drop-in for the render-node job interface of Appearance:
 *   NodeRenderer              crates/appearance-render-loop/src/node.rs:14-25
 *   DistributedRenderer       crates/appearance-distributed-renderer/src/lib.rs:60-82
 *   PathTracerGpu::{new, handle_visible_world_action, render}
 *                             crates/appearance-path-tracer-gpu/src/lib.rs:260-615
 * and for the hardware ray query those kernels call
 *   rayQueryInitialize/Proceed/GetCommittedIntersection
 *                             crates/appearance-path-tracer-gpu/assets/shaders/trace.wgsl:159-162,
 *                             helpers/trace.wgsl:20-26 (terminate-on-first-hit shadow query)
 *   BLAS/TLAS build           src/scene_resources/scene_model.rs:53-88, scene_resources/mod.rs:218-276
 *
 * Conventions
 *   - Every entry point returns apt_status and never aborts; on failure the
 *     reason is readable through apt_last_error() on the calling thread.
 *   - Structs are plain C PODs; sizes are fixed and checked by static asserts
 *     in the implementation.  A Rust -sys crate mirrors them 1:1.
 *   - Pointers named d_* are device (HBM) pointers; every other pointer is host
 *     memory owned by the caller.
 *   - One apt_renderer is not re-entrant (the reference's &mut self); callers
 *     serialise calls on it.  Distinct renderers may be used from distinct
 *     threads.
 */
#ifndef APT_H
#define APT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APT_ABI_VERSION 1u

typedef enum apt_status {
  APT_OK = 0,
  APT_ERR_INVALID_ARG = 1,
  APT_ERR_OOM = 2,
  APT_ERR_DEVICE = 3,
  APT_ERR_UNSUPPORTED = 4,
  APT_ERR_BUSY = 5,
  APT_ERR_NOT_FOUND = 6
} apt_status;

/* Message of the last failed call on this thread ("" if none). */
const char* apt_last_error(void);
uint32_t apt_abi_version(void);

typedef struct apt_device_info {
  char name[64];
  char arch[32];         /* e.g. "gfx950" */
  uint32_t compute_units;
  uint32_t wavefront_size;
  uint64_t hbm_bytes;
  uint64_t lds_bytes_per_block;
} apt_device_info;
apt_status apt_device_query(apt_device_info* out);

/* ------------------------------------------------------------------------
 * Ray-query level.  Replaces the driver BVH build and rayQuery* traversal.
 * ------------------------------------------------------------------------ */

/* 32 B ray: origin, tmin, direction (need not be normalised), tmax. */
typedef struct apt_ray {
  float org[3];
  float tmin;
  float dir[3];
  float tmax;
} apt_ray;

/* 16 B closest hit.  prim = APT_MISS when nothing is hit; u, v are the weights
 * of the second and third vertex (the reference's intersection.barycentrics,
 * trace.wgsl:168).  The hit is the lexicographic minimum of (t, prim) over all
 * triangles with tmin < t < tmax under the watertight test, so it does not
 * depend on the BVH width or on traversal order. */
typedef struct apt_hit {
  float t;
  float u;
  float v;
  uint32_t prim;
} apt_hit;

#define APT_MISS 0xFFFFFFFFu

typedef enum apt_builder {
  APT_BUILDER_HOST = 0,  /* binned SAH on the host threads of libapt */
  APT_BUILDER_GPU = 1    /* binned SAH, collapse and quantization on the device; the same bytes */
} apt_builder;

typedef struct apt_bvh_opts {
  uint32_t width;     /* 2, 4 or 8 children per node (0 = 4) */
  uint32_t max_leaf;  /* triangles per leaf, 1..8 (0 = 4) */
  uint32_t bins;      /* SAH bins per axis, 4..64 (0 = 32) */
  uint32_t builder;   /* apt_builder */
  uint32_t lds_stack; /* traversal stack entries kept in LDS: 8, 12 or 16 (0 = 8) */
  uint32_t refill;    /* ray queries: persistent waves refetch rays once this many of their
                         64 lanes are idle, 1..64; 65 = one ray per thread (0 = 32) */
} apt_bvh_opts;

typedef struct apt_bvh_stats {
  uint32_t n_tris;
  uint32_t n_nodes;
  uint32_t n_leaves;
  uint32_t width;
  uint32_t depth;       /* levels of wide nodes */
  uint32_t max_stack;   /* worst-case traversal stack entries */
  uint64_t device_bytes;
  double build_ms;
  double sah_cost;      /* SAH cost of the binary tree before collapse (the device builder sums
                           it per 256 nodes, so it may differ from the host's in the last bits) */
} apt_bvh_stats;

typedef struct apt_bvh apt_bvh;

/* Build from n_tris triangles, tri_xyz = n_tris * 9 floats (v0 v1 v2) in host
 * memory.  The triangles are copied to the device; prim ids are their input
 * indices. */
apt_status apt_bvh_build(const float* tri_xyz, uint32_t n_tris, const apt_bvh_opts* opts,
                         apt_bvh** out);
/* Same, with tri_xyz already resident in device memory (with APT_BUILDER_GPU
 * nothing but level sizes and statistics crosses to the host). */
apt_status apt_bvh_build_device(const float* d_tri_xyz, uint32_t n_tris,
                                const apt_bvh_opts* opts, apt_bvh** out);
/* Builds on the host threads and reports the tree's statistics without
 * touching the device, for checking a scene before it is uploaded.  Returns
 * APT_ERR_UNSUPPORTED, with *out filled in, when the tree needs a deeper
 * traversal stack (out->max_stack) than the kernels hold
 * (apt_trace_stack_capacity()); every entry point that traces a BVH (the ray
 * queries, apt_scene_set and world actions, apt_server_add_scene,
 * apt_autotune) refuses such a tree the same way instead of losing subtrees.
 * No counterpart in the reference (the driver builds and traverses there). */
apt_status apt_bvh_plan(const float* tri_xyz, uint32_t n_tris, const apt_bvh_opts* opts,
                        apt_bvh_stats* out);
uint32_t apt_trace_stack_capacity(void);
void apt_bvh_destroy(apt_bvh* bvh);
apt_status apt_bvh_get_stats(const apt_bvh* bvh, apt_bvh_stats* out);
/* Copies the device layout out: n_nodes wide nodes of 48 / 64 / 96 bytes for
 * width 2 / 4 / 8 (float grid origin[3]; uint8 biased grid exponent[3] and one
 * pad byte; uint32 child ref[width]; uint8 quantized box rows lo_x, lo_y,
 * lo_z, hi_x, hi_y, hi_z of [width] bytes each; zero padding to 16 bytes) and
 * 3 * n_tris float4 triangle records in leaf order (prim id in the first
 * record's w).
 * No counterpart in the reference (its BVH is opaque inside the driver); for
 * callers that cache builds, and for comparing builders. */
apt_status apt_bvh_export(const apt_bvh* bvh, void* nodes, size_t nodes_len, void* tris,
                          size_t tris_len);

/* Closest-hit query (ray flags 0 in the reference): d_rays/d_hits are device
 * arrays of n entries; stream is a hipStream_t (NULL = the library stream).
 * The call is asynchronous on that stream. */
apt_status apt_trace_closest(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n,
                             apt_hit* d_hits, void* stream);
/* Any-hit query (RAY_FLAG_TERMINATE_ON_FIRST_HIT, helpers/trace.wgsl:22):
 * d_occluded[i] = 1 when some triangle has tmin < t < tmax, else 0. */
apt_status apt_trace_any(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n,
                         uint32_t* d_occluded, void* stream);
/* Synchronous variants on host arrays (copies in and out). */
apt_status apt_trace_closest_host(const apt_bvh* bvh, const apt_ray* rays, uint32_t n,
                                  apt_hit* hits);
apt_status apt_trace_any_host(const apt_bvh* bvh, const apt_ray* rays, uint32_t n,
                              uint32_t* occluded);

/* apt_trace_closest_host that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * rays, hits and ray-pool counters are staged in a fixed device array of the
 * library, 3 MiB + 1 KiB per GPU, held until the process ends; n above
 * APT_TRACE_NOALLOC_MAX_RAYS returns APT_ERR_INVALID_ARG.  Every call on a
 * GPU shares that GPU's array (it keeps the last call's rays and hits), so
 * calls that overlap on one GPU race: serialise them.  The BVH must already
 * exist (apt_bvh_build allocates). */
#define APT_TRACE_NOALLOC_MAX_RAYS 65536u
apt_status apt_trace_closest_host_noalloc(const apt_bvh* bvh, const apt_ray* rays, uint32_t n,
                                          apt_hit* hits);

/* ------------------------------------------------------------------------
 * Two-level scenes: a top-level BVH (TLAS) over instances of bottom-level
 * BVHs (BLAS), the reference's acceleration structures
 * (src/scene_resources/scene_model.rs:53-88: one BLAS per mesh, built once;
 * scene_resources/mod.rs:159-276: a TLAS of at most MAX_TLAS_INSTANCES = 8192
 * instances rebuilt from the instance transforms every frame).
 *
 * A hit is traced in the instance's object space: the ray is mapped by the
 * inverse of the instance transform (computed in double, rounded to float)
 * as o' = ((m0 x + m1 y) + m2 z) + m3 per row, without fused multiply-adds,
 * and tested against the BLAS triangles with the watertight test; t is the
 * same ray parameter in both spaces (affine map) and u, v the same
 * barycentrics.  An instance whose transform is exactly the identity uses
 * the ray as it is.  The closest hit is the lexicographic minimum of
 * (t, custom_index + BLAS primitive) over all instances.
 * ------------------------------------------------------------------------ */

#define APT_TLAS_MAX_INSTANCES 8192u

/* 64 B.  transform: object-to-world, row-major 3x4 (the reference's
 * TlasInstance transform, scene_resources/mod.rs:177-191).  blas: index
 * returned by apt_tlas_add_blas.  custom_index: added to the BLAS primitive
 * index in every hit.  The reference's instance custom data selects a
 * vertex-pool slice whose first triangle the shader adds to primitive_index
 * (trace.wgsl:106-120); custom_index is that first triangle. */
typedef struct apt_instance {
  float transform[12];
  uint32_t blas;
  uint32_t custom_index;
  uint32_t _pad[2];
} apt_instance;

typedef struct apt_tlas apt_tlas;

typedef struct apt_tlas_stats {
  uint32_t n_blas;
  uint32_t n_instances;
  uint32_t n_tlas_nodes;
  uint32_t width;
  uint32_t max_stack;       /* worst-case traversal stack: TLAS path + deepest BLAS */
  uint32_t _pad;
  uint64_t blas_tris;       /* triangles of all BLASes */
  uint64_t device_bytes;
  double ms_last_update;    /* wall time of the last apt_tlas_set_instances */
} apt_tlas_stats;

/* An empty two-level scene; opts->width (0 = 4) is the width of every BLAS
 * it takes and of its TLAS; lds_stack and refill as for apt_bvh_opts. */
apt_status apt_tlas_create(const apt_bvh_opts* opts, apt_tlas** out);
/* Copies a BLAS (built with the same width) into the scene's device arena
 * and returns its index; the apt_bvh may be destroyed afterwards. */
apt_status apt_tlas_add_blas(apt_tlas* t, const apt_bvh* blas, uint32_t* blas_index);
/* Rebuilds the top level over n instances (n <= APT_TLAS_MAX_INSTANCES;
 * n = 0 makes every ray miss).  Synchronous: on return the device holds the
 * new TLAS.  APT_ERR_UNSUPPORTED when a transform is singular or the two
 * levels need a deeper traversal stack than apt_trace_stack_capacity(). */
apt_status apt_tlas_set_instances(apt_tlas* t, const apt_instance* instances, uint32_t n);
void apt_tlas_destroy(apt_tlas* t);
apt_status apt_tlas_get_stats(const apt_tlas* t, apt_tlas_stats* out);
/* The ray queries of apt_trace_closest / apt_trace_any over a two-level
 * scene (device arrays, asynchronous on stream), and their host-array
 * variants. */
apt_status apt_trace_closest_tlas(const apt_tlas* t, const apt_ray* d_rays, uint32_t n, apt_hit* d_hits,
                                  void* stream);
apt_status apt_trace_any_tlas(const apt_tlas* t, const apt_ray* d_rays, uint32_t n, uint32_t* d_occluded,
                              void* stream);
apt_status apt_trace_closest_tlas_host(const apt_tlas* t, const apt_ray* rays, uint32_t n, apt_hit* hits);
apt_status apt_trace_any_tlas_host(const apt_tlas* t, const apt_ray* rays, uint32_t n, uint32_t* occluded);

/* ------------------------------------------------------------------------
 * Scene data.  Layouts follow the reference's GPU pools.
 * ------------------------------------------------------------------------ */

#define APT_INVALID_TEXTURE 0xFFFFFFFFu

/* 160 B; field for field the reference's MaterialDescriptor
 * (src/scene_resources/material_pool.rs:13-46, shared/material/material_pool.wgsl:6-41).
 * bsdf_model occupies the reference's _padding0: 0 = Disney BSDF (the
 * reference's only model), 1 = Lambertian (the procedural Cornell box). */
typedef struct apt_material {
  float color[3];
  uint32_t color_texture;
  float metallic;
  float roughness;
  uint32_t metallic_roughness_texture;
  float normal_scale;
  float emission[3];
  uint32_t normal_texture;
  uint32_t emission_texture;
  float transmission;
  float eta;
  float subsurface;
  float absorption[3];
  float specular;
  float specular_tint[3];
  float anisotropic;
  float sheen;
  uint32_t sheen_texture;
  float clearcoat;
  uint32_t clearcoat_texture;
  float clearcoat_roughness;
  uint32_t clearcoat_roughness_texture;
  /* Carried for layout only: every hit is opaque, as in the reference, whose
   * MAX_NON_OPAQUE_DEPTH and MAX_NON_OPAQUE_SHADOW_DEPTH are 1
   * (helpers/trace.wgsl:7-8), so its alpha test (trace.wgsl:124-136) always
   * takes the opaque branch and shadow rays use the opaque query. */
  float alpha_cutoff;
  uint32_t sheen_tint_texture;
  uint32_t clearcoat_normal_texture;
  uint32_t bsdf_model;
  uint32_t _padding1;
  uint32_t _padding2;
  float sheen_tint[3];
  uint32_t transmission_texture;
} apt_material;

#define APT_BSDF_DISNEY 0u
#define APT_BSDF_LAMBERT 1u

/* Sets every field to the reference's glTF defaults: white, metallic 0,
 * roughness 1, eta 1/1.5 (crates/appearance-model/src/asset.rs:229), no
 * textures, Disney model. */
void apt_material_default(apt_material* m);

/* Sun (the reference's SunInfo, src/scene_resources/sky.rs:11-22, 32 B)
 * plus an analytic gradient sky that stands in for the HDR panorama. */
typedef struct apt_sky {
  float sun_direction[3]; /* direction the sunlight travels (reference convention) */
  float sun_size;
  float sun_color[3];
  float sun_intensity;
  float zenith[3];        /* sky radiance straight up */
  float sun_enabled;      /* 0 or 1 */
  float horizon[3];       /* sky radiance at the horizon */
  float _pad0;
  float ground[3];        /* radiance below the horizon */
  float _pad1;
} apt_sky;

/* A flat triangle soup with per-triangle material ids. */
typedef struct apt_mesh_desc {
  const float* positions;        /* n_tris * 9 */
  const float* normals;          /* n_tris * 9 vertex normals, or NULL = geometric */
  const uint32_t* material_ids;  /* n_tris, or NULL = all 0 */
  uint32_t n_tris;
  uint32_t n_materials;
  const apt_material* materials; /* n_materials */
} apt_mesh_desc;

/* Texture sampling and vertex attributes of a textured scene.
 *
 * A texture is width x height RGBA8 texels (R in the low byte), row 0 at
 * v = 0 (glTF's image origin).  Materials index the scene's texture list
 * (APT_INVALID_TEXTURE = none) and the texture scales the material's factor as
 * the reference's MaterialDescriptor accessors do
 * (shared/material/material_pool_bindings.wgsl:19-107): colour x sRGB-to-
 * linear(rgba) (alpha -> the Disney luminance), emission x rgb, metallic x b
 * and roughness x g, clearcoat x r, clearcoat roughness x g, transmission x r,
 * sheen x r, sheen tint x sRGB-to-linear(rgb), each only when the factor is
 * nonzero; the normal texture (rgb * 2 - 1, in the interpolated tangent frame,
 * blended by normal_scale below 1, :89-123) and the clearcoat normal texture
 * bend the shading normals.  Sampling is bilinear with repeat addressing at
 * level 0 (the reference's sampler, src/scene_resources/material_pool.rs:68-74),
 * in FP32 on the decoded texels (the reference samples BC-compressed copies).
 * Emission textures also weight light samples on emissive triangles
 * (helpers/nee.wgsl:16-57). */
typedef struct apt_texture {
  uint32_t width;
  uint32_t height;
  const uint32_t* rgba8;  /* width * height texels, row major */
} apt_texture;

/* Per-vertex attributes beyond positions and normals (every pointer may be
 * NULL): uvs = n_tris * 6 floats (TEXCOORD_0), tangents = n_tris * 12 floats
 * (xyz and handedness w per vertex, glTF TANGENT), bitangents = n_tris * 9
 * floats (NULL: cross(normal, tangent) * w per vertex,
 * shared/vertex_pool_bindings.wgsl:43-58), and the scene's textures. */
typedef struct apt_mesh_attribs {
  const float* uvs;
  const float* tangents;
  const float* bitangents;
  const apt_texture* textures;
  uint32_t n_textures;
  uint32_t _pad;
} apt_mesh_attribs;

/* The reference's CameraUpdateData (crates/appearance-world/src/visible_world_action.rs:9-15),
 * 80 B: column-major world transform of the camera (it looks down its -Z),
 * vertical fov in degrees, near, far. */
typedef struct apt_camera {
  float transform[16];
  float fov_deg;
  float near_plane;
  float far_plane;
  uint32_t _padding;
} apt_camera;

/* Procedural scenes generated inside the library from a seed:
 *   "cornell"   Cornell box, 36 triangles, Lambertian, quad light (17,12,4)
 *   "bvh_stress" 1,048,576 random triangles in [-10,10]^3
 *   "disney"    256 UV spheres x 2048 triangles + ground, Disney materials
 *   "restir"    "disney" + 4096 small emissive triangles
 * The same names work as asset paths "procedural://<name>?seed=<n>" in a
 * SpawnModel world action. */
typedef struct apt_scene_data apt_scene_data;
apt_status apt_scene_generate(const char* name, uint32_t seed, apt_scene_data** out);
/* Reads a glTF 2.0 model (.glb, or .gltf with external or base64 buffers)
 * into a triangle soup with materials, as the reference's asset loader
 * (crates/appearance-model/src/asset.rs:19-380) builds a Model: the default
 * scene's nodes with their transforms applied, triangle primitives with
 * POSITION, NORMAL (generated when absent, mesh.rs:58-79) and indices, and
 * each material's factors mapped as asset.rs:199-339 maps them (emissive
 * strength, volume absorption, transmission, eta = 1/ior, specular,
 * clearcoat, sheen), TEXCOORD_0, TANGENT (generated when absent with the
 * reference's per-vertex method, crates/appearance-model/src/mesh.rs:82-160)
 * and the textures the materials use (PNG images decoded to RGBA8; an image
 * in another format leaves its texture slot empty, so the material uses its
 * factor).  The same paths work as SpawnModel asset paths. */
apt_status apt_scene_load_gltf(const char* path, apt_scene_data** out);
/* Views into the library-owned arrays; valid until apt_scene_data_destroy. */
apt_status apt_scene_data_view(const apt_scene_data* s, apt_mesh_desc* mesh, apt_camera* camera,
                               apt_sky* sky);
/* The scene's uvs, tangents, bitangents and textures (NULL / 0 when it has
 * none); valid until apt_scene_data_destroy. */
apt_status apt_scene_data_attribs(const apt_scene_data* s, apt_mesh_attribs* out);
void apt_scene_data_destroy(apt_scene_data* s);

/* The bvh_stress ray sets: n_coherent primary rays of a side x side pinhole camera at
 * (0,0,-25) looking +z with a 40 degree fov (side*side == n_coherent), then
 * n_incoherent rays with origins uniform in [-10,10]^3 and uniform-sphere
 * directions, all with tmin 0 and tmax 1e30. */
apt_status apt_rays_bvh_stress(uint32_t seed, uint32_t side, uint32_t n_incoherent, apt_ray* out);

/* ------------------------------------------------------------------------
 * Renderer (PathTracerGpu / NodeRenderer).
 * ------------------------------------------------------------------------ */

typedef enum apt_integrator {
  /* Wavefront unidirectional path tracer: jittered samples, NEE + MIS,
   * Russian roulette, many samples per call; no G-buffer, so no post chain. */
  APT_INTEGRATOR_PLAIN = 0,
  /* The reference's integrator (src/lib.rs:354-576): pixel-centre samples,
   * DI by RIS over light + BSDF candidates, the continuation chosen by a GI
   * reservoir, spatiotemporal reuse at the first vertex (restir_di,
   * restir_gi), then demodulation, firefly filter, SVGF and TAA. */
  APT_INTEGRATOR_RESTIR = 1
} apt_integrator;

/* PathTracerGpuConfig (src/lib.rs:219-245), made public, plus the knobs the
 * reference keeps as shader constants and the MI355X tuning parameters. */
typedef struct apt_config {
  uint32_t max_bounces;       /* path segments traced per sample */
  uint32_t sample_count;      /* samples per pixel per render call */
  uint32_t accum_frames;      /* progressive accumulation across calls */
  uint32_t rr_start_bounce;   /* PLAIN: Russian roulette from this bounce on (RESTIR: bounce 2, trace.wgsl:220) */
  uint32_t integrator;        /* apt_integrator */
  uint32_t restir_di;         /* RESTIR: temporal + spatial reuse of DI reservoirs */
  uint32_t restir_gi;         /* RESTIR: temporal reuse of GI reservoirs */
  uint32_t svgf;              /* RESTIR: SVGF denoiser */
  uint32_t firefly_filter;    /* RESTIR: 3x3 luma clamp before SVGF */
  uint32_t taa;               /* RESTIR: temporal antialiasing */
  uint32_t di_candidates;     /* RESTIR: light candidates of the DI RIS (plus one BSDF candidate), <= 64 */
  uint32_t di_mcap;           /* RESTIR: temporal M cap, times the current M */
  uint32_t di_spatial_neighbours; /* RESTIR: spatial taps, <= 16 */
  float di_spatial_radius;    /* RESTIR: spatial radius in pixels (0 = 5% of (width + rows) / 2) */
  uint32_t gi_mcap;           /* RESTIR: GI temporal M cap */
  uint32_t svgf_iterations;   /* RESTIR: a-trous passes, <= 8 */
  uint32_t svgf_history;      /* RESTIR: max frames of SVGF moment history */
  uint32_t bvh_width;         /* 2, 4, 8 */
  uint32_t queue_chunk;       /* PLAIN: paths per wavefront pass (0 = 2^24) */
  uint32_t lds_stack_depth;   /* traversal stack entries kept in LDS: 8, 12, 16 */
  uint32_t seed;              /* base seed mixed into every path seed */
  uint32_t frame_idx;         /* starting frame index */
  uint32_t gi_spatial_passes; /* RESTIR: ReSTIR GI spatial reuse passes (restir_gi_spatial.wgsl), <= 4;
                                 0 as the reference ships it (lib.rs:441) */
} apt_config;

/* PathTracerGpuConfig::default (src/lib.rs:230-245): the RESTIR integrator,
 * 2 bounces, 1 sample, no accumulation, restir_di, restir_gi and the firefly
 * filter on, SVGF and TAA off; the shader constants of the reference
 * (4 light candidates, M caps 20 and 30, 3 spatial taps, 5 a-trous passes,
 * 24 history frames) and MI355X tuning defaults. */
void apt_config_default(apt_config* cfg);

typedef struct apt_renderer apt_renderer;

apt_status apt_create(const apt_config* cfg, apt_renderer** out);
void apt_destroy(apt_renderer* r);
apt_status apt_set_config(apt_renderer* r, const apt_config* cfg);

/* VisibleWorldActionType::from_ty_and_bytes (visible_world_action.rs:87-96):
 * ty 0 CameraUpdate (80 B), 1 SpawnModel (336 B), 2 TransformModel (80 B),
 * 3 DestroyModel (16 B), 4 Clear (4 B).  SpawnModel accepts
 * "procedural://<name>?seed=<n>" asset paths and glTF files (.glb, .gltf;
 * apt_scene_load_gltf).  Spawned models form a two-level scene as in the
 * reference (scene_resources/mod.rs:115-276): an asset path is loaded once
 * and its BLAS built once in object space (later spawns of the path share
 * them, until Clear); every world action rebuilds the TLAS over the models
 * (at most APT_TLAS_MAX_INSTANCES) and re-places their world-space
 * attributes.  A singular model transform makes the next render return
 * APT_ERR_UNSUPPORTED. */
apt_status apt_world_action(apt_renderer* r, uint32_t ty, const void* bytes, size_t len);

/* Replace the whole scene with a triangle soup (procedural or caller-made). */
apt_status apt_scene_set(apt_renderer* r, const apt_mesh_desc* mesh, const apt_sky* sky);
/* apt_scene_set with uvs, tangent frames and textures (attribs may be NULL). */
apt_status apt_scene_set_attribs(apt_renderer* r, const apt_mesh_desc* mesh, const apt_mesh_attribs* attribs,
                                 const apt_sky* sky);

/* The sky panorama (the reference's sky texture: an equirectangular HDR image,
 * ::evening_road_01_puresky_4k.hdr, set in src/scene_resources/mod.rs:70-75 and
 * sky.rs:93-105; its blob is missing from the reference, so by default the
 * analytic gradient of apt_sky stands in).  width x height linear RGB texels,
 * 3 floats each, row 0 at the zenith.  Rays that leave the scene read it
 * (Sky::sky, shared/sky_bindings.wgsl:68-70) bilinearly with repeat addressing
 * at u = (atan2(d.z, d.x) + pi) / 2 pi, v = acos(d.y) / pi
 * (unit_vector_to_panorama_coords, assets/shaders/math.wgsl:101-106), in FP32
 * (the reference samples a BC6H-compressed copy); the sun stays analytic.  The
 * texels are copied to the device (16 B each).  rgb NULL or a zero size
 * removes the panorama.  Negative or non-finite texels are refused.  A new
 * sky restarts accumulation and the ReSTIR history. */
#define APT_SKY_PANORAMA_MAX 16384u
apt_status apt_set_sky_panorama(apt_renderer* r, uint32_t width, uint32_t height, const float* rgb);

/* Radiance HDR (.hdr, RGBE) decoding: the "#?RADIANCE" / "#?RGBE" header with
 * FORMAT=32-bit_rle_rgbe, the "-Y height +X width" orientation, and flat,
 * run-length (new style) or repeat-marker (old style) scanlines.  A texel
 * (r, g, b, e) is (r, g, b) * 2^(e - 136), zero when e = 0 (the image crate's
 * decoder, which the reference loads HDR images with, crates/appearance-texture/src/asset.rs:10-14).
 * *out_rgb receives width * height * 3 floats, row major, freed with
 * apt_image_free. */
apt_status apt_image_decode_hdr(const uint8_t* bytes, size_t len, uint32_t* width, uint32_t* height,
                                float** out_rgb);
void apt_image_free(float* rgb);

/* Texture image decoding, as glTF assets load their images (the reference
 * decodes them with the image crate, crates/appearance-model/src/asset.rs:382-420):
 * PNG, or baseline JPEG (8-bit sequential Huffman; the samples of
 * libjpeg-turbo's default decompression).  width/height receive the size;
 * out_rgba, when not NULL, receives width * height RGBA8 texels (out_len
 * bytes at least), row major from the top. */
apt_status apt_image_decode_rgba8(const uint8_t* bytes, size_t len, uint32_t* width, uint32_t* height,
                                  uint8_t* out_rgba, size_t out_len);

/* PathTracerGpu::render: renders rows [row_start, row_end) of a width x height
 * frame.  Writes RGBA8 in 64x64 block-major order (resolve.wgsl:56-57,
 * block.wgsl:3-15) to out_rgba (width*(row_end-row_start)*4 bytes; width and
 * slab height must then be multiples of 64, host.rs:476) and, when
 * out_accum_rgb is not NULL, the FP32 linear radiance mean per pixel in row
 * order (3 floats per pixel).  Either output may be NULL. */
apt_status apt_render(apt_renderer* r, uint32_t width, uint32_t height, uint32_t row_start,
                      uint32_t row_end, uint8_t* out_rgba, size_t out_len, float* out_accum_rgb,
                      size_t accum_len);

/* The closure form of render (NodeRenderer::render's result_callback). */
typedef void (*apt_result_cb)(const uint8_t* pixels, size_t len, void* user);
apt_status apt_render_cb(apt_renderer* r, uint32_t width, uint32_t height, uint32_t row_start,
                         uint32_t row_end, apt_result_cb cb, void* user);

/* StartRenderData (crates/appearance-render-loop/src/host.rs:85-93), 20 B:
 * the render-node job the host sends (HostToNodeMessage::StartRender). */
typedef struct apt_start_render_data {
  uint32_t width;
  uint32_t height;
  uint32_t row_start;
  uint32_t row_end;
  uint32_t frame_idx;
} apt_start_render_data;

/* One 64x64 block of a rendered slab, the unit Node::start_render sends as a
 * RenderPartialFinished message (node.rs:55-93): the block's first frame row,
 * its column, the job's frame_idx and 64*64*4 RGBA8 bytes, valid only during
 * the call. */
typedef void (*apt_block_cb)(uint32_t row, uint32_t column_block, uint32_t frame_idx, const uint8_t* pixels,
                             size_t len, void* user);

/* Node::start_render (crates/appearance-render-loop/src/node.rs:44-105)
 * without the transport: start_render_data holds the 20 bytes of
 * StartRenderData as they arrive; the slab is rendered (NodeRenderer::render)
 * and cb receives its blocks in the order the node sends them (block rows top
 * to bottom, columns left to right).  The UDP messages stay with the
 * caller; apt_start_render_jpeg below also does the node's compression.
 * Width and slab height must be multiples of 64 (host.rs:476). */
apt_status apt_start_render(apt_renderer* r, const void* start_render_data, size_t len, apt_block_cb cb,
                            void* user);

/* JPEG compression of 64x64 RGBX blocks on the device, the node's
 * turbojpeg::compress(image, quality, Subsamp::Sub2x2) of every block it sends
 * (node.rs:65-79, quality 95; NODE_PIXEL_FORMAT RGBX, host.rs:22).  Each
 * block becomes one complete baseline JFIF file, byte for byte what
 * libjpeg-turbo's default compression (accurate integer DCT, the standard
 * Huffman tables, 4:2:0) writes for it. */
#define APT_JPEG_QUALITY 95
#define APT_JPEG_FILE_MAX 41600u   /* bytes one block's file can take (623 B header + segment) */
#define APT_JPEG_MAX_BLOCKS 16384u /* blocks per call */

/* Node::start_render with the node's compression (node.rs:44-105): as
 * apt_start_render, but each callback's pixels/len are the block's JPEG file
 * (RenderPartialFinishedData::compressed_pixel_bytes, host.rs:24-29),
 * compressed on the device at `quality` (1..100). */
apt_status apt_start_render_jpeg(apt_renderer* r, const void* start_render_data, size_t len, int quality,
                                 apt_block_cb cb, void* user);

/* turbojpeg::compress(image, quality, Sub2x2) for n_blocks blocks of 64x64
 * RGBX pixels (16384 bytes each, rows top to bottom, X ignored) in host
 * memory.  Block b's file is out[offsets[b] .. offsets[b+1]); offsets holds
 * n_blocks + 1 entries.  out_cap below the total fails with
 * APT_ERR_INVALID_ARG and offsets[n_blocks] set to the bytes needed
 * (n_blocks * APT_JPEG_FILE_MAX always suffices).  ms_device, when not NULL,
 * receives the device time of the compression kernels.  Each call allocates
 * and frees its device buffers (about 100 KB per block) and runs on the
 * library stream; a renderer's apt_start_render_jpeg keeps its own. */
apt_status apt_jpeg_compress_blocks(const uint8_t* rgbx, uint32_t n_blocks, int quality, uint8_t* out,
                                    size_t out_cap, uint64_t* offsets, float* ms_device);

/* Restarts the frame sequence at frame_idx (seeds are frame*spp + sample). */
apt_status apt_set_frame_index(apt_renderer* r, uint32_t frame_idx);

typedef struct apt_render_stats {
  uint64_t paths;            /* camera paths started in the last render */
  uint64_t extension_rays;   /* closest-hit rays traced */
  uint64_t shadow_rays;      /* any-hit rays traced */
  uint64_t segments;         /* shaded path vertices */
  double ms_total;           /* device time of the last render (hipEvents) */
  double ms_extend;          /* APT_INTEGRATOR_PLAIN: device time of the frame's extend (closest
                              * hit), shade and shadow (any hit) launches, summed over bounces and
                              * passes (hipEvents around each launch; 0 when a frame has more than
                              * 256 launches of a stage); ms_other is the rest of ms_total */
  double ms_shade;
  double ms_shadow;
  double ms_other;
  double ms_scene;           /* wall time of the last scene rebuild after a world action (BLAS builds
                              * of newly loaded assets, TLAS, placement, emitter list) or apt_scene_set
                              * (upload and device BVH build); 0 if none yet */
} apt_render_stats;
apt_status apt_get_render_stats(const apt_renderer* r, apt_render_stats* out);

/* ------------------------------------------------------------------------
 * Autotuner: per-scene choice of BVH width (2/4/8), traversal stack entries
 * kept in LDS (8/12/16) and wavefront path budget, by timing a probe frame of
 * the integrator on the device for each candidate.  Every candidate must give
 * the identical image (checked); the fastest is returned with a JSON report.
 * ------------------------------------------------------------------------ */
typedef struct apt_tune_opts {
  uint32_t probe_width;    /* 0 = 512 */
  uint32_t probe_height;   /* 0 = 256 */
  uint32_t probe_spp;      /* 0 = 2 */
  uint32_t probe_bounces;  /* 0 = 4 */
  uint32_t repeats;        /* timed renders per candidate, 0 = 2 */
} apt_tune_opts;

typedef struct apt_tune_result {
  uint32_t bvh_width;
  uint32_t lds_stack_depth;
  uint32_t queue_chunk;
  uint32_t trials;
  double best_ms;
} apt_tune_result;

apt_status apt_autotune(const apt_mesh_desc* mesh, const apt_sky* sky, const apt_camera* camera,
                        const apt_tune_opts* opts, apt_tune_result* out, char* json,
                        size_t json_len);

/* ------------------------------------------------------------------------
 * Serving: continuous batching of render requests over resident scenes.
 *
 * Replaces running one render-node job per request (NodeRenderer::render,
 * node.rs:44-105) with a server that keeps scene BVHs resident in HBM, admits
 * requests by memory footprint as earlier ones finish, and packs tiles of all
 * admitted requests (any scene, camera, resolution, spp and depth) into shared
 * wavefront batches replayed from a captured hipGraph.  Each request's result
 * does not depend on what it was batched with: paths are seeded by (request
 * seed, pixel, sample) and radiance is summed in 64-bit fixed point.
 * All apt_server_* / apt_submit / apt_poll / apt_wait / apt_cancel calls are
 * thread-safe.
 * ------------------------------------------------------------------------ */

typedef struct apt_server apt_server;

typedef struct apt_server_config {
  uint32_t path_slots;        /* paths per wavefront batch (0 = 2^24) */
  uint32_t max_active;        /* admitted requests at a time (0 = 64) */
  uint64_t accum_budget;      /* bytes of the accumulator arena allocated at creation, from which each
                                 admitted request takes 36 B per pixel (0 = 4 GiB); a request waits in
                                 the queue until a range of that size is free */
  uint32_t max_bounces;       /* deepest path any request may ask for (0 = 8) */
  uint32_t tile_paths;        /* paths per scheduling tile (0 = 65536) */
  uint32_t bvh_width;         /* 2, 4, 8 (0 = 4) */
  uint32_t lds_stack_depth;   /* 8, 12, 16 (0 = 8) */
  uint32_t policy;            /* 0 = requests of at most path_slots samples shortest remaining work
                                 first, then the larger ones in arrival order; 1 = FIFO;
                                 2 = shortest remaining work first */
  uint32_t persistent;        /* 0 = a batch is wavefront stages (raygen, extend/shade/shadow per
                                 bounce, film); 1 = one persistent kernel per batch whose lanes
                                 pull the batch's paths from a counter and run each through all
                                 its bounces.  Both give the same bits. */
} apt_server_config;

/* A render request: scene, camera, frame size, samples, depth and seed. */
typedef struct apt_request {
  uint32_t scene_id;
  uint32_t width;
  uint32_t height;
  uint32_t spp;
  uint32_t max_bounces;
  uint32_t rr_start_bounce;
  uint32_t seed;
  uint32_t _pad;
  apt_camera camera;
} apt_request;

typedef enum apt_request_state {
  APT_REQ_QUEUED = 0,
  APT_REQ_ACTIVE = 1,
  APT_REQ_DONE = 2,
  APT_REQ_CANCELLED = 3,
  APT_REQ_FAILED = 4
} apt_request_state;

typedef struct apt_result {
  uint32_t state;             /* apt_request_state */
  uint32_t status;            /* apt_status of a failed request */
  double submit_ms;           /* server clock at apt_submit */
  double admit_ms;            /* ... when admitted (footprint reserved) */
  double done_ms;             /* ... when the result was in out_rgb */
  uint64_t paths;             /* camera paths traced */
} apt_result;

apt_status apt_server_create(const apt_server_config* cfg, apt_server** out);
void apt_server_destroy(apt_server* s);
/* Uploads a scene (BVH + attributes) and keeps it resident. */
apt_status apt_server_add_scene(apt_server* s, const apt_mesh_desc* mesh, const apt_sky* sky,
                                uint32_t* scene_id);
/* Queues a request; out_rgb (width*height*3 floats, host memory owned by the
 * caller) receives the linear radiance mean when the request is done. */
apt_status apt_submit(apt_server* s, const apt_request* req, float* out_rgb, uint64_t* id);
apt_status apt_poll(apt_server* s, uint64_t id, apt_result* out);
/* Blocks until the request is done, cancelled or failed, or timeout_ms elapses
 * (APT_ERR_BUSY). */
apt_status apt_wait(apt_server* s, uint64_t id, uint32_t timeout_ms, apt_result* out);
apt_status apt_cancel(apt_server* s, uint64_t id);
/* Forgets a finished (done, cancelled or failed) request: its record is
 * freed and later calls with its id return APT_ERR_NOT_FOUND; APT_ERR_BUSY
 * while it is queued or running.  A long-running caller releases every
 * request once it has read the result. */
apt_status apt_release(apt_server* s, uint64_t id);

typedef struct apt_server_stats {
  uint64_t batches;
  uint64_t paths;
  uint64_t requests_done;
  uint32_t active;
  uint32_t queued;
  double gpu_busy_ms;         /* summed device time of the batches */
} apt_server_stats;
apt_status apt_server_get_stats(apt_server* s, apt_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* APT_H */
