//! Raw bindings to libapt (`include/apt.h`), the MI355X path-tracing hot path
//! behind a C ABI.  Every type is `#[repr(C)]` and mirrors its C struct field
//! for field; every function is declared as in the header.  The safe
//! `NodeRenderer` on top of it is in `node.rs` (feature `node-renderer`).
//!
//! What each entry point replaces in Appearance:
//! * `apt_create` / `apt_world_action` / `apt_render_cb`: `PathTracerGpu::{new,
//!   handle_visible_world_action, render}` behind `NodeRenderer`
//!   (crates/appearance-render-loop/src/node.rs:14-25).
//! * `apt_start_render`: `Node::start_render` (node.rs:44-105) without the
//!   transport; `apt_start_render_jpeg` with the node's per-block JPEG
//!   compression (node.rs:65-79), `apt_jpeg_compress_blocks` that compression
//!   alone.
//! * `apt_bvh_*` / `apt_trace_*`: the driver's BLAS build and ray query;
//!   `apt_tlas_*`: its TLAS over BLAS instances.
//! * `apt_server_*`, `apt_submit`, ...: the serving loop (no counterpart).
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

pub mod node;

// ---- status, versions --------------------------------------------------------

pub type apt_status = c_int;
pub const APT_OK: apt_status = 0;
pub const APT_ERR_INVALID_ARG: apt_status = 1;
pub const APT_ERR_OOM: apt_status = 2;
pub const APT_ERR_DEVICE: apt_status = 3;
pub const APT_ERR_UNSUPPORTED: apt_status = 4;
pub const APT_ERR_BUSY: apt_status = 5;
pub const APT_ERR_NOT_FOUND: apt_status = 6;

pub const APT_ABI_VERSION: u32 = 1;
pub const APT_MISS: u32 = 0xFFFF_FFFF;
pub const APT_TRACE_NOALLOC_MAX_RAYS: u32 = 65536;
pub const APT_JPEG_QUALITY: u32 = 95;
pub const APT_JPEG_FILE_MAX: u32 = 41600;
pub const APT_JPEG_MAX_BLOCKS: u32 = 16384;
pub const APT_SKY_PANORAMA_MAX: u32 = 16384;
pub const APT_TLAS_MAX_INSTANCES: u32 = 8192;
pub const APT_INVALID_TEXTURE: u32 = 0xFFFF_FFFF;
pub const APT_BSDF_DISNEY: u32 = 0;
pub const APT_BSDF_LAMBERT: u32 = 1;

pub type apt_builder = c_int;
pub const APT_BUILDER_HOST: apt_builder = 0;
pub const APT_BUILDER_GPU: apt_builder = 1;

pub type apt_integrator = c_int;
pub const APT_INTEGRATOR_PLAIN: apt_integrator = 0;
pub const APT_INTEGRATOR_RESTIR: apt_integrator = 1;

pub type apt_request_state = c_int;
pub const APT_REQ_QUEUED: apt_request_state = 0;
pub const APT_REQ_ACTIVE: apt_request_state = 1;
pub const APT_REQ_DONE: apt_request_state = 2;
pub const APT_REQ_CANCELLED: apt_request_state = 3;
pub const APT_REQ_FAILED: apt_request_state = 4;

// ---- opaque handles ----------------------------------------------------------

#[repr(C)]
pub struct apt_bvh {
    _p: [u8; 0],
}
#[repr(C)]
pub struct apt_scene_data {
    _p: [u8; 0],
}
#[repr(C)]
pub struct apt_renderer {
    _p: [u8; 0],
}
#[repr(C)]
pub struct apt_tlas {
    _p: [u8; 0],
}
#[repr(C)]
pub struct apt_server {
    _p: [u8; 0],
}

// ---- plain data --------------------------------------------------------------

#[repr(C)]
#[derive(Clone, Copy)]
pub struct apt_device_info {
    pub name: [c_char; 64],
    pub arch: [c_char; 32],
    pub compute_units: u32,
    pub wavefront_size: u32,
    pub hbm_bytes: u64,
    pub lds_bytes_per_block: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_ray {
    pub org: [f32; 3],
    pub tmin: f32,
    pub dir: [f32; 3],
    pub tmax: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_hit {
    pub t: f32,
    pub u: f32,
    pub v: f32,
    pub prim: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_bvh_opts {
    pub width: u32,
    pub max_leaf: u32,
    pub bins: u32,
    pub builder: u32,
    pub lds_stack: u32,
    pub refill: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_bvh_stats {
    pub n_tris: u32,
    pub n_nodes: u32,
    pub n_leaves: u32,
    pub width: u32,
    pub depth: u32,
    pub max_stack: u32,
    pub device_bytes: u64,
    pub build_ms: f64,
    pub sah_cost: f64,
}

/// One instance of a two-level scene (the reference's TlasInstance,
/// scene_resources/mod.rs:177-191), 64 B: object-to-world 3x4 row major.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_instance {
    pub transform: [f32; 12],
    pub blas: u32,
    pub custom_index: u32,
    pub _pad: [u32; 2],
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_tlas_stats {
    pub n_blas: u32,
    pub n_instances: u32,
    pub n_tlas_nodes: u32,
    pub width: u32,
    pub max_stack: u32,
    pub _pad: u32,
    pub blas_tris: u64,
    pub device_bytes: u64,
    pub ms_last_update: f64,
}

/// The reference's MaterialDescriptor (src/scene_resources/material_pool.rs:13-46), 160 B.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_material {
    pub color: [f32; 3],
    pub color_texture: u32,
    pub metallic: f32,
    pub roughness: f32,
    pub metallic_roughness_texture: u32,
    pub normal_scale: f32,
    pub emission: [f32; 3],
    pub normal_texture: u32,
    pub emission_texture: u32,
    pub transmission: f32,
    pub eta: f32,
    pub subsurface: f32,
    pub absorption: [f32; 3],
    pub specular: f32,
    pub specular_tint: [f32; 3],
    pub anisotropic: f32,
    pub sheen: f32,
    pub sheen_texture: u32,
    pub clearcoat: f32,
    pub clearcoat_texture: u32,
    pub clearcoat_roughness: f32,
    pub clearcoat_roughness_texture: u32,
    pub alpha_cutoff: f32,
    pub sheen_tint_texture: u32,
    pub clearcoat_normal_texture: u32,
    pub bsdf_model: u32,
    pub _padding1: u32,
    pub _padding2: u32,
    pub sheen_tint: [f32; 3],
    pub transmission_texture: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_sky {
    pub sun_direction: [f32; 3],
    pub sun_size: f32,
    pub sun_color: [f32; 3],
    pub sun_intensity: f32,
    pub zenith: [f32; 3],
    pub sun_enabled: f32,
    pub horizon: [f32; 3],
    pub _pad0: f32,
    pub ground: [f32; 3],
    pub _pad1: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct apt_mesh_desc {
    pub positions: *const f32,
    pub normals: *const f32,
    pub material_ids: *const u32,
    pub n_tris: u32,
    pub n_materials: u32,
    pub materials: *const apt_material,
}

/// One decoded texture: width * height RGBA8 texels, row major.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct apt_texture {
    pub width: u32,
    pub height: u32,
    pub rgba8: *const u32,
}

/// Per-vertex uvs, tangent frames and the scene's textures (NULL / 0 when absent).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct apt_mesh_attribs {
    pub uvs: *const f32,
    pub tangents: *const f32,
    pub bitangents: *const f32,
    pub textures: *const apt_texture,
    pub n_textures: u32,
    pub _pad: u32,
}

/// CameraUpdateData (crates/appearance-world/src/visible_world_action.rs:9-15), 80 B.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_camera {
    pub transform: [f32; 16],
    pub fov_deg: f32,
    pub near_plane: f32,
    pub far_plane: f32,
    pub _padding: u32,
}

/// PathTracerGpuConfig (src/lib.rs:219-245), made public.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_config {
    pub max_bounces: u32,
    pub sample_count: u32,
    pub accum_frames: u32,
    pub rr_start_bounce: u32,
    pub integrator: u32,
    pub restir_di: u32,
    pub restir_gi: u32,
    pub svgf: u32,
    pub firefly_filter: u32,
    pub taa: u32,
    pub di_candidates: u32,
    pub di_mcap: u32,
    pub di_spatial_neighbours: u32,
    pub di_spatial_radius: f32,
    pub gi_mcap: u32,
    pub svgf_iterations: u32,
    pub svgf_history: u32,
    pub bvh_width: u32,
    pub queue_chunk: u32,
    pub lds_stack_depth: u32,
    pub seed: u32,
    pub frame_idx: u32,
    pub gi_spatial_passes: u32,
}

/// StartRenderData (crates/appearance-render-loop/src/host.rs:85-93), 20 B.
#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_start_render_data {
    pub width: u32,
    pub height: u32,
    pub row_start: u32,
    pub row_end: u32,
    pub frame_idx: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_render_stats {
    pub paths: u64,
    pub extension_rays: u64,
    pub shadow_rays: u64,
    pub segments: u64,
    pub ms_total: f64,
    pub ms_extend: f64,
    pub ms_shade: f64,
    pub ms_shadow: f64,
    pub ms_other: f64,
    pub ms_scene: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_tune_opts {
    pub probe_width: u32,
    pub probe_height: u32,
    pub probe_spp: u32,
    pub probe_bounces: u32,
    pub repeats: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_tune_result {
    pub bvh_width: u32,
    pub lds_stack_depth: u32,
    pub queue_chunk: u32,
    pub trials: u32,
    pub best_ms: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_server_config {
    pub path_slots: u32,
    pub max_active: u32,
    pub accum_budget: u64,
    pub max_bounces: u32,
    pub tile_paths: u32,
    pub bvh_width: u32,
    pub lds_stack_depth: u32,
    pub policy: u32,
    pub persistent: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_request {
    pub scene_id: u32,
    pub width: u32,
    pub height: u32,
    pub spp: u32,
    pub max_bounces: u32,
    pub rr_start_bounce: u32,
    pub seed: u32,
    pub _pad: u32,
    pub camera: apt_camera,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_result {
    pub state: u32,
    pub status: u32,
    pub submit_ms: f64,
    pub admit_ms: f64,
    pub done_ms: f64,
    pub paths: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Default, Debug)]
pub struct apt_server_stats {
    pub batches: u64,
    pub paths: u64,
    pub requests_done: u64,
    pub active: u32,
    pub queued: u32,
    pub gpu_busy_ms: f64,
}

// ---- callbacks ---------------------------------------------------------------

/// NodeRenderer::render's result_callback: the RGBA8 slab in 64x64 block order.
pub type apt_result_cb = extern "C" fn(pixels: *const u8, len: usize, user: *mut c_void);
/// One 64x64 block of a slab, as Node::start_render sends it.
pub type apt_block_cb = extern "C" fn(
    row: u32,
    column_block: u32,
    frame_idx: u32,
    pixels: *const u8,
    len: usize,
    user: *mut c_void,
);

// ---- functions ---------------------------------------------------------------

#[link(name = "apt")]
extern "C" {
    pub fn apt_last_error() -> *const c_char;
    pub fn apt_abi_version() -> u32;
    pub fn apt_device_query(out: *mut apt_device_info) -> apt_status;

    // ray-query level
    pub fn apt_bvh_build(tri_xyz: *const f32, n_tris: u32, opts: *const apt_bvh_opts, out: *mut *mut apt_bvh)
        -> apt_status;
    pub fn apt_bvh_build_device(d_tri_xyz: *const f32, n_tris: u32, opts: *const apt_bvh_opts,
                                out: *mut *mut apt_bvh) -> apt_status;
    pub fn apt_bvh_plan(tri_xyz: *const f32, n_tris: u32, opts: *const apt_bvh_opts, out: *mut apt_bvh_stats)
        -> apt_status;
    pub fn apt_trace_stack_capacity() -> u32;
    pub fn apt_bvh_destroy(bvh: *mut apt_bvh);
    pub fn apt_bvh_get_stats(bvh: *const apt_bvh, out: *mut apt_bvh_stats) -> apt_status;
    pub fn apt_bvh_export(bvh: *const apt_bvh, nodes: *mut c_void, nodes_len: usize, tris: *mut c_void,
                          tris_len: usize) -> apt_status;
    pub fn apt_trace_closest(bvh: *const apt_bvh, d_rays: *const apt_ray, n: u32, d_hits: *mut apt_hit,
                             stream: *mut c_void) -> apt_status;
    pub fn apt_trace_any(bvh: *const apt_bvh, d_rays: *const apt_ray, n: u32, d_occluded: *mut u32,
                         stream: *mut c_void) -> apt_status;
    pub fn apt_trace_closest_host(bvh: *const apt_bvh, rays: *const apt_ray, n: u32, hits: *mut apt_hit)
        -> apt_status;
    pub fn apt_trace_any_host(bvh: *const apt_bvh, rays: *const apt_ray, n: u32, occluded: *mut u32)
        -> apt_status;
    pub fn apt_trace_closest_host_noalloc(bvh: *const apt_bvh, rays: *const apt_ray, n: u32,
                                          hits: *mut apt_hit) -> apt_status;

    // two-level scenes
    pub fn apt_tlas_create(opts: *const apt_bvh_opts, out: *mut *mut apt_tlas) -> apt_status;
    pub fn apt_tlas_add_blas(t: *mut apt_tlas, blas: *const apt_bvh, blas_index: *mut u32) -> apt_status;
    pub fn apt_tlas_set_instances(t: *mut apt_tlas, instances: *const apt_instance, n: u32) -> apt_status;
    pub fn apt_tlas_destroy(t: *mut apt_tlas);
    pub fn apt_tlas_get_stats(t: *const apt_tlas, out: *mut apt_tlas_stats) -> apt_status;
    pub fn apt_trace_closest_tlas(t: *const apt_tlas, d_rays: *const apt_ray, n: u32, d_hits: *mut apt_hit,
                                  stream: *mut c_void) -> apt_status;
    pub fn apt_trace_any_tlas(t: *const apt_tlas, d_rays: *const apt_ray, n: u32, d_occluded: *mut u32,
                              stream: *mut c_void) -> apt_status;
    pub fn apt_trace_closest_tlas_host(t: *const apt_tlas, rays: *const apt_ray, n: u32, hits: *mut apt_hit)
        -> apt_status;
    pub fn apt_trace_any_tlas_host(t: *const apt_tlas, rays: *const apt_ray, n: u32, occluded: *mut u32)
        -> apt_status;

    // scene data
    pub fn apt_material_default(m: *mut apt_material);
    pub fn apt_scene_generate(name: *const c_char, seed: u32, out: *mut *mut apt_scene_data) -> apt_status;
    pub fn apt_scene_load_gltf(path: *const c_char, out: *mut *mut apt_scene_data) -> apt_status;
    pub fn apt_scene_data_view(s: *const apt_scene_data, mesh: *mut apt_mesh_desc, camera: *mut apt_camera,
                               sky: *mut apt_sky) -> apt_status;
    pub fn apt_scene_data_attribs(s: *const apt_scene_data, out: *mut apt_mesh_attribs) -> apt_status;
    pub fn apt_scene_data_destroy(s: *mut apt_scene_data);
    pub fn apt_rays_bvh_stress(seed: u32, side: u32, n_incoherent: u32, out: *mut apt_ray) -> apt_status;

    // renderer (PathTracerGpu / NodeRenderer)
    pub fn apt_config_default(cfg: *mut apt_config);
    pub fn apt_create(cfg: *const apt_config, out: *mut *mut apt_renderer) -> apt_status;
    pub fn apt_destroy(r: *mut apt_renderer);
    pub fn apt_set_config(r: *mut apt_renderer, cfg: *const apt_config) -> apt_status;
    pub fn apt_world_action(r: *mut apt_renderer, ty: u32, bytes: *const c_void, len: usize) -> apt_status;
    pub fn apt_scene_set(r: *mut apt_renderer, mesh: *const apt_mesh_desc, sky: *const apt_sky) -> apt_status;
    pub fn apt_scene_set_attribs(r: *mut apt_renderer, mesh: *const apt_mesh_desc, attribs: *const apt_mesh_attribs,
                                 sky: *const apt_sky) -> apt_status;
    pub fn apt_set_sky_panorama(r: *mut apt_renderer, width: u32, height: u32, rgb: *const f32) -> apt_status;
    pub fn apt_image_decode_hdr(bytes: *const u8, len: usize, width: *mut u32, height: *mut u32,
                                out_rgb: *mut *mut f32) -> apt_status;
    pub fn apt_image_free(rgb: *mut f32);
    pub fn apt_image_decode_rgba8(bytes: *const u8, len: usize, width: *mut u32, height: *mut u32,
                                  out_rgba: *mut u8, out_len: usize) -> apt_status;
    pub fn apt_render(r: *mut apt_renderer, width: u32, height: u32, row_start: u32, row_end: u32,
                      out_rgba: *mut u8, out_len: usize, out_accum_rgb: *mut f32, accum_len: usize)
        -> apt_status;
    pub fn apt_render_cb(r: *mut apt_renderer, width: u32, height: u32, row_start: u32, row_end: u32,
                         cb: apt_result_cb, user: *mut c_void) -> apt_status;
    pub fn apt_start_render(r: *mut apt_renderer, start_render_data: *const c_void, len: usize,
                            cb: apt_block_cb, user: *mut c_void) -> apt_status;
    pub fn apt_start_render_jpeg(r: *mut apt_renderer, start_render_data: *const c_void, len: usize,
                                 quality: c_int, cb: apt_block_cb, user: *mut c_void) -> apt_status;
    pub fn apt_jpeg_compress_blocks(rgbx: *const u8, n_blocks: u32, quality: c_int, out: *mut u8,
                                    out_cap: usize, offsets: *mut u64, ms_device: *mut f32) -> apt_status;
    pub fn apt_set_frame_index(r: *mut apt_renderer, frame_idx: u32) -> apt_status;
    pub fn apt_get_render_stats(r: *const apt_renderer, out: *mut apt_render_stats) -> apt_status;

    // autotuner
    pub fn apt_autotune(mesh: *const apt_mesh_desc, sky: *const apt_sky, camera: *const apt_camera,
                        opts: *const apt_tune_opts, out: *mut apt_tune_result, json: *mut c_char,
                        json_len: usize) -> apt_status;

    // serving loop (thread-safe)
    pub fn apt_server_create(cfg: *const apt_server_config, out: *mut *mut apt_server) -> apt_status;
    pub fn apt_server_destroy(s: *mut apt_server);
    pub fn apt_server_add_scene(s: *mut apt_server, mesh: *const apt_mesh_desc, sky: *const apt_sky,
                                scene_id: *mut u32) -> apt_status;
    pub fn apt_submit(s: *mut apt_server, req: *const apt_request, out_rgb: *mut f32, id: *mut u64)
        -> apt_status;
    pub fn apt_poll(s: *mut apt_server, id: u64, out: *mut apt_result) -> apt_status;
    pub fn apt_wait(s: *mut apt_server, id: u64, timeout_ms: u32, out: *mut apt_result) -> apt_status;
    pub fn apt_cancel(s: *mut apt_server, id: u64) -> apt_status;
    pub fn apt_release(s: *mut apt_server, id: u64) -> apt_status;
    pub fn apt_server_get_stats(s: *mut apt_server, out: *mut apt_server_stats) -> apt_status;
}

// Layout checks against the sizes include/apt.h fixes.
const _: () = assert!(std::mem::size_of::<apt_ray>() == 32);
const _: () = assert!(std::mem::size_of::<apt_hit>() == 16);
const _: () = assert!(std::mem::size_of::<apt_material>() == 160);
const _: () = assert!(std::mem::size_of::<apt_camera>() == 80);
const _: () = assert!(std::mem::size_of::<apt_start_render_data>() == 20);
const _: () = assert!(std::mem::size_of::<apt_sky>() == 80);
