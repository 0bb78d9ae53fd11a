// C ABI: error reporting, device query, BVH build and the ray-query entry
// points of include/apt.h.
#include <chrono>
#include <memory>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "apt_internal.h"
#include "tlas.h"

namespace apt {

namespace {
thread_local char g_err[1024] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void clear_error() { g_err[0] = 0; }

hipStream_t lib_stream() {
  // One non-blocking stream at the default priority per device.
  static std::mutex mu;
  static hipStream_t streams[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!streams[dev]) {
    if (hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking) != hipSuccess)
      streams[dev] = nullptr;
  }
  return streams[dev];
}

}  // namespace apt

using namespace apt;

static_assert(sizeof(apt_ray) == 32, "apt_ray must be 32 B");
static_assert(sizeof(apt_hit) == 16, "apt_hit must be 16 B");
static_assert(sizeof(apt_material) == 160, "apt_material must match MaterialDescriptor (160 B)");
static_assert(sizeof(apt_camera) == 80, "apt_camera must match CameraUpdateData (80 B)");
static_assert(sizeof(apt_start_render_data) == 20, "apt_start_render_data must match StartRenderData (20 B)");

extern "C" {

const char* apt_last_error(void) { return g_err; }

uint32_t apt_abi_version(void) { return APT_ABI_VERSION; }

apt_status apt_device_query(apt_device_info* out) {
  APT_REQUIRE(out, APT_ERR_INVALID_ARG, "apt_device_query: out is NULL");
  std::memset(out, 0, sizeof(*out));
  int dev = 0;
  APT_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  APT_HIP(hipGetDeviceProperties(&p, dev));
  std::snprintf(out->name, sizeof(out->name), "%s", p.name);
  std::snprintf(out->arch, sizeof(out->arch), "%s", p.gcnArchName);
  out->compute_units = (uint32_t)p.multiProcessorCount;
  out->wavefront_size = (uint32_t)p.warpSize;
  out->hbm_bytes = (uint64_t)p.totalGlobalMem;
  out->lds_bytes_per_block = (uint64_t)p.sharedMemPerBlock;
  return APT_OK;
}

static apt_status resolve_opts(const apt_bvh_opts* opts, apt_bvh_opts& o) {
  o = opts ? *opts : apt_bvh_opts{};
  if (o.width == 0) o.width = 4;
  if (o.max_leaf == 0) o.max_leaf = 4;
  if (o.bins == 0) o.bins = 32;
  if (o.lds_stack == 0) o.lds_stack = 8;
  if (o.refill == 0) o.refill = 32;
  APT_REQUIRE(o.refill <= 65, APT_ERR_INVALID_ARG, "bvh: refill %u outside 1..65", o.refill);
  APT_REQUIRE(o.lds_stack == 8 || o.lds_stack == 12 || o.lds_stack == 16, APT_ERR_INVALID_ARG,
              "bvh: lds_stack %u is not 8, 12 or 16", o.lds_stack);
  APT_REQUIRE(o.width == 2 || o.width == 4 || o.width == 8, APT_ERR_INVALID_ARG,
              "bvh: width %u is not 2, 4 or 8", o.width);
  APT_REQUIRE(o.max_leaf >= 1 && o.max_leaf <= 8, APT_ERR_INVALID_ARG,
              "bvh: max_leaf %u outside 1..8", o.max_leaf);
  APT_REQUIRE(o.bins >= 4 && o.bins <= 64, APT_ERR_INVALID_ARG, "bvh: bins %u outside 4..64",
              o.bins);
  APT_REQUIRE(o.builder == APT_BUILDER_HOST || o.builder == APT_BUILDER_GPU,
              APT_ERR_INVALID_ARG, "bvh: unknown builder %u", o.builder);
  return APT_OK;
}

// Builds the binary tree (host or device builder), collapses it to the wide
// layout and uploads it.  d_tri_xyz, when given, holds the same triangles on
// the device.
static apt_status build_impl(const float* tri_xyz, const float* d_tri_xyz, uint32_t n_tris,
                             const apt_bvh_opts* opts, apt_bvh** out) {
  APT_REQUIRE(n_tris > 0, APT_ERR_INVALID_ARG, "apt_bvh_build: empty triangle list");
  APT_REQUIRE(n_tris < (1u << 28), APT_ERR_INVALID_ARG, "apt_bvh_build: %u triangles exceed 2^28",
              n_tris);
  apt_bvh_opts o;
  apt_status st = resolve_opts(opts, o);
  if (st != APT_OK) return st;
  auto t0 = std::chrono::steady_clock::now();
  std::unique_ptr<apt_bvh> bvh(new apt_bvh());  // freed on every early return
  if (o.builder == APT_BUILDER_GPU) {
    float* d_tmp = nullptr;
    if (!d_tri_xyz) {
      APT_HIP(hipMalloc(&d_tmp, (size_t)n_tris * 9 * sizeof(float)));
      hipError_t e = hipMemcpyAsync(d_tmp, tri_xyz, (size_t)n_tris * 9 * sizeof(float),
                                    hipMemcpyHostToDevice, lib_stream());
      if (e != hipSuccess) {
        (void)hipFree(d_tmp);
        set_error("apt_bvh_build: upload failed: %s", hipGetErrorString(e));
        return APT_ERR_DEVICE;
      }
      d_tri_xyz = d_tmp;
    }
    st = build_wide_device(d_tri_xyz, n_tris, o.max_leaf, o.bins, o.width, lib_stream(), bvh.get());
    if (d_tmp) (void)hipFree(d_tmp);
  } else {
    std::vector<float> h;
    if (!tri_xyz) {  // device triangles, host builder
      h.resize((size_t)n_tris * 9);
      APT_HIP(hipMemcpy(h.data(), d_tri_xyz, h.size() * sizeof(float), hipMemcpyDeviceToHost));
      tri_xyz = h.data();
    }
    BvhHostResult b2;
    st = build_bvh2_host(tri_xyz, n_tris, o.max_leaf, o.bins, b2);
    if (st == APT_OK) st = upload_wide_bvh(tri_xyz, n_tris, b2, o.width, scene_pad(tri_xyz, n_tris), bvh.get());
  }
  if (st != APT_OK) return st;
  bvh->lds_stack = o.lds_stack;
  bvh->refill = o.refill;
  bvh->stats.build_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = bvh.release();
  return APT_OK;
}

apt_status apt_bvh_build(const float* tri_xyz, uint32_t n_tris, const apt_bvh_opts* opts,
                         apt_bvh** out) {
  clear_error();
  APT_REQUIRE(out && (tri_xyz || n_tris == 0), APT_ERR_INVALID_ARG, "apt_bvh_build: null argument");
  return build_impl(tri_xyz, nullptr, n_tris, opts, out);
}

apt_status apt_bvh_plan(const float* tri_xyz, uint32_t n_tris, const apt_bvh_opts* opts, apt_bvh_stats* out) {
  clear_error();
  APT_REQUIRE(out && tri_xyz && n_tris > 0, APT_ERR_INVALID_ARG, "apt_bvh_plan: null or empty argument");
  APT_REQUIRE(n_tris < (1u << 28), APT_ERR_INVALID_ARG, "apt_bvh_plan: %u triangles exceed 2^28", n_tris);
  apt_bvh_opts o;
  apt_status st = resolve_opts(opts, o);
  if (st != APT_OK) return st;
  auto t0 = std::chrono::steady_clock::now();
  BvhHostResult b2;
  st = build_bvh2_host(tri_xyz, n_tris, o.max_leaf, o.bins, b2);
  apt_bvh plan;
  if (st == APT_OK)
    st = upload_wide_bvh(tri_xyz, n_tris, b2, o.width, scene_pad(tri_xyz, n_tris), &plan, false);
  if (st != APT_OK) return st;
  plan.stats.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  *out = plan.stats;
  return check_trace_stack(plan.stats, "apt_bvh_plan");
}

uint32_t apt_trace_stack_capacity(void) { return trace_stack_capacity(); }

apt_status apt_bvh_build_device(const float* d_tri_xyz, uint32_t n_tris, const apt_bvh_opts* opts,
                                apt_bvh** out) {
  clear_error();
  APT_REQUIRE(out && d_tri_xyz && n_tris > 0, APT_ERR_INVALID_ARG,
              "apt_bvh_build_device: null or empty argument");
  return build_impl(nullptr, d_tri_xyz, n_tris, opts, out);
}

void apt_bvh_destroy(apt_bvh* bvh) {
  if (!bvh) return;
  if (bvh->d_nodes) (void)hipFree(bvh->d_nodes);
  if (bvh->d_tris) (void)hipFree(bvh->d_tris);
  if (bvh->d_counters) (void)hipFree(bvh->d_counters);
  delete bvh;
}

apt_status apt_bvh_get_stats(const apt_bvh* bvh, apt_bvh_stats* out) {
  APT_REQUIRE(bvh && out, APT_ERR_INVALID_ARG, "apt_bvh_get_stats: null argument");
  *out = bvh->stats;
  return APT_OK;
}

apt_status apt_bvh_export(const apt_bvh* bvh, void* nodes, size_t nodes_len, void* tris,
                          size_t tris_len) {
  clear_error();
  APT_REQUIRE(bvh, APT_ERR_INVALID_ARG, "apt_bvh_export: null bvh");
  const size_t nb = (size_t)bvh->n_nodes * wide_node_bytes(bvh->width);
  const size_t tb = (size_t)bvh->n_tris * 3 * sizeof(float4);
  APT_REQUIRE(!nodes || nodes_len >= nb, APT_ERR_INVALID_ARG, "apt_bvh_export: nodes need %zu bytes", nb);
  APT_REQUIRE(!tris || tris_len >= tb, APT_ERR_INVALID_ARG, "apt_bvh_export: triangles need %zu bytes", tb);
  if (nodes && nb) APT_HIP(hipMemcpy(nodes, bvh->d_nodes, nb, hipMemcpyDeviceToHost));
  if (tris && tb) APT_HIP(hipMemcpy(tris, bvh->d_tris, tb, hipMemcpyDeviceToHost));
  return APT_OK;
}

apt_status apt_trace_closest(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n,
                             apt_hit* d_hits, void* stream) {
  clear_error();
  hipStream_t s = stream ? (hipStream_t)stream : lib_stream();
  return launch_trace_closest(bvh, d_rays, n, d_hits, s);
}

apt_status apt_trace_any(const apt_bvh* bvh, const apt_ray* d_rays, uint32_t n,
                         uint32_t* d_occluded, void* stream) {
  clear_error();
  hipStream_t s = stream ? (hipStream_t)stream : lib_stream();
  return launch_trace_any(bvh, d_rays, n, d_occluded, s);
}

apt_status apt_trace_closest_tlas(const apt_tlas* t, const apt_ray* d_rays, uint32_t n, apt_hit* d_hits,
                                  void* stream) {
  clear_error();
  hipStream_t s = stream ? (hipStream_t)stream : lib_stream();
  return launch_trace_tlas(t, d_rays, n, d_hits, nullptr, false, s);
}

apt_status apt_trace_any_tlas(const apt_tlas* t, const apt_ray* d_rays, uint32_t n, uint32_t* d_occluded,
                              void* stream) {
  clear_error();
  hipStream_t s = stream ? (hipStream_t)stream : lib_stream();
  return launch_trace_tlas(t, d_rays, n, nullptr, d_occluded, true, s);
}

// One BVH or one two-level scene (the other NULL), host arrays in and out.
static apt_status trace_host(const apt_bvh* bvh, const apt_tlas* tlas, const apt_ray* rays, uint32_t n,
                             void* out, bool any) {
  clear_error();
  APT_REQUIRE((bvh || tlas) && rays && out, APT_ERR_INVALID_ARG, "trace_host: null argument");
  if (n == 0) return APT_OK;
  const size_t out_bytes = (size_t)n * (any ? sizeof(uint32_t) : sizeof(apt_hit));
  void* d_rays = nullptr;
  void* d_out = nullptr;
  APT_HIP(hipMalloc(&d_rays, (size_t)n * sizeof(apt_ray)));
  hipError_t e = hipMalloc(&d_out, out_bytes);
  if (e != hipSuccess) {
    (void)hipFree(d_rays);
    set_error("trace_host: hipMalloc failed: %s", hipGetErrorString(e));
    return APT_ERR_OOM;
  }
  hipStream_t s = lib_stream();
  apt_status st = APT_OK;
  e = hipMemcpyAsync(d_rays, rays, (size_t)n * sizeof(apt_ray), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    if (tlas)
      st = launch_trace_tlas(tlas, (const apt_ray*)d_rays, n, any ? nullptr : (apt_hit*)d_out,
                             any ? (uint32_t*)d_out : nullptr, any, s);
    else
      st = any ? launch_trace_any(bvh, (const apt_ray*)d_rays, n, (uint32_t*)d_out, s)
               : launch_trace_closest(bvh, (const apt_ray*)d_rays, n, (apt_hit*)d_out, s);
    if (st == APT_OK) e = hipMemcpyAsync(out, d_out, out_bytes, hipMemcpyDeviceToHost, s);
    if (st == APT_OK && e == hipSuccess) e = hipStreamSynchronize(s);
  }
  (void)hipFree(d_rays);
  (void)hipFree(d_out);
  if (st != APT_OK) return st;
  if (e != hipSuccess) {
    set_error("trace_host: %s", hipGetErrorString(e));
    return APT_ERR_DEVICE;
  }
  return APT_OK;
}

apt_status apt_trace_closest_host(const apt_bvh* bvh, const apt_ray* rays, uint32_t n,
                                  apt_hit* hits) {
  return trace_host(bvh, nullptr, rays, n, hits, false);
}

apt_status apt_trace_any_host(const apt_bvh* bvh, const apt_ray* rays, uint32_t n,
                              uint32_t* occluded) {
  return trace_host(bvh, nullptr, rays, n, occluded, true);
}

apt_status apt_trace_closest_tlas_host(const apt_tlas* t, const apt_ray* rays, uint32_t n, apt_hit* hits) {
  APT_REQUIRE(t, APT_ERR_INVALID_ARG, "apt_trace_closest_tlas_host: null scene");
  return trace_host(nullptr, t, rays, n, hits, false);
}

apt_status apt_trace_any_tlas_host(const apt_tlas* t, const apt_ray* rays, uint32_t n, uint32_t* occluded) {
  APT_REQUIRE(t, APT_ERR_INVALID_ARG, "apt_trace_any_tlas_host: null scene");
  return trace_host(nullptr, t, rays, n, occluded, true);
}

}  // extern "C"
