// Two-level scenes: BLAS arena, relocation and the TLAS over instances.
//
// The reference builds one BLAS per mesh when a model is first spawned
// (src/scene_resources/scene_model.rs:53-88) and rebuilds the TLAS from the
// instance transforms every frame (scene_resources/mod.rs:218-276, at most
// MAX_TLAS_INSTANCES = 8192, :20).  Here a BLAS is a wide BVH built by the
// device builder in object space and copied once into the scene's arenas
// (refs relocated by a kernel); a TLAS update is a host binned-SAH build over
// the instances' world boxes (a few thousand boxes at most, microseconds) plus
// one upload of its nodes and instance records.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "tlas.h"

namespace apt {

bool affine_inverse(const float* o2w, float* w2o, double* cond) {
  const double a = o2w[0], b = o2w[1], c = o2w[2], tx = o2w[3];
  const double d = o2w[4], e = o2w[5], f = o2w[6], ty = o2w[7];
  const double g = o2w[8], h = o2w[9], i = o2w[10], tz = o2w[11];
  const double A = e * i - f * h, B = f * g - d * i, C = d * h - e * g;
  const double det = a * A + b * B + c * C;
  if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
  // rows of the inverse: the transposed cofactor matrix over det
  const double inv[9] = {A / det, (c * h - b * i) / det, (b * f - c * e) / det,
                         B / det, (a * i - c * g) / det, (c * d - a * f) / det,
                         C / det, (b * g - a * h) / det, (a * e - b * d) / det};
  double na = 0.0, ni = 0.0;
  const double m[9] = {a, b, c, d, e, f, g, h, i};
  for (int k = 0; k < 9; ++k) {
    na += m[k] * m[k];
    ni += inv[k] * inv[k];
  }
  for (int r = 0; r < 3; ++r) {
    for (int k = 0; k < 3; ++k) w2o[4 * r + k] = (float)inv[3 * r + k];
    w2o[4 * r + 3] = (float)(-((inv[3 * r] * tx + inv[3 * r + 1] * ty) + inv[3 * r + 2] * tz));
  }
  for (int k = 0; k < 12; ++k)
    if (!std::isfinite(w2o[k])) return false;
  if (cond) *cond = std::sqrt(na) * std::sqrt(ni);
  return true;
}

namespace {

// Copies a BLAS's nodes into the arena: internal refs move by the BLAS's node
// base, leaf refs by its triangle base (the first-triangle field).
template <int W>
__global__ void k_relocate(const WNode<W>* __restrict__ src, WNode<W>* __restrict__ dst, uint32_t n,
                           uint32_t node_base, uint32_t tri_base) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    WNode<W> nd = src[i];
#pragma unroll
    for (int k = 0; k < W; ++k) {
      const uint32_t r = nd.child[k];
      if (r != REF_EMPTY) nd.child[k] = ref_is_leaf(r) ? r + tri_base : r + node_base;
    }
    dst[i] = nd;
  }
}

uint32_t relocate_ref(uint32_t r, uint32_t node_base, uint32_t tri_base) {
  return r == REF_EMPTY ? r : ref_is_leaf(r) ? r + tri_base : r + node_base;
}

apt_status dev_alloc(void** p, size_t bytes) {
  const hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
  if (e != hipSuccess) {
    *p = nullptr;
    set_error("two-level scene: hipMalloc(%zu B) failed: %s", bytes, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? APT_ERR_OOM : APT_ERR_DEVICE;
  }
  return APT_OK;
}

// Grows an arena to hold `need` elements of `elem` bytes, keeping the first
// `used` elements.
apt_status grow(void*& buf, size_t& cap, size_t used, size_t need, size_t elem, hipStream_t s) {
  if (need <= cap && buf) return APT_OK;
  const size_t ncap = std::max(need, cap + cap / 2);
  void* nb = nullptr;
  apt_status st = dev_alloc(&nb, ncap * elem);
  if (st != APT_OK) return st;
  if (buf && used) {
    const hipError_t e = hipMemcpyAsync(nb, buf, used * elem, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) (void)hipStreamSynchronize(s);
    if (e != hipSuccess) {
      (void)hipFree(nb);
      set_error("two-level scene: arena copy failed: %s", hipGetErrorString(e));
      return APT_ERR_DEVICE;
    }
  }
  if (buf) (void)hipFree(buf);
  buf = nb;
  cap = ncap;
  return APT_OK;
}

// Object-space bounds of a BLAS: the decoded child boxes of its root node, or
// the triangles of a root leaf.
template <int W>
apt_status blas_bounds(const void* d_nodes, const float4* d_tris, uint32_t root, float* lo, float* hi) {
  double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
  if (ref_is_leaf(root)) {
    const uint32_t first = ref_leaf_first(root), cnt = ref_leaf_count(root);
    std::vector<float4> t(3 * (size_t)cnt);
    APT_HIP(hipMemcpy(t.data(), d_tris + 3 * (size_t)first, t.size() * sizeof(float4), hipMemcpyDeviceToHost));
    for (const float4& v : t) {
      const double p[3] = {v.x, v.y, v.z};
      for (int a = 0; a < 3; ++a) {
        l[a] = std::min(l[a], p[a]);
        h[a] = std::max(h[a], p[a]);
      }
    }
  } else {
    WNode<W> nd;
    APT_HIP(hipMemcpy(&nd, static_cast<const WNode<W>*>(d_nodes) + root, sizeof(nd), hipMemcpyDeviceToHost));
    const double org[3] = {nd.ox, nd.oy, nd.oz};
    const double step[3] = {qstep(nd.ex), qstep(nd.ey), qstep(nd.ez)};
    for (int k = 0; k < W; ++k) {
      if (nd.child[k] == REF_EMPTY) continue;
      for (int a = 0; a < 3; ++a) {
        l[a] = std::min(l[a], org[a] + nd.q[a][k] * step[a]);
        h[a] = std::max(h[a], org[a] + nd.q[a + 3][k] * step[a]);
      }
    }
  }
  for (int a = 0; a < 3; ++a) {
    lo[a] = std::nextafter((float)l[a], -INFINITY);
    hi[a] = std::nextafter((float)h[a], INFINITY);
  }
  return APT_OK;
}

bool is_identity(const float* m) {
  static const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
  for (int k = 0; k < 12; ++k)
    if (!(m[k] == id[k]) || std::signbit(m[k])) return false;
  return true;
}

}  // namespace

apt_status tlas_add_blas_device(apt_tlas* t, const void* d_nodes, uint32_t n_nodes, const float4* d_tris,
                                uint32_t n_tris, uint32_t root, uint32_t max_stack, uint32_t* index,
                                hipStream_t s) {
  const size_t nb = wide_node_bytes(t->width);
  if (t->node_used == 0) t->node_used = TLAS_NODE_CAP;
  APT_REQUIRE(t->tri_used + n_tris < (1u << 28), APT_ERR_INVALID_ARG,
              "two-level scene: %zu BLAS triangles exceed 2^28", t->tri_used + n_tris);
  APT_REQUIRE(t->node_used + n_nodes < 0x80000000u, APT_ERR_INVALID_ARG, "two-level scene: node arena full");
  apt_status st = grow(t->d_nodes, t->node_cap, t->node_used, t->node_used + n_nodes, nb, s);
  void* tris = t->d_tris;
  if (st == APT_OK) st = grow(tris, t->tri_cap, t->tri_used, t->tri_used + n_tris, 3 * sizeof(float4), s);
  t->d_tris = static_cast<float4*>(tris);
  if (st != APT_OK) return st;
  apt_tlas::Blas b{};
  b.node_base = (uint32_t)t->node_used;
  b.tri_base = (uint32_t)t->tri_used;
  b.n_nodes = n_nodes;
  b.n_tris = n_tris;
  b.root = relocate_ref(root, b.node_base, b.tri_base);
  b.max_stack = max_stack;
  switch (t->width) {
    case 2: st = blas_bounds<2>(d_nodes, d_tris, root, b.lo, b.hi); break;
    case 4: st = blas_bounds<4>(d_nodes, d_tris, root, b.lo, b.hi); break;
    default: st = blas_bounds<8>(d_nodes, d_tris, root, b.lo, b.hi); break;
  }
  if (st != APT_OK) return st;
  if (n_nodes) {
    const dim3 grid(std::min<uint32_t>((n_nodes + 255) / 256, 4096));
    char* dst = static_cast<char*>(t->d_nodes) + (size_t)b.node_base * nb;
    switch (t->width) {
      case 2:
        hipLaunchKernelGGL(k_relocate<2>, grid, dim3(256), 0, s, static_cast<const WNode<2>*>(d_nodes),
                           reinterpret_cast<WNode<2>*>(dst), n_nodes, b.node_base, b.tri_base);
        break;
      case 4:
        hipLaunchKernelGGL(k_relocate<4>, grid, dim3(256), 0, s, static_cast<const WNode<4>*>(d_nodes),
                           reinterpret_cast<WNode<4>*>(dst), n_nodes, b.node_base, b.tri_base);
        break;
      default:
        hipLaunchKernelGGL(k_relocate<8>, grid, dim3(256), 0, s, static_cast<const WNode<8>*>(d_nodes),
                           reinterpret_cast<WNode<8>*>(dst), n_nodes, b.node_base, b.tri_base);
        break;
    }
    APT_HIP(hipGetLastError());
  }
  if (n_tris)
    APT_HIP(hipMemcpyAsync(t->d_tris + 3 * (size_t)b.tri_base, d_tris, (size_t)n_tris * 3 * sizeof(float4),
                           hipMemcpyDeviceToDevice, s));
  APT_HIP(hipStreamSynchronize(s));
  t->node_used += n_nodes;
  t->tri_used += n_tris;
  *index = (uint32_t)t->blas.size();
  t->blas.push_back(b);
  return APT_OK;
}

}  // namespace apt

using namespace apt;

extern "C" {

apt_status apt_tlas_create(const apt_bvh_opts* opts, apt_tlas** out) {
  clear_error();
  APT_REQUIRE(out, APT_ERR_INVALID_ARG, "apt_tlas_create: out is NULL");
  apt_bvh_opts o{};
  if (opts) o = *opts;
  const uint32_t width = o.width ? o.width : 4u;
  const uint32_t lds = o.lds_stack ? o.lds_stack : 8u;
  const uint32_t refill = o.refill ? o.refill : 32u;
  APT_REQUIRE(width == 2 || width == 4 || width == 8, APT_ERR_INVALID_ARG, "apt_tlas_create: width %u", width);
  APT_REQUIRE(lds == 8 || lds == 16 || lds == 24, APT_ERR_INVALID_ARG, "apt_tlas_create: lds_stack %u", lds);
  APT_REQUIRE(refill >= 1 && refill <= 65, APT_ERR_INVALID_ARG, "apt_tlas_create: refill %u", refill);
  apt_tlas* t = new apt_tlas();
  t->width = width;
  t->lds_stack = lds;
  t->refill = refill;
  t->node_used = TLAS_NODE_CAP;
  apt_status st = dev_alloc(&t->d_nodes, (size_t)TLAS_NODE_CAP * wide_node_bytes(width));
  if (st == APT_OK) {
    t->node_cap = TLAS_NODE_CAP;
    st = dev_alloc((void**)&t->d_inst, (size_t)TLAS_MAX_INSTANCES * sizeof(InstRec));
  }
  if (st == APT_OK) st = dev_alloc((void**)&t->d_counters, (size_t)APT_BVH_COUNTERS * POOL_WORDS * sizeof(uint32_t));
  if (st != APT_OK) {
    apt_tlas_destroy(t);
    return st;
  }
  *out = t;
  return APT_OK;
}

void apt_tlas_destroy(apt_tlas* t) {
  if (!t) return;
  for (void* p : {t->d_nodes, (void*)t->d_tris, (void*)t->d_inst, (void*)t->d_counters})
    if (p) (void)hipFree(p);
  delete t;
}

apt_status apt_tlas_add_blas(apt_tlas* t, const apt_bvh* blas, uint32_t* blas_index) {
  clear_error();
  APT_REQUIRE(t && blas && blas_index, APT_ERR_INVALID_ARG, "apt_tlas_add_blas: null argument");
  APT_REQUIRE(blas->width == t->width, APT_ERR_INVALID_ARG,
              "apt_tlas_add_blas: the BLAS is %u wide, the scene %u", blas->width, t->width);
  APT_REQUIRE(blas->n_tris > 0, APT_ERR_INVALID_ARG, "apt_tlas_add_blas: empty BLAS");
  return tlas_add_blas_device(t, blas->d_nodes, blas->n_nodes, blas->d_tris, blas->n_tris, blas->root,
                              blas->stats.max_stack, blas_index, lib_stream());
}

apt_status apt_tlas_set_instances(apt_tlas* t, const apt_instance* inst, uint32_t n) {
  clear_error();
  APT_REQUIRE(t && (inst || n == 0), APT_ERR_INVALID_ARG, "apt_tlas_set_instances: null argument");
  APT_REQUIRE(n <= TLAS_MAX_INSTANCES, APT_ERR_INVALID_ARG, "apt_tlas_set_instances: %u instances exceed %u", n,
              TLAS_MAX_INSTANCES);
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<InstRec> rec(n);
  std::vector<double> box(6 * (size_t)n), cond(n);
  uint32_t blas_stack = 0;
  double R = 0.0;
  for (uint32_t k = 0; k < n; ++k) {
    const apt_instance& in = inst[k];
    APT_REQUIRE(in.blas < t->blas.size(), APT_ERR_INVALID_ARG, "instance %u: BLAS %u of %zu", k, in.blas,
                t->blas.size());
    const apt_tlas::Blas& b = t->blas[in.blas];
    APT_REQUIRE((uint64_t)in.custom_index + b.n_tris < 0xFFFFFFFFull, APT_ERR_INVALID_ARG,
                "instance %u: custom_index %u + %u triangles overflow the prim ids", k, in.custom_index, b.n_tris);
    for (int j = 0; j < 12; ++j)
      APT_REQUIRE(std::isfinite(in.transform[j]), APT_ERR_INVALID_ARG, "instance %u: non-finite transform", k);
    InstRec& r = rec[k];
    std::memset(&r, 0, sizeof(r));
    if (is_identity(in.transform)) {
      std::memcpy(r.w2o, in.transform, sizeof(r.w2o));
      r.flags = INST_IDENTITY;
      cond[k] = 1.0;
    } else {
      APT_REQUIRE(affine_inverse(in.transform, r.w2o, &cond[k]), APT_ERR_UNSUPPORTED,
                  "instance %u: the transform is singular", k);
    }
    r.blas_root = b.root;
    r.prim_base = in.custom_index;
    blas_stack = std::max(blas_stack, b.max_stack);
    // world box: the 8 corners of the BLAS bounds through the transform
    double* bx = &box[6 * (size_t)k];
    for (int a = 0; a < 3; ++a) {
      bx[a] = INFINITY;
      bx[3 + a] = -INFINITY;
    }
    for (int c = 0; c < 8; ++c) {
      const double p[3] = {(c & 1) ? b.hi[0] : b.lo[0], (c & 2) ? b.hi[1] : b.lo[1], (c & 4) ? b.hi[2] : b.lo[2]};
      for (int a = 0; a < 3; ++a) {
        const float* m = in.transform + 4 * a;
        const double w = (double)m[0] * p[0] + (double)m[1] * p[1] + (double)m[2] * p[2] + (double)m[3];
        bx[a] = std::min(bx[a], w);
        bx[3 + a] = std::max(bx[3 + a], w);
      }
    }
    for (int a = 0; a < 6; ++a) R = std::max(R, std::fabs(bx[a]));
  }
  // Padding of each instance box: a hit found in object space lies on the
  // world ray only to within about cond(A) * 2^-23 of the ray's reach, so each
  // box grows by 2^-16 cond(A) R (R: the largest world coordinate of any box;
  // rays are assumed to start within ~100x of it).
  std::vector<float> pseudo(9 * (size_t)n);
  for (uint32_t k = 0; k < n; ++k) {
    const double* bx = &box[6 * (size_t)k];
    const double pad = std::ldexp(1.0, -16) * std::max(1.0, cond[k]) * std::max(R, 1e-30);
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::nextafter((float)(bx[a] - pad), -INFINITY);
      hi[a] = std::nextafter((float)(bx[3 + a] + pad), INFINITY);
      APT_REQUIRE(std::isfinite(lo[a]) && std::isfinite(hi[a]), APT_ERR_INVALID_ARG,
                  "instance %u: world box is not finite", k);
    }
    // a box as the degenerate triangle (lo, hi, lo): its bounds are the box
    float* p = &pseudo[9 * (size_t)k];
    for (int a = 0; a < 3; ++a) {
      p[a] = lo[a];
      p[3 + a] = hi[a];
      p[6 + a] = lo[a];
    }
  }
  std::vector<InstRec> ordered(n);
  std::vector<uint8_t> nodes;
  uint32_t root = REF_EMPTY;
  apt_bvh_stats ts{};
  if (n > 0) {
    BvhHostResult b2;
    apt_status st = build_bvh2_host(pseudo.data(), n, 1, 32, b2);
    if (st != APT_OK) return st;
    st = collapse_wide_host(b2, t->width, 0.0f, nodes, root, ts);
    if (st != APT_OK) return st;
    APT_REQUIRE(ts.n_nodes <= TLAS_NODE_CAP, APT_ERR_DEVICE, "TLAS of %u nodes exceeds its %u slots", ts.n_nodes,
                TLAS_NODE_CAP);
    for (uint32_t pos = 0; pos < n; ++pos) ordered[pos] = rec[b2.order[pos]];
  }
  const uint32_t max_stack = (n ? ts.max_stack : 0u) + blas_stack;
  APT_REQUIRE(max_stack <= trace_stack_capacity(), APT_ERR_UNSUPPORTED,
              "two-level scene: the TLAS and its deepest BLAS need a %u-entry traversal stack; the kernels hold %u",
              max_stack, trace_stack_capacity());
  hipStream_t s = lib_stream();
  if (!nodes.empty()) APT_HIP(hipMemcpyAsync(t->d_nodes, nodes.data(), nodes.size(), hipMemcpyHostToDevice, s));
  if (n) APT_HIP(hipMemcpyAsync(t->d_inst, ordered.data(), (size_t)n * sizeof(InstRec), hipMemcpyHostToDevice, s));
  APT_HIP(hipStreamSynchronize(s));
  t->root = root;
  t->n_tlas_nodes = n ? ts.n_nodes : 0u;
  t->n_inst = n;
  t->max_stack = max_stack;
  t->h_inst.swap(ordered);
  t->ms_update = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return APT_OK;
}

apt_status apt_tlas_get_stats(const apt_tlas* t, apt_tlas_stats* out) {
  clear_error();
  APT_REQUIRE(t && out, APT_ERR_INVALID_ARG, "apt_tlas_get_stats: null argument");
  std::memset(out, 0, sizeof(*out));
  out->n_blas = (uint32_t)t->blas.size();
  out->n_instances = t->n_inst;
  out->n_tlas_nodes = t->n_tlas_nodes;
  out->width = t->width;
  out->max_stack = t->max_stack;
  out->blas_tris = t->tri_used;
  out->device_bytes = t->node_cap * wide_node_bytes(t->width) + t->tri_cap * 3 * sizeof(float4) +
                      (size_t)TLAS_MAX_INSTANCES * sizeof(InstRec);
  out->ms_last_update = t->ms_update;
  return APT_OK;
}

}  // extern "C"
