// Device binned-SAH BVH build for gfx950 (replaces the driver BLAS build,
// src/scene_resources/scene_model.rs:53-88, and tinybvh's CPU build).
//
// The build is level-synchronous: every node of a tree level is processed by
// one 64-lane wavefront, which
//   1. bins the node's triangles by centroid on all three axes in LDS
//      (min/max of boxes with integer atomics on order-preserving keys, counts
//      with integer adds: order-independent, so the bins are exact),
//   2. evaluates the SAH cost of every bin boundary in the same order and with
//      the same float operations as the host builder (lane 0),
//   3. partitions the node's index range stably with wave ballots (left
//      triangles keep their order, then right ones), reducing both children's
//      triangle and centroid bounds on the way.
// A single-block scan then numbers the children of the splitting nodes and
// forms the next level.  Every choice depends only on the set of triangles in
// a node and their input order, so the binary tree, its leaves and their
// triangle order equal the host builder's (which uses a stable partition too);
// the breadth-first collapse to 2/4/8-wide nodes is shared with it.
#include <chrono>
#include <vector>

#include "apt_internal.h"

namespace apt {
namespace {

constexpr int WAVE = 64;
constexpr int MAX_BINS = 64;
constexpr int MAX_DEPTH = 56;  // the host builder's limit: median splits below it
constexpr int SCAN_BLOCK = 1024;

__device__ __forceinline__ int okey(float f) {
  const int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ofloat(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF); }
// std::min / std::max semantics of the host builder
__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return a < b ? b : a; }

struct GBox {
  float lo[3], hi[3];
};

__device__ __forceinline__ float box_area(const float* lo, const float* hi) {
  const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Per-triangle bounds and centroid (9 floats), and a flag for non-finite input.
__global__ void k_prims(const float* __restrict__ tri, uint32_t n, float* __restrict__ prim,
                        uint32_t* __restrict__ bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* t = tri + (size_t)i * 9;
  float* p = prim + (size_t)i * 9;
  bool ok = true;
  for (int k = 0; k < 3; ++k) {
    const float a = t[k], b = t[3 + k], c = t[6 + k];
    ok = ok && isfinite(a) && isfinite(b) && isfinite(c);
    const float lo = smin(a, smin(b, c)), hi = smax(a, smax(b, c));
    p[k] = lo;
    p[3 + k] = hi;
    p[6 + k] = (lo + hi) * 0.5f;
  }
  if (!ok) atomicMin(bad, i);
}

// Root triangle and centroid bounds: keys[0..5] box lo/hi, keys[6..11] centroid lo/hi.
__global__ void k_root_bounds(const float* __restrict__ prim, uint32_t n, int* __restrict__ keys) {
  float v[12];
  for (int k = 0; k < 3; ++k) {
    v[k] = v[6 + k] = INFINITY;
    v[3 + k] = v[9 + k] = -INFINITY;
  }
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float* p = prim + (size_t)i * 9;
    for (int k = 0; k < 3; ++k) {
      v[k] = smin(v[k], p[k]);
      v[3 + k] = smax(v[3 + k], p[3 + k]);
      v[6 + k] = smin(v[6 + k], p[6 + k]);
      v[9 + k] = smax(v[9 + k], p[6 + k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    atomicMin(&keys[k], okey(v[k]));
    atomicMax(&keys[3 + k], okey(v[3 + k]));
    atomicMin(&keys[6 + k], okey(v[6 + k]));
    atomicMax(&keys[9 + k], okey(v[9 + k]));
  }
}

__global__ void k_root_init(const int* __restrict__ keys, uint32_t n, uint32_t* act_node, uint32_t* act_range,
                            float* act_box) {
  act_node[0] = 0;
  act_range[0] = 0;
  act_range[1] = n;
  for (int k = 0; k < 12; ++k) act_box[k] = ofloat(keys[k]);
}

struct Level {
  const uint32_t* act_node;   // [a]
  const uint32_t* act_range;  // [2a], [2a+1]: begin, end
  const float* act_box;       // [12a]: box lo/hi, centroid lo/hi
  uint32_t n_active;
  uint32_t depth;
  uint32_t max_leaf, bins;
  const float* prim;
  const uint32_t* order_in;
  uint32_t* order_out;
  Node2* nodes;
  uint32_t* split;      // [a]: 1 when the node was split
  uint32_t* mid;        // [a]
  float* child_box;     // [24a]: left box+centroid box, right box+centroid box
};

__device__ __forceinline__ int bin_of(float c, float lo, float scale, int bins) {
  int b = (int)((c - lo) * scale);
  return b < 0 ? 0 : (b > bins - 1 ? bins - 1 : b);
}

__device__ __forceinline__ float wave_min(float v) {
  for (int m = 32; m >= 1; m >>= 1) v = smin(v, __shfl_xor(v, m));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
  for (int m = 32; m >= 1; m >>= 1) v = smax(v, __shfl_xor(v, m));
  return v;
}

__global__ __launch_bounds__(WAVE) void k_level(Level L) {
  __shared__ int s_lo[3][MAX_BINS][3];
  __shared__ int s_hi[3][MAX_BINS][3];
  __shared__ uint32_t s_cnt[3][MAX_BINS];
  __shared__ int s_axis, s_split;
  __shared__ uint32_t s_nleft;
  const uint32_t a = blockIdx.x;
  const uint32_t lane = threadIdx.x;
  const uint32_t begin = L.act_range[2 * a], end = L.act_range[2 * a + 1];
  const uint32_t n = end - begin;
  const float* box = L.act_box + 12 * (size_t)a;
  const uint32_t node = L.act_node[a];
  Node2& nd = L.nodes[node];
  if (lane == 0) {
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] = box[k];
      nd.hi[k] = box[3 + k];
    }
  }
  if (n <= L.max_leaf) {
    for (uint32_t i = begin + lane; i < end; i += WAVE) L.order_out[i] = L.order_in[i];
    if (lane == 0) {
      nd.first = begin;
      nd.count = n;
      nd.left = nd.right = 0;
      L.split[a] = 0;
    }
    return;
  }
  const int bins = (int)L.bins;
  float cblo[3], scale[3];
  bool use[3];
  for (int k = 0; k < 3; ++k) {
    cblo[k] = box[6 + k];
    const float ext = box[9 + k] - box[6 + k];
    use[k] = ext > 0.0f && L.depth < (uint32_t)MAX_DEPTH;
    scale[k] = use[k] ? (float)bins / ext : 0.0f;
  }
  if (lane == 0) s_axis = -1;
  if (use[0] || use[1] || use[2]) {
    for (int j = lane; j < 3 * MAX_BINS; j += WAVE) {
      const int ax = j / MAX_BINS, b = j % MAX_BINS;
      for (int k = 0; k < 3; ++k) {
        s_lo[ax][b][k] = okey(INFINITY);
        s_hi[ax][b][k] = okey(-INFINITY);
      }
      s_cnt[ax][b] = 0;
    }
    __syncthreads();
    for (uint32_t i = begin + lane; i < end; i += WAVE) {
      const float* p = L.prim + (size_t)L.order_in[i] * 9;
      const int kl[3] = {okey(p[0]), okey(p[1]), okey(p[2])};
      const int kh[3] = {okey(p[3]), okey(p[4]), okey(p[5])};
      for (int ax = 0; ax < 3; ++ax) {
        if (!use[ax]) continue;
        const int b = bin_of(p[6 + ax], cblo[ax], scale[ax], bins);
        for (int k = 0; k < 3; ++k) {
          atomicMin(&s_lo[ax][b][k], kl[k]);
          atomicMax(&s_hi[ax][b][k], kh[k]);
        }
        atomicAdd(&s_cnt[ax][b], 1u);
      }
    }
    __syncthreads();
    if (lane == 0) {
      // SAH sweep of the host builder: right-to-left areas and counts, then
      // left-to-right costs, first strictly smaller cost wins
      float best_cost = INFINITY;
      int best_axis = -1, best_split = 0;
      uint32_t best_left = 0;
      float right_area[MAX_BINS];
      uint32_t right_cnt[MAX_BINS];
      for (int ax = 0; ax < 3; ++ax) {
        if (!use[ax]) continue;
        GBox acc;
        for (int k = 0; k < 3; ++k) {
          acc.lo[k] = INFINITY;
          acc.hi[k] = -INFINITY;
        }
        uint32_t cnt = 0;
        for (int b = bins - 1; b > 0; --b) {
          for (int k = 0; k < 3; ++k) {
            acc.lo[k] = smin(acc.lo[k], ofloat(s_lo[ax][b][k]));
            acc.hi[k] = smax(acc.hi[k], ofloat(s_hi[ax][b][k]));
          }
          cnt += s_cnt[ax][b];
          right_area[b] = box_area(acc.lo, acc.hi);
          right_cnt[b] = cnt;
        }
        for (int k = 0; k < 3; ++k) {
          acc.lo[k] = INFINITY;
          acc.hi[k] = -INFINITY;
        }
        cnt = 0;
        for (int b = 1; b < bins; ++b) {
          for (int k = 0; k < 3; ++k) {
            acc.lo[k] = smin(acc.lo[k], ofloat(s_lo[ax][b - 1][k]));
            acc.hi[k] = smax(acc.hi[k], ofloat(s_hi[ax][b - 1][k]));
          }
          cnt += s_cnt[ax][b - 1];
          if (cnt == 0 || right_cnt[b] == 0) continue;
          const float cost = box_area(acc.lo, acc.hi) * (float)cnt + right_area[b] * (float)right_cnt[b];
          if (cost < best_cost) {
            best_cost = cost;
            best_axis = ax;
            best_split = b;
            best_left = cnt;
          }
        }
      }
      s_axis = best_axis;
      s_split = best_split;
      s_nleft = best_left;
    }
  }
  __syncthreads();
  const int axis = s_axis, split = s_split;
  // object median on the (sorted) index range when no split applies
  const uint32_t n_left = axis >= 0 ? s_nleft : n / 2;
  float lb[12], rb[12];
  for (int k = 0; k < 3; ++k) {
    lb[k] = lb[6 + k] = rb[k] = rb[6 + k] = INFINITY;
    lb[3 + k] = lb[9 + k] = rb[3 + k] = rb[9 + k] = -INFINITY;
  }
  const float alo = axis >= 0 ? cblo[axis] : 0.0f, asc = axis >= 0 ? scale[axis] : 0.0f;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t lbase = 0, rbase = 0;
  for (uint32_t i0 = begin; i0 < end; i0 += WAVE) {
    const uint32_t i = i0 + lane;
    const bool valid = i < end;
    const uint32_t idx = valid ? L.order_in[i] : 0u;
    const float* p = L.prim + (size_t)idx * 9;
    bool left = false;
    if (valid) left = axis >= 0 ? bin_of(p[6 + axis], alo, asc, bins) < split : (i - begin) < n_left;
    const unsigned long long ml = __ballot(valid && left), mr = __ballot(valid && !left);
    if (valid) {
      float* bx = left ? lb : rb;
      for (int k = 0; k < 3; ++k) {
        bx[k] = smin(bx[k], p[k]);
        bx[3 + k] = smax(bx[3 + k], p[3 + k]);
        bx[6 + k] = smin(bx[6 + k], p[6 + k]);
        bx[9 + k] = smax(bx[9 + k], p[6 + k]);
      }
      const uint32_t dst = left ? begin + lbase + (uint32_t)__popcll(ml & below)
                                : begin + n_left + rbase + (uint32_t)__popcll(mr & below);
      if (dst < end) L.order_out[dst] = idx;  // always true: the bins counted the same test
    }
    lbase += (uint32_t)__popcll(ml);
    rbase += (uint32_t)__popcll(mr);
  }
  float* cb = L.child_box + 24 * (size_t)a;
  for (int k = 0; k < 12; ++k) {
    const bool is_min = (k % 6) < 3;
    const float l = is_min ? wave_min(lb[k]) : wave_max(lb[k]);
    const float r = is_min ? wave_min(rb[k]) : wave_max(rb[k]);
    if (lane == 0) {
      cb[k] = l;
      cb[12 + k] = r;
    }
  }
  if (lane == 0) {
    nd.first = 0;
    nd.count = 0;
    L.split[a] = 1;
    L.mid[a] = begin + n_left;
  }
}

// Numbers the children of the split nodes in active order and writes the
// next level's active list; counter[0] receives its size.
__global__ __launch_bounds__(SCAN_BLOCK) void k_next_level(Level L, uint32_t node_base, uint32_t* next_node,
                                                          uint32_t* next_range, float* next_box,
                                                          uint32_t* counter) {
  __shared__ uint32_t s_part[SCAN_BLOCK];
  const uint32_t t = threadIdx.x;
  const uint32_t per = (L.n_active + SCAN_BLOCK - 1) / SCAN_BLOCK;
  const uint32_t a0 = t * per, a1 = min(L.n_active, a0 + per);
  uint32_t c = 0;
  for (uint32_t a = a0; a < a1; ++a) c += L.split[a];
  s_part[t] = c;
  __syncthreads();
  for (uint32_t off = 1; off < SCAN_BLOCK; off <<= 1) {
    const uint32_t v = t >= off ? s_part[t - off] : 0u;
    __syncthreads();
    s_part[t] += v;
    __syncthreads();
  }
  uint32_t r = s_part[t] - c;  // exclusive prefix
  for (uint32_t a = a0; a < a1; ++a) {
    if (!L.split[a]) continue;
    const uint32_t l = node_base + 2 * r;
    Node2& nd = L.nodes[L.act_node[a]];
    nd.left = l;
    nd.right = l + 1;
    const uint32_t b = L.act_range[2 * a], e = L.act_range[2 * a + 1], m = L.mid[a];
    next_node[2 * r] = l;
    next_node[2 * r + 1] = l + 1;
    next_range[4 * r + 0] = b;
    next_range[4 * r + 1] = m;
    next_range[4 * r + 2] = m;
    next_range[4 * r + 3] = e;
    const float* cb = L.child_box + 24 * (size_t)a;
    for (int k = 0; k < 24; ++k) next_box[24 * (size_t)r + k] = cb[k];
    ++r;
  }
  if (t == SCAN_BLOCK - 1) counter[0] = 2 * s_part[t];
}

template <typename T>
struct DBuf {
  T* p = nullptr;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t n) { return hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)); }
};

}  // namespace

// Builds the binary tree on the device from device-resident triangles and
// returns it (nodes, triangle order, SAH cost) like build_bvh2_host.
apt_status build_bvh2_device(const float* d_tri, uint32_t n, uint32_t max_leaf, uint32_t bins,
                             hipStream_t s, BvhHostResult& out) {
  APT_REQUIRE(bins <= (uint32_t)MAX_BINS, APT_ERR_INVALID_ARG, "bvh: bins %u above %d", bins, MAX_BINS);
  DBuf<float> prim, box[2], child_box;
  DBuf<uint32_t> order[2], node_ids[2], range[2], split, mid, counter, bad;
  DBuf<Node2> nodes;
  DBuf<int> keys;
  const size_t cap_nodes = 2 * (size_t)n + 1;
  APT_HIP(prim.alloc((size_t)n * 9));
  for (int k = 0; k < 2; ++k) {
    APT_HIP(order[k].alloc(n));
    APT_HIP(node_ids[k].alloc(n));
    APT_HIP(range[k].alloc(2 * (size_t)n));
    APT_HIP(box[k].alloc(12 * (size_t)n));
  }
  APT_HIP(split.alloc(n));
  APT_HIP(mid.alloc(n));
  APT_HIP(child_box.alloc(24 * (size_t)n));
  APT_HIP(counter.alloc(1));
  APT_HIP(bad.alloc(1));
  APT_HIP(nodes.alloc(cap_nodes));
  APT_HIP(keys.alloc(12));

  const uint32_t nb = (n + 255) / 256;
  const uint32_t bad_init = 0xFFFFFFFFu;
  APT_HIP(hipMemcpyAsync(bad.p, &bad_init, 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_prims, dim3(nb), dim3(256), 0, s, d_tri, n, prim.p, bad.p);
  std::vector<uint32_t> iota(n);
  for (uint32_t i = 0; i < n; ++i) iota[i] = i;
  APT_HIP(hipMemcpyAsync(order[0].p, iota.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
  int kinit[12];
  for (int k = 0; k < 12; ++k) kinit[k] = ((k % 6) < 3) ? 0x7FFFFFFF : (int)0x80000000;
  APT_HIP(hipMemcpyAsync(keys.p, kinit, sizeof(kinit), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_root_bounds, dim3(std::min(nb, 1024u)), dim3(256), 0, s, prim.p, n, keys.p);
  hipLaunchKernelGGL(k_root_init, dim3(1), dim3(1), 0, s, keys.p, n, node_ids[0].p, range[0].p, box[0].p);
  uint32_t bad_h = 0;
  APT_HIP(hipMemcpyAsync(&bad_h, bad.p, 4, hipMemcpyDeviceToHost, s));
  APT_HIP(hipStreamSynchronize(s));
  APT_REQUIRE(bad_h == 0xFFFFFFFFu, APT_ERR_INVALID_ARG, "triangle %u has a non-finite coordinate", bad_h);

  uint32_t n_active = 1, node_base = 1, depth = 0;
  int cur = 0, ord = 0;
  while (n_active > 0) {
    Level L;
    L.act_node = node_ids[cur].p;
    L.act_range = range[cur].p;
    L.act_box = box[cur].p;
    L.n_active = n_active;
    L.depth = depth;
    L.max_leaf = max_leaf;
    L.bins = bins;
    L.prim = prim.p;
    L.order_in = order[ord].p;
    L.order_out = order[ord ^ 1].p;
    L.nodes = nodes.p;
    L.split = split.p;
    L.mid = mid.p;
    L.child_box = child_box.p;
    hipLaunchKernelGGL(k_level, dim3(n_active), dim3(WAVE), 0, s, L);
    hipLaunchKernelGGL(k_next_level, dim3(1), dim3(SCAN_BLOCK), 0, s, L, node_base, node_ids[cur ^ 1].p,
                       range[cur ^ 1].p, box[cur ^ 1].p, counter.p);
    uint32_t next = 0;
    APT_HIP(hipMemcpyAsync(&next, counter.p, 4, hipMemcpyDeviceToHost, s));
    APT_HIP(hipStreamSynchronize(s));
    APT_REQUIRE(node_base + next <= cap_nodes, APT_ERR_DEVICE, "bvh: device build ran out of nodes");
    node_base += next;
    n_active = next;
    cur ^= 1;
    ord ^= 1;
    depth++;
  }
  APT_HIP(hipGetLastError());
  out.nodes.resize(node_base);
  out.order.resize(n);
  APT_HIP(hipMemcpyAsync(out.nodes.data(), nodes.p, (size_t)node_base * sizeof(Node2), hipMemcpyDeviceToHost, s));
  APT_HIP(hipMemcpyAsync(out.order.data(), order[ord].p, (size_t)n * 4, hipMemcpyDeviceToHost, s));
  APT_HIP(hipStreamSynchronize(s));
  out.sah_cost = bvh2_sah_cost(out.nodes);
  return APT_OK;
}

}  // namespace apt
