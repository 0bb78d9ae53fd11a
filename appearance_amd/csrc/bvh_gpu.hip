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
// Nodes of more than LARGE_NODE triangles (the top levels) run the same three
// steps over many blocks instead of one wave.  A single-block scan then
// numbers the children of the splitting nodes and forms the next level.  Every choice depends only on the set of triangles in
// a node and their input order, so the binary tree, its leaves and their
// triangle order equal the host builder's (which uses a stable partition too).
// The breadth-first collapse to 2/4/8-wide quantized nodes then runs on the
// device as well (collapse_device), byte-identical to the host collapse.
#include <algorithm>
#include <chrono>
#include <cmath>
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
                        uint32_t* __restrict__ order, uint32_t* __restrict__ bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  order[i] = i;  // the root's triangles in input order
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

// Root triangle and centroid bounds as order-preserving keys ([0..5] box
// lo/hi, [6..11] centroid lo/hi): every work-group reduces its share to one
// set of 12 keys in `part` (integer min / max: exact), and k_root_init
// reduces those.  (Atomics on 12 global words from every wave would queue
// thousands of adds on each word.)
constexpr int ROOT_BLOCK = 256;
constexpr uint32_t ROOT_GRID = 1024;

__device__ __forceinline__ void reduce_keys_wave(int (&key)[12]) {
  for (int m = WAVE / 2; m >= 1; m >>= 1) {
    for (int k = 0; k < 12; ++k) {
      const int o = __shfl_xor(key[k], m);
      key[k] = (k % 6) < 3 ? min(key[k], o) : max(key[k], o);
    }
  }
}

__global__ __launch_bounds__(ROOT_BLOCK) void k_root_bounds(const float* __restrict__ prim, uint32_t n,
                                                             int* __restrict__ part) {
  __shared__ int s_key[ROOT_BLOCK / WAVE][12];
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
  int key[12];
  for (int k = 0; k < 12; ++k) key[k] = okey(v[k]);
  reduce_keys_wave(key);
  if ((threadIdx.x & (WAVE - 1)) == 0)
    for (int k = 0; k < 12; ++k) s_key[threadIdx.x / WAVE][k] = key[k];
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = (int)threadIdx.x;
    int r = s_key[0][k];
    for (int w = 1; w < ROOT_BLOCK / WAVE; ++w) r = (k % 6) < 3 ? min(r, s_key[w][k]) : max(r, s_key[w][k]);
    part[(size_t)blockIdx.x * 12 + k] = r;
  }
}

// One wave: reduces the work-groups' keys and sets up the root's range and box.
__global__ __launch_bounds__(WAVE) void k_root_init(const int* __restrict__ part, uint32_t n_part, uint32_t n,
                                                    uint32_t* act_node, uint32_t* act_range, float* act_box) {
  int key[12];
  for (int k = 0; k < 12; ++k) key[k] = (k % 6) < 3 ? 0x7FFFFFFF : (int)0x80000000;
  for (uint32_t j = threadIdx.x; j < n_part; j += WAVE)
    for (int k = 0; k < 12; ++k) key[k] = (k % 6) < 3 ? min(key[k], part[j * 12 + k]) : max(key[k], part[j * 12 + k]);
  reduce_keys_wave(key);
  if (threadIdx.x == 0) {
    act_node[0] = 0;
    act_range[0] = 0;
    act_range[1] = n;
    for (int k = 0; k < 12; ++k) act_box[k] = ofloat(key[k]);
  }
}

struct Level {
  const uint32_t* act_node;   // [a]
  const uint32_t* act_range;  // [2a], [2a+1]: begin, end
  const float* act_box;       // [12a]: box lo/hi, centroid lo/hi
  uint32_t n_active;
  uint32_t depth;
  uint32_t max_leaf, bins;
  uint32_t large;             // nodes above this many triangles use the multi-block path
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

// The host builder's SAH sweep over three axes of bins (lo/hi box keys and
// counts) with one lane per bin: for every boundary b the cost is
// area(bins < b) * count(bins < b) + area(bins >= b) * count(bins >= b), and
// the first strictly smallest cost in (axis, boundary) order wins.  The
// prefix and suffix unions and counts come from wave scans (min, max and
// integer adds are exact, so every boundary sees the boxes and counts of the
// host's sequential sweep), each cost from the same float expression, and the
// winner from a wave reduction.  Every lane of the wave calls it and gets the
// result.
__device__ void sah_sweep_wave(const int (*blo)[MAX_BINS][3], const int (*bhi)[MAX_BINS][3],
                               const uint32_t (*bcnt)[MAX_BINS], const bool* use, int bins, uint32_t lane,
                               int& axis_out, int& split_out, uint32_t& nleft_out) {
  float best_cost = INFINITY;
  int best_key = 0x7FFFFFFF;  // axis * MAX_BINS + boundary
  uint32_t best_left = 0;
  const bool in = (int)lane < bins;
  for (int ax = 0; ax < 3; ++ax) {
    if (!use[ax]) continue;
    float plo[3], phi[3], slo[3], shi[3];
    uint32_t pc = in ? bcnt[ax][lane] : 0u, sc = pc;
    for (int k = 0; k < 3; ++k) {
      plo[k] = slo[k] = in ? ofloat(blo[ax][lane][k]) : INFINITY;
      phi[k] = shi[k] = in ? ofloat(bhi[ax][lane][k]) : -INFINITY;
    }
    // inclusive prefix (bins 0..lane) and suffix (bins lane..) scans
    for (uint32_t d = 1; d < (uint32_t)WAVE; d <<= 1) {
      const bool up = lane >= d, down = lane + d < (uint32_t)WAVE;
      for (int k = 0; k < 3; ++k) {
        const float ul = __shfl_up(plo[k], d), uh = __shfl_up(phi[k], d);
        const float dl = __shfl_down(slo[k], d), dh = __shfl_down(shi[k], d);
        if (up) {
          plo[k] = smin(ul, plo[k]);
          phi[k] = smax(uh, phi[k]);
        }
        if (down) {
          slo[k] = smin(slo[k], dl);
          shi[k] = smax(shi[k], dh);
        }
      }
      const uint32_t uc = __shfl_up(pc, d), dc = __shfl_down(sc, d);
      if (up) pc += uc;
      if (down) sc += dc;
    }
    // boundary b = lane: left = bins 0..b-1 (the prefix one lane down), right = bins b..
    float llo[3], lhi[3];
    for (int k = 0; k < 3; ++k) {
      llo[k] = __shfl_up(plo[k], 1);
      lhi[k] = __shfl_up(phi[k], 1);
    }
    const uint32_t lc = __shfl_up(pc, 1);
    if (lane >= 1 && in && lc != 0 && sc != 0) {
      const float cost = box_area(llo, lhi) * (float)lc + box_area(slo, shi) * (float)sc;
      if (cost < best_cost) {  // a lane's earlier axis wins a tie
        best_cost = cost;
        best_key = ax * MAX_BINS + (int)lane;
        best_left = lc;
      }
    }
  }
  for (int m = WAVE / 2; m >= 1; m >>= 1) {
    const float oc = __shfl_xor(best_cost, m);
    const int ok = __shfl_xor(best_key, m);
    const uint32_t ol = __shfl_xor(best_left, m);
    if (oc < best_cost || (oc == best_cost && ok < best_key)) {
      best_cost = oc;
      best_key = ok;
      best_left = ol;
    }
  }
  const bool found = best_cost < INFINITY;
  axis_out = found ? best_key / MAX_BINS : -1;
  split_out = found ? best_key % MAX_BINS : 0;
  nleft_out = found ? best_left : 0u;
}

// Binning set-up of a node from its centroid box: the axes that can split
// (non-flat, above the depth limit) and their bin scale.
__device__ __forceinline__ void bin_setup(const float* box, uint32_t depth, int bins, float* cblo, float* scale,
                                          bool* use) {
  for (int k = 0; k < 3; ++k) {
    cblo[k] = box[6 + k];
    const float ext = box[9 + k] - box[6 + k];
    use[k] = ext > 0.0f && depth < (uint32_t)MAX_DEPTH;
    scale[k] = use[k] ? (float)bins / ext : 0.0f;
  }
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
  if (n > L.large) return;  // split by the multi-block kernels below
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
  bin_setup(box, L.depth, bins, cblo, scale, use);
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
    int ax_w, split_w;
    uint32_t nleft_w;
    sah_sweep_wave(s_lo, s_hi, s_cnt, use, bins, lane, ax_w, split_w, nleft_w);
    if (lane == 0) {
      s_axis = ax_w;
      s_split = split_w;
      s_nleft = nleft_w;
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

// ---------------------------------------------------------------------------
// Multi-block path for large nodes (the top levels, where a level has few
// nodes of many triangles and one wave per node would leave the GPU idle).
// The same three steps, spread over blocks of LARGE_CHUNK triangles: the bins
// and child boxes are merged across blocks with integer atomics on
// order-preserving keys (exact and order-independent), the sweep is the same
// function on the merged bins, and the stable partition places each block's
// left and right triangles after those of the blocks before it.
// ---------------------------------------------------------------------------
constexpr uint32_t LARGE_NODE = 16384;  // triangles above which a node takes this path
constexpr uint32_t LARGE_CHUNK = 4096;  // triangles per block
constexpr int LBLOCK = 256;
constexpr int BIN_WORDS = 3 * MAX_BINS * 3 * 2 + 3 * MAX_BINS;  // lo keys, hi keys, counts

struct Large {
  uint32_t n;                  // large nodes of this level
  const uint32_t* act;         // [n] active index of each
  const uint32_t* blk_node;    // [blocks] large node of each block
  const uint32_t* blk_first;   // [n] first block of each large node
  int* bins;                   // [n][BIN_WORDS]
  int* info;                   // [n][4]: axis, split, n_left
  uint32_t* blk_left;          // [blocks]: left triangles per block, then their exclusive prefix
  int* cbox;                   // [n][24] child box keys
  uint32_t blocks;
};

__device__ __forceinline__ int* bins_lo(int* b) { return b; }
__device__ __forceinline__ int* bins_hi(int* b) { return b + 3 * MAX_BINS * 3; }
__device__ __forceinline__ uint32_t* bins_cnt(int* b) { return reinterpret_cast<uint32_t*>(b + 2 * 3 * MAX_BINS * 3); }

__global__ __launch_bounds__(LBLOCK) void k_large_init(Large G) {
  int* b = G.bins + (size_t)blockIdx.x * BIN_WORDS;
  for (int i = threadIdx.x; i < 3 * MAX_BINS * 3; i += LBLOCK) {
    bins_lo(b)[i] = okey(INFINITY);
    bins_hi(b)[i] = okey(-INFINITY);
  }
  for (int i = threadIdx.x; i < 3 * MAX_BINS; i += LBLOCK) bins_cnt(b)[i] = 0u;
  if (threadIdx.x < 24) {
    const int k = threadIdx.x % 12;
    G.cbox[24 * blockIdx.x + threadIdx.x] = (k % 6) < 3 ? okey(INFINITY) : okey(-INFINITY);
  }
}

// Triangle range of block `blk` of large node j.
__device__ __forceinline__ void large_chunk(const Level& L, const Large& G, uint32_t blk, uint32_t& j,
                                            uint32_t& a, uint32_t& c0, uint32_t& c1) {
  j = G.blk_node[blk];
  a = G.act[j];
  const uint32_t begin = L.act_range[2 * a], end = L.act_range[2 * a + 1];
  c0 = begin + (blk - G.blk_first[j]) * LARGE_CHUNK;
  c1 = min(end, c0 + LARGE_CHUNK);
}

__global__ __launch_bounds__(LBLOCK) void k_large_bin(Level L, Large G) {
  __shared__ int s_lo[3][MAX_BINS][3];
  __shared__ int s_hi[3][MAX_BINS][3];
  __shared__ uint32_t s_cnt[3][MAX_BINS];
  uint32_t j, a, c0, c1;
  large_chunk(L, G, blockIdx.x, j, a, c0, c1);
  const int bins = (int)L.bins;
  float cblo[3], scale[3];
  bool use[3];
  bin_setup(L.act_box + 12 * (size_t)a, L.depth, bins, cblo, scale, use);
  if (!(use[0] || use[1] || use[2])) return;
  for (int i = threadIdx.x; i < 3 * MAX_BINS; i += LBLOCK) {
    const int ax = i / MAX_BINS, b = i % MAX_BINS;
    for (int k = 0; k < 3; ++k) {
      s_lo[ax][b][k] = okey(INFINITY);
      s_hi[ax][b][k] = okey(-INFINITY);
    }
    s_cnt[ax][b] = 0;
  }
  __syncthreads();
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += LBLOCK) {
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
  int* gb = G.bins + (size_t)j * BIN_WORDS;
  for (int i = threadIdx.x; i < 3 * MAX_BINS; i += LBLOCK) {
    const int ax = i / MAX_BINS, b = i % MAX_BINS;
    if (!s_cnt[ax][b]) continue;
    for (int k = 0; k < 3; ++k) {
      atomicMin(&bins_lo(gb)[(ax * MAX_BINS + b) * 3 + k], s_lo[ax][b][k]);
      atomicMax(&bins_hi(gb)[(ax * MAX_BINS + b) * 3 + k], s_hi[ax][b][k]);
    }
    atomicAdd(&bins_cnt(gb)[ax * MAX_BINS + b], s_cnt[ax][b]);
  }
}

// One wave per large node (the bins are merged in global memory).
__global__ __launch_bounds__(WAVE) void k_large_sweep(Level L, Large G) {
  const uint32_t j = blockIdx.x, a = G.act[j];
  const uint32_t n = L.act_range[2 * a + 1] - L.act_range[2 * a];
  float cblo[3], scale[3];
  bool use[3];
  bin_setup(L.act_box + 12 * (size_t)a, L.depth, (int)L.bins, cblo, scale, use);
  int axis = -1, split = 0;
  uint32_t nleft = 0;
  int* gb = G.bins + (size_t)j * BIN_WORDS;
  if (use[0] || use[1] || use[2])
    sah_sweep_wave(reinterpret_cast<const int (*)[MAX_BINS][3]>(bins_lo(gb)),
                   reinterpret_cast<const int (*)[MAX_BINS][3]>(bins_hi(gb)),
                   reinterpret_cast<const uint32_t (*)[MAX_BINS]>(bins_cnt(gb)), use, (int)L.bins, threadIdx.x, axis,
                   split, nleft);
  if (threadIdx.x == 0) {
    G.info[4 * j + 0] = axis;
    G.info[4 * j + 1] = split;
    G.info[4 * j + 2] = (int)(axis >= 0 ? nleft : n / 2);  // object median when nothing splits
  }
}

// The split test of large node j (axis, split bin, or the object median).
struct LargeSplit {
  int axis, split;
  uint32_t begin, n_left;
  float lo, scale;
};

__device__ __forceinline__ LargeSplit large_split(const Level& L, const Large& G, uint32_t j, uint32_t a) {
  LargeSplit t;
  t.axis = G.info[4 * j];
  t.split = G.info[4 * j + 1];
  t.n_left = (uint32_t)G.info[4 * j + 2];
  t.begin = L.act_range[2 * a];
  float cblo[3], scale[3];
  bool use[3];
  bin_setup(L.act_box + 12 * (size_t)a, L.depth, (int)L.bins, cblo, scale, use);
  t.lo = t.axis >= 0 ? cblo[t.axis] : 0.0f;
  t.scale = t.axis >= 0 ? scale[t.axis] : 0.0f;
  return t;
}

__device__ __forceinline__ bool large_left(const Level& L, const LargeSplit& t, uint32_t i, uint32_t idx) {
  if (t.axis < 0) return i - t.begin < t.n_left;
  return bin_of(L.prim[(size_t)idx * 9 + 6 + t.axis], t.lo, t.scale, (int)L.bins) < t.split;
}

__global__ __launch_bounds__(LBLOCK) void k_large_count(Level L, Large G) {
  __shared__ uint32_t s_sum;
  uint32_t j, a, c0, c1;
  large_chunk(L, G, blockIdx.x, j, a, c0, c1);
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  const LargeSplit t = large_split(L, G, j, a);
  uint32_t c = 0;
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += LBLOCK) c += large_left(L, t, i, L.order_in[i]) ? 1u : 0u;
  atomicAdd(&s_sum, c);
  __syncthreads();
  if (threadIdx.x == 0) G.blk_left[blockIdx.x] = s_sum;
}

// Exclusive prefix of the per-block left counts within each large node.
__global__ void k_large_scan(Large G) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t run = 0, node = 0xFFFFFFFFu;
  for (uint32_t b = 0; b < G.blocks; ++b) {
    if (G.blk_node[b] != node) {
      node = G.blk_node[b];
      run = 0;
    }
    const uint32_t c = G.blk_left[b];
    G.blk_left[b] = run;
    run += c;
  }
}

__global__ __launch_bounds__(LBLOCK) void k_large_scatter(Level L, Large G) {
  __shared__ uint32_t s_wl[LBLOCK / WAVE], s_wr[LBLOCK / WAVE];
  __shared__ float s_box[LBLOCK / WAVE][24];
  uint32_t j, a, c0, c1;
  large_chunk(L, G, blockIdx.x, j, a, c0, c1);
  const LargeSplit t = large_split(L, G, j, a);
  const uint32_t begin = t.begin, n_left = t.n_left, end = L.act_range[2 * a + 1];
  const uint32_t lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
  const unsigned long long below = (1ull << lane) - 1ull;
  uint32_t lbase = G.blk_left[blockIdx.x];
  uint32_t rbase = (c0 - begin) - lbase;  // right triangles of the blocks before this one
  float lb[12], rb[12];
  for (int k = 0; k < 3; ++k) {
    lb[k] = lb[6 + k] = rb[k] = rb[6 + k] = INFINITY;
    lb[3 + k] = lb[9 + k] = rb[3 + k] = rb[9 + k] = -INFINITY;
  }
  for (uint32_t i0 = c0; i0 < c1; i0 += LBLOCK) {
    const uint32_t i = i0 + threadIdx.x;
    const bool valid = i < c1;
    const uint32_t idx = valid ? L.order_in[i] : 0u;
    const bool left = valid && large_left(L, t, i, idx);
    const unsigned long long ml = __ballot(valid && left), mr = __ballot(valid && !left);
    if (lane == 0) {
      s_wl[wv] = (uint32_t)__popcll(ml);
      s_wr[wv] = (uint32_t)__popcll(mr);
    }
    __syncthreads();
    uint32_t wl = 0, wr = 0, tl = 0, tr = 0;
    for (uint32_t w = 0; w < LBLOCK / WAVE; ++w) {
      if (w < wv) {
        wl += s_wl[w];
        wr += s_wr[w];
      }
      tl += s_wl[w];
      tr += s_wr[w];
    }
    if (valid) {
      const float* p = L.prim + (size_t)idx * 9;
      float* bx = left ? lb : rb;
      for (int k = 0; k < 3; ++k) {
        bx[k] = smin(bx[k], p[k]);
        bx[3 + k] = smax(bx[3 + k], p[3 + k]);
        bx[6 + k] = smin(bx[6 + k], p[6 + k]);
        bx[9 + k] = smax(bx[9 + k], p[6 + k]);
      }
      const uint32_t dst = left ? begin + lbase + wl + (uint32_t)__popcll(ml & below)
                                : begin + n_left + rbase + wr + (uint32_t)__popcll(mr & below);
      if (dst < end) L.order_out[dst] = idx;  // always true: the count pass used the same test
    }
    lbase += tl;
    rbase += tr;
    __syncthreads();
  }
  for (int k = 0; k < 12; ++k) {
    const bool is_min = (k % 6) < 3;
    const float l = is_min ? wave_min(lb[k]) : wave_max(lb[k]);
    const float r = is_min ? wave_min(rb[k]) : wave_max(rb[k]);
    if (lane == 0) {
      s_box[wv][k] = l;
      s_box[wv][12 + k] = r;
    }
  }
  __syncthreads();
  if (threadIdx.x < 24) {
    const int k = threadIdx.x;
    const bool is_min = (k % 6) < 3;
    int key = is_min ? okey(INFINITY) : okey(-INFINITY);
    for (uint32_t w = 0; w < LBLOCK / WAVE; ++w) {
      const int v = okey(s_box[w][k]);
      key = is_min ? min(key, v) : max(key, v);
    }
    if (is_min) atomicMin(&G.cbox[24 * j + k], key);
    else atomicMax(&G.cbox[24 * j + k], key);
  }
}

__global__ void k_large_final(Level L, Large G) {
  const uint32_t j = blockIdx.x, a = G.act[j];
  const float* box = L.act_box + 12 * (size_t)a;
  Node2& nd = L.nodes[L.act_node[a]];
  for (int k = 0; k < 3; ++k) {
    nd.lo[k] = box[k];
    nd.hi[k] = box[3 + k];
  }
  nd.first = 0;
  nd.count = 0;
  float* cb = L.child_box + 24 * (size_t)a;
  for (int k = 0; k < 24; ++k) cb[k] = ofloat(G.cbox[24 * j + k]);
  L.split[a] = 1;
  L.mid[a] = L.act_range[2 * a] + (uint32_t)G.info[4 * j + 2];
}

// Numbers the children of the split nodes in active order and writes the
// next level's active list: per-block split counts, their exclusive scan
// (one block; counter[0] receives the next level's size), then each block
// writes its nodes' children from its offset.
constexpr uint32_t NEXT_CHUNK = 1024;  // active nodes per block, at least (a multiple of NEXT_BLOCK)
constexpr int NEXT_BLOCK = 256;
constexpr uint32_t NEXT_MAX_BLOCKS = 4096;  // one scan block of SCAN_BLOCK threads, 4 each

__global__ __launch_bounds__(NEXT_BLOCK) void k_next_count(Level L, uint32_t chunk, uint32_t* blk_sum) {
  __shared__ uint32_t s_sum;
  if (threadIdx.x == 0) s_sum = 0;
  __syncthreads();
  const uint32_t a0 = blockIdx.x * chunk, a1 = min(L.n_active, a0 + chunk);
  uint32_t c = 0;
  for (uint32_t a = a0 + threadIdx.x; a < a1; a += NEXT_BLOCK) c += L.split[a];
  atomicAdd(&s_sum, c);
  __syncthreads();
  if (threadIdx.x == 0) blk_sum[blockIdx.x] = s_sum;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_next_scan(uint32_t* blk_sum, uint32_t nb, uint32_t* counter) {
  __shared__ uint32_t s_part[SCAN_BLOCK];
  const uint32_t t = threadIdx.x;
  constexpr uint32_t PER = NEXT_MAX_BLOCKS / SCAN_BLOCK;
  uint32_t v[PER], c = 0;
  for (uint32_t k = 0; k < PER; ++k) {
    const uint32_t i = t * PER + k;
    v[k] = i < nb ? blk_sum[i] : 0u;
    c += v[k];
  }
  s_part[t] = c;
  __syncthreads();
  for (uint32_t off = 1; off < SCAN_BLOCK; off <<= 1) {
    const uint32_t x = t >= off ? s_part[t - off] : 0u;
    __syncthreads();
    s_part[t] += x;
    __syncthreads();
  }
  uint32_t r = s_part[t] - c;
  for (uint32_t k = 0; k < PER; ++k) {
    const uint32_t i = t * PER + k;
    if (i < nb) blk_sum[i] = r;
    r += v[k];
  }
  if (t == SCAN_BLOCK - 1) counter[0] = 2 * s_part[t];
}

__global__ __launch_bounds__(NEXT_BLOCK) void k_next_write(Level L, uint32_t chunk, const uint32_t* __restrict__ blk_off,
                                                          uint32_t node_base, uint32_t* next_node,
                                                          uint32_t* next_range, float* next_box) {
  const uint32_t PER = chunk / NEXT_BLOCK;  // consecutive active nodes per thread
  __shared__ uint32_t s_part[NEXT_BLOCK];
  const uint32_t t = threadIdx.x;
  const uint32_t a0 = blockIdx.x * chunk + t * PER;
  uint32_t c = 0;
  for (uint32_t k = 0; k < PER; ++k)
    if (a0 + k < L.n_active) c += L.split[a0 + k];
  s_part[t] = c;
  __syncthreads();
  for (uint32_t off = 1; off < NEXT_BLOCK; off <<= 1) {
    const uint32_t x = t >= off ? s_part[t - off] : 0u;
    __syncthreads();
    s_part[t] += x;
    __syncthreads();
  }
  uint32_t r = blk_off[blockIdx.x] + s_part[t] - c;
  for (uint32_t k = 0; k < PER; ++k) {
    const uint32_t a = a0 + k;
    if (a >= L.n_active || !L.split[a]) continue;
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
    for (int k2 = 0; k2 < 24; ++k2) next_box[24 * (size_t)r + k2] = cb[k2];
    ++r;
  }
}

__global__ void k_gather_tris(const float* __restrict__ tri, const uint32_t* __restrict__ order, uint32_t n,
                              float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = order[i];
  const float* t = tri + (size_t)p * 9;
  out[3 * (size_t)i + 0] = make_float4(t[0], t[1], t[2], __uint_as_float(p));
  out[3 * (size_t)i + 1] = make_float4(t[3], t[4], t[5], 0.0f);
  out[3 * (size_t)i + 2] = make_float4(t[6], t[7], t[8], 0.0f);
}

// ---------------------------------------------------------------------------
// Collapse to W-wide quantized nodes on the device: the host collapse
// (bvh_host.cpp) restated level by level.  Each wide node of a level opens
// its BVH2 node greedily (the internal child of largest surface area first,
// the same float operations), a prefix sum over the internal children numbers
// the next level in breadth-first queue order, and a last pass quantizes
// every node's child boxes to its grid in double precision with the host's
// floor/ceil and nudge loops, so the nodes are byte-identical to the host
// collapse's.
constexpr uint32_t COL_BLOCK = 256;

// bvh_host.cpp's box_area (no empty-box test: collapse boxes are valid)
__device__ __forceinline__ float col_area(const float* lo, const float* hi) {
  const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// Opens wide node base + i (BVH2 node front[i]): its up to W BVH2 children
// into wkids, nk | (internal children << 8) into nkni, and the block's count
// of internal children into blk_sum.
template <int W>
__global__ __launch_bounds__(COL_BLOCK) void k_col_open(const Node2* __restrict__ bn,
                                                        const uint32_t* __restrict__ front, uint32_t m,
                                                        uint32_t base, uint32_t* __restrict__ wkids,
                                                        uint32_t* __restrict__ nkni,
                                                        uint32_t* __restrict__ blk_sum) {
  __shared__ uint32_t s_sum[COL_BLOCK / WAVE];
  const uint32_t i = blockIdx.x * COL_BLOCK + threadIdx.x;
  uint32_t ni = 0;
  if (i < m) {
    uint32_t kids[W];
    int nk = 0;
    kids[nk++] = bn[front[i]].left;
    kids[nk++] = bn[front[i]].right;
    while (nk < W) {
      int best = -1;
      float best_a = -1.0f;
      for (int k = 0; k < nk; ++k) {
        const Node2& c = bn[kids[k]];
        if (c.count) continue;
        const float a = col_area(c.lo, c.hi);
        if (a > best_a) {
          best_a = a;
          best = k;
        }
      }
      if (best < 0) break;
      const uint32_t x = kids[best];
      kids[best] = bn[x].left;
      kids[nk++] = bn[x].right;
    }
    for (int k = 0; k < W; ++k) {
      const bool in = k < nk;
      wkids[((size_t)base + i) * W + k] = in ? kids[k] : 0xFFFFFFFFu;
      if (in && !bn[kids[k]].count) ++ni;
    }
    nkni[i] = (uint32_t)nk | (ni << 8);
  }
  uint32_t v = ni;
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  if ((threadIdx.x & (WAVE - 1)) == 0) s_sum[threadIdx.x / WAVE] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t w = 0; w < COL_BLOCK / WAVE; ++w) t += s_sum[w];
    blk_sum[blockIdx.x] = t;
  }
}

// Exclusive scan of blk[0, nb) in place by one work-group; the sum goes to *total.
__global__ __launch_bounds__(SCAN_BLOCK) void k_col_scan(uint32_t* __restrict__ blk, uint32_t nb,
                                                         uint32_t* __restrict__ total) {
  __shared__ uint32_t s[SCAN_BLOCK];
  const uint32_t t = threadIdx.x;
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < nb; c0 += SCAN_BLOCK) {
    const uint32_t i = c0 + t;
    const uint32_t v = i < nb ? blk[i] : 0u;
    s[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < SCAN_BLOCK; off <<= 1) {
      const uint32_t x = t >= off ? s[t - off] : 0u;
      __syncthreads();
      s[t] += x;
      __syncthreads();
    }
    if (i < nb) blk[i] = carry + s[t] - v;
    carry += s[SCAN_BLOCK - 1];
    __syncthreads();
  }
  if (t == 0) *total = carry;
}

// Writes the child refs of wide node base + i (internal children numbered
// base + m + their rank among the level's internal children) and the next
// level's BVH2 nodes and stack depths; stats[0] = max stack, stats[1] =
// leaves (as the host collapse counts them).
template <int W>
__global__ __launch_bounds__(COL_BLOCK) void k_col_write(const Node2* __restrict__ bn,
                                                         const uint32_t* __restrict__ fstack, uint32_t m,
                                                         uint32_t base, const uint32_t* __restrict__ wkids,
                                                         const uint32_t* __restrict__ nkni,
                                                         const uint32_t* __restrict__ blk_off,
                                                         WNode<W>* __restrict__ wide,
                                                         uint32_t* __restrict__ nfront,
                                                         uint32_t* __restrict__ nstack,
                                                         uint32_t* __restrict__ stats) {
  __shared__ uint32_t s[COL_BLOCK];
  const uint32_t t = threadIdx.x, i = blockIdx.x * COL_BLOCK + t;
  const uint32_t code = i < m ? nkni[i] : 0u;
  const uint32_t nk = code & 0xFFu, ni = code >> 8;
  const uint32_t fs = i < m ? fstack[i] : 0u;
  s[t] = ni;
  __syncthreads();
  for (uint32_t off = 1; off < COL_BLOCK; off <<= 1) {
    const uint32_t x = t >= off ? s[t - off] : 0u;
    __syncthreads();
    s[t] += x;
    __syncthreads();
  }
  uint32_t smax_ = i < m ? fs + nk : 0u, sleaf = nk - ni;
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    smax_ = max(smax_, (uint32_t)__shfl_xor(smax_, o, WAVE));
    sleaf += __shfl_xor(sleaf, o, WAVE);
  }
  if ((t & (WAVE - 1)) == 0) {
    atomicMax(&stats[0], smax_);
    if (sleaf) atomicAdd(&stats[1], sleaf);
  }
  if (i >= m) return;
  const uint32_t here = fs + nk - 1u, next_base = base + m;
  uint32_t j = blk_off[blockIdx.x] + s[t] - ni;
  WNode<W>& w = wide[(size_t)base + i];
  for (int k = 0; k < W; ++k) {
    const uint32_t kid = wkids[((size_t)base + i) * W + k];
    uint32_t ref;
    if (kid == 0xFFFFFFFFu) {
      ref = REF_EMPTY;
    } else if (bn[kid].count) {
      ref = make_leaf_ref(bn[kid].first, bn[kid].count);
    } else {
      ref = next_base + j;
      nfront[j] = kid;
      nstack[j] = here;
      ++j;
    }
    w.child[k] = ref;
  }
}

// quantize_node (bvh_host.cpp) for every wide node: the padded child boxes on
// the node's power-of-two grid, rounded outwards and checked in double.
template <int W>
__global__ __launch_bounds__(COL_BLOCK) void k_col_quant(const Node2* __restrict__ bn,
                                                         const uint32_t* __restrict__ wkids, uint32_t n_wide,
                                                         float pad, WNode<W>* __restrict__ wide) {
  const uint32_t i = blockIdx.x * COL_BLOCK + threadIdx.x;
  if (i >= n_wide) return;
  const uint32_t* kids = wkids + (size_t)i * W;
  int nk = 0;
  while (nk < W && kids[nk] != 0xFFFFFFFFu) ++nk;
  float clo[W][3], chi[W][3], nlo[3] = {INFINITY, INFINITY, INFINITY};
  float nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int k = 0; k < nk; ++k) {
    const Node2& c = bn[kids[k]];
    for (int a = 0; a < 3; ++a) {
      clo[k][a] = c.lo[a] - pad;
      chi[k][a] = c.hi[a] + pad;
      nlo[a] = smin(nlo[a], clo[k][a]);
      nhi[a] = smax(nhi[a], chi[k][a]);
    }
  }
  WNode<W>& w = wide[i];
  for (int a = 0; a < 3; ++a) {
    const double lo = nk ? (double)nlo[a] : 0.0, ext = nk ? (double)nhi[a] - lo : 0.0;
    int e = -126;
    if (ext > 0.0) {
      (void)frexp(ext / 255.0, &e);  // 2^(e-1) <= ext/255 < 2^e
      e = e < -126 ? -126 : (e > 127 ? 127 : e);
      while (e > -126 && ext <= 255.0 * ldexp(1.0, e - 1)) --e;
      while (e < 127 && ext > 255.0 * ldexp(1.0, e)) ++e;
    }
    if (a == 0) {
      w.ox = (float)lo;
      w.ex = (uint8_t)(e + 127);
    } else if (a == 1) {
      w.oy = (float)lo;
      w.ey = (uint8_t)(e + 127);
    } else {
      w.oz = (float)lo;
      w.ez = (uint8_t)(e + 127);
    }
    const double step = ldexp(1.0, e);
    for (int k = 0; k < W; ++k) {
      if (k >= nk) {
        w.q[a][k] = 255;
        w.q[a + 3][k] = 0;
        continue;
      }
      double ql = floor(((double)clo[k][a] - lo) / step);
      double qh = ceil(((double)chi[k][a] - lo) / step);
      ql = fmin(255.0, fmax(0.0, ql));
      qh = fmin(255.0, fmax(0.0, qh));
      while (ql > 0.0 && lo + ql * step > (double)clo[k][a]) ql -= 1.0;
      while (qh < 255.0 && lo + qh * step < (double)chi[k][a]) qh += 1.0;
      w.q[a][k] = (uint8_t)ql;
      w.q[a + 3][k] = (uint8_t)qh;
    }
  }
}

// Per-block partial sums of the binary tree's SAH cost terms (relative area,
// times the triangle count for a leaf), each block's nodes in index order.
__global__ __launch_bounds__(COL_BLOCK) void k_col_sah(const Node2* __restrict__ bn, uint32_t n_nodes,
                                                       double root_area, double* __restrict__ part) {
  __shared__ double s[COL_BLOCK];
  const uint32_t i = blockIdx.x * COL_BLOCK + threadIdx.x;
  double v = 0.0;
  if (i < n_nodes) {
    const Node2& nd = bn[i];
    const double a = (double)col_area(nd.lo, nd.hi) / root_area;
    v = nd.count ? a * nd.count : a;
  }
  s[threadIdx.x] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (uint32_t k = 0; k < COL_BLOCK; ++k) t += s[k];
    part[blockIdx.x] = t;
  }
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

namespace {

// Collapses the binary tree bn[0, n_bin) (device) into W-wide quantized nodes
// and fills `out` like the host path's upload_wide_bvh; d_order is the leaf
// order of the triangles d_tri.
template <int W>
apt_status collapse_device(const Node2* bn, uint32_t n_bin, const uint32_t* d_order, const float* d_tri,
                           uint32_t n, hipStream_t s, apt_bvh* out) {
  Node2 root;
  APT_HIP(hipMemcpyAsync(&root, bn, sizeof(Node2), hipMemcpyDeviceToHost, s));
  APT_HIP(hipStreamSynchronize(s));
  // SAH cost relative to the root area (bvh2_sah_cost), summed per block of
  // 256 nodes and then over the blocks in order
  double sah_cost = 0.0;
  {
    const float dx = root.hi[0] - root.lo[0], dy = root.hi[1] - root.lo[1], dz = root.hi[2] - root.lo[2];
    const double root_area = std::max(1e-30f, 2.0f * (dx * dy + dy * dz + dz * dx));
    const uint32_t nb = (n_bin + COL_BLOCK - 1) / COL_BLOCK;
    DBuf<double> part;
    APT_HIP(part.alloc(nb));
    hipLaunchKernelGGL(k_col_sah, dim3(nb), dim3(COL_BLOCK), 0, s, bn, n_bin, root_area, part.p);
    std::vector<double> h(nb);
    APT_HIP(hipMemcpyAsync(h.data(), part.p, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, s));
    APT_HIP(hipStreamSynchronize(s));
    for (double v : h) sah_cost += v;
  }
  // scene_pad: 2^-14 of the largest coordinate magnitude, which the root box holds
  float m_abs = 0.0f;
  for (int a = 0; a < 3; ++a) m_abs = std::max(m_abs, std::max(std::fabs(root.lo[a]), std::fabs(root.hi[a])));
  const float pad = std::max(m_abs * (1.0f / 16384.0f), 1e-30f);

  uint32_t root_ref = 0, n_wide = 0, depth = 0, stats_h[2] = {1u, 0u};
  DBuf<WNode<W>> wide;
  DBuf<uint32_t> wkids;
  if (root.count) {
    root_ref = make_leaf_ref(root.first, root.count);
    stats_h[1] = 1;
  } else {
    const uint32_t cap_w = n_bin;  // wide nodes <= internal binary nodes
    const uint32_t cap_f = n + 1;  // nodes of one level
    DBuf<uint32_t> front[2], fstack[2], nkni, blk, stats, total;
    APT_HIP(wide.alloc(cap_w));
    APT_HIP(wkids.alloc((size_t)cap_w * W));
    for (int k = 0; k < 2; ++k) {
      APT_HIP(front[k].alloc(cap_f));
      APT_HIP(fstack[k].alloc(cap_f));
    }
    APT_HIP(nkni.alloc(cap_f));
    APT_HIP(blk.alloc((cap_f + COL_BLOCK - 1) / COL_BLOCK));
    APT_HIP(stats.alloc(2));
    APT_HIP(total.alloc(1));
    APT_HIP(hipMemsetAsync(wide.p, 0, (size_t)cap_w * sizeof(WNode<W>), s));
    const uint32_t zero = 0u;
    APT_HIP(hipMemcpyAsync(front[0].p, &zero, 4, hipMemcpyHostToDevice, s));
    APT_HIP(hipMemcpyAsync(fstack[0].p, &zero, 4, hipMemcpyHostToDevice, s));
    APT_HIP(hipMemcpyAsync(stats.p, stats_h, sizeof(stats_h), hipMemcpyHostToDevice, s));
    uint32_t m = 1, base = 0, cur = 0;
    depth = 1;
    while (true) {
      APT_REQUIRE(base + m <= cap_w, APT_ERR_DEVICE, "bvh: device collapse ran out of nodes");
      const uint32_t nb = (m + COL_BLOCK - 1) / COL_BLOCK;
      hipLaunchKernelGGL((k_col_open<W>), dim3(nb), dim3(COL_BLOCK), 0, s, bn, front[cur].p, m, base, wkids.p,
                         nkni.p, blk.p);
      hipLaunchKernelGGL(k_col_scan, dim3(1), dim3(SCAN_BLOCK), 0, s, blk.p, nb, total.p);
      hipLaunchKernelGGL((k_col_write<W>), dim3(nb), dim3(COL_BLOCK), 0, s, bn, fstack[cur].p, m, base, wkids.p,
                         nkni.p, blk.p, wide.p, front[cur ^ 1].p, fstack[cur ^ 1].p, stats.p);
      uint32_t next = 0;
      APT_HIP(hipMemcpyAsync(&next, total.p, 4, hipMemcpyDeviceToHost, s));
      APT_HIP(hipStreamSynchronize(s));
      base += m;
      if (next == 0) break;
      APT_REQUIRE(next <= cap_f, APT_ERR_DEVICE, "bvh: device collapse level overflow");
      m = next;
      cur ^= 1;
      ++depth;
    }
    n_wide = base;
    hipLaunchKernelGGL((k_col_quant<W>), dim3((n_wide + COL_BLOCK - 1) / COL_BLOCK), dim3(COL_BLOCK), 0, s, bn,
                       wkids.p, n_wide, pad, wide.p);
    APT_HIP(hipMemcpyAsync(stats_h, stats.p, sizeof(stats_h), hipMemcpyDeviceToHost, s));
  }

  const size_t node_bytes = std::max<size_t>(1, n_wide) * sizeof(WNode<W>);
  const size_t tri_bytes = std::max<size_t>(1, (size_t)n * 3) * sizeof(float4);
  void* d_nodes = nullptr;
  void* d_tris = nullptr;
  APT_HIP(hipMalloc(&d_nodes, node_bytes));
  hipError_t e = hipMalloc(&d_tris, tri_bytes);
  if (e != hipSuccess) {
    (void)hipFree(d_nodes);
    set_error("hipMalloc of %zu triangle bytes failed: %s", tri_bytes, hipGetErrorString(e));
    return APT_ERR_OOM;
  }
  e = n_wide ? hipMemcpyAsync(d_nodes, wide.p, (size_t)n_wide * sizeof(WNode<W>), hipMemcpyDeviceToDevice, s)
             : hipSuccess;
  if (e == hipSuccess && n) {
    hipLaunchKernelGGL(k_gather_tris, dim3((n + 255) / 256), dim3(256), 0, s, d_tri, d_order, n,
                       static_cast<float4*>(d_tris));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(d_nodes);
    (void)hipFree(d_tris);
    set_error("bvh: device collapse failed: %s", hipGetErrorString(e));
    return APT_ERR_DEVICE;
  }
  out->width = W;
  out->n_tris = n;
  out->n_nodes = n_wide;
  out->root = root_ref;
  out->d_nodes = d_nodes;
  out->d_tris = (float4*)d_tris;
  e = hipMalloc(&out->d_counters, APT_BVH_COUNTERS * POOL_WORDS * sizeof(uint32_t));
  if (e != hipSuccess) {
    out->d_counters = nullptr;
    set_error("hipMalloc of the ray-pool counters failed: %s", hipGetErrorString(e));
    return APT_ERR_OOM;
  }
  out->stats.n_tris = n;
  out->stats.n_nodes = n_wide;
  out->stats.n_leaves = stats_h[1];
  out->stats.width = W;
  out->stats.depth = n_wide ? depth : 0;
  out->stats.max_stack = stats_h[0];
  out->stats.device_bytes = node_bytes + tri_bytes;
  out->stats.sah_cost = sah_cost;
  return APT_OK;
}

}  // namespace


// Builds the binary tree on the device from device-resident triangles, then
// collapses it there (collapse_device): only level sizes and statistics come
// back to the host.
apt_status build_wide_device(const float* d_tri, uint32_t n, uint32_t max_leaf, uint32_t bins, uint32_t width,
                             hipStream_t s, apt_bvh* out) {
  APT_REQUIRE(width == 2 || width == 4 || width == 8, APT_ERR_INVALID_ARG, "bvh: width %u", width);
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
  DBuf<uint32_t> blk_sum;
  APT_HIP(blk_sum.alloc(NEXT_MAX_BLOCKS));

  APT_HIP(bad.alloc(1));
  APT_HIP(nodes.alloc(cap_nodes));
  APT_HIP(keys.alloc(12 * (size_t)ROOT_GRID));  // per-work-group root keys

  const uint32_t nb = (n + 255) / 256;
  const uint32_t bad_init = 0xFFFFFFFFu;
  APT_HIP(hipMemcpyAsync(bad.p, &bad_init, 4, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_prims, dim3(nb), dim3(256), 0, s, d_tri, n, prim.p, order[0].p, bad.p);
  const uint32_t n_part = std::min(nb, ROOT_GRID);
  hipLaunchKernelGGL(k_root_bounds, dim3(n_part), dim3(ROOT_BLOCK), 0, s, prim.p, n, keys.p);
  hipLaunchKernelGGL(k_root_init, dim3(1), dim3(WAVE), 0, s, keys.p, n_part, n, node_ids[0].p, range[0].p,
                     box[0].p);
  uint32_t bad_h = 0;
  APT_HIP(hipMemcpyAsync(&bad_h, bad.p, 4, hipMemcpyDeviceToHost, s));
  APT_HIP(hipStreamSynchronize(s));
  APT_REQUIRE(bad_h == 0xFFFFFFFFu, APT_ERR_INVALID_ARG, "triangle %u has a non-finite coordinate", bad_h);

  // multi-block path buffers: at most n / LARGE_NODE large nodes per level
  const uint32_t max_large = n / LARGE_NODE + 1, max_blocks = n / LARGE_CHUNK + max_large;
  DBuf<uint32_t> l_act, l_blk_node, l_blk_first, l_blk_left;
  DBuf<int> l_bins, l_info, l_cbox;
  APT_HIP(l_act.alloc(max_large));
  APT_HIP(l_blk_first.alloc(max_large));
  APT_HIP(l_blk_node.alloc(max_blocks));
  APT_HIP(l_blk_left.alloc(max_blocks));
  APT_HIP(l_bins.alloc((size_t)max_large * BIN_WORDS));
  APT_HIP(l_info.alloc(4 * (size_t)max_large));
  APT_HIP(l_cbox.alloc(24 * (size_t)max_large));
  std::vector<uint32_t> h_range, h_act, h_blk_node, h_blk_first;

  uint32_t n_active = 1, node_base = 1, depth = 0;
  int cur = 0, ord = 0;
  while (n_active > 0) {
    // Large nodes of this level (looked up while a level has few nodes: a
    // node above LARGE_NODE deeper down is split by k_level alone).
    h_act.clear();
    h_blk_node.clear();
    h_blk_first.clear();
    if (n_active <= 2 * max_large) {
      h_range.resize(2 * (size_t)n_active);
      APT_HIP(hipMemcpyAsync(h_range.data(), range[cur].p, h_range.size() * 4, hipMemcpyDeviceToHost, s));
      APT_HIP(hipStreamSynchronize(s));
      for (uint32_t a = 0; a < n_active && h_act.size() < max_large; ++a) {
        const uint32_t len = h_range[2 * a + 1] - h_range[2 * a];
        if (len <= LARGE_NODE) continue;
        h_blk_first.push_back((uint32_t)h_blk_node.size());
        for (uint32_t c = 0; c < (len + LARGE_CHUNK - 1) / LARGE_CHUNK; ++c)
          h_blk_node.push_back((uint32_t)h_act.size());
        h_act.push_back(a);
      }
    }
    const uint32_t n_large = (uint32_t)h_act.size(), blocks = (uint32_t)h_blk_node.size();
    if (n_large) {
      APT_HIP(hipMemcpyAsync(l_act.p, h_act.data(), n_large * 4, hipMemcpyHostToDevice, s));
      APT_HIP(hipMemcpyAsync(l_blk_first.p, h_blk_first.data(), n_large * 4, hipMemcpyHostToDevice, s));
      APT_HIP(hipMemcpyAsync(l_blk_node.p, h_blk_node.data(), blocks * 4, hipMemcpyHostToDevice, s));
    }
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
    L.large = n_large ? LARGE_NODE : 0xFFFFFFFFu;
    hipLaunchKernelGGL(k_level, dim3(n_active), dim3(WAVE), 0, s, L);
    if (n_large) {
      Large G;
      G.n = n_large;
      G.act = l_act.p;
      G.blk_node = l_blk_node.p;
      G.blk_first = l_blk_first.p;
      G.bins = l_bins.p;
      G.info = l_info.p;
      G.blk_left = l_blk_left.p;
      G.cbox = l_cbox.p;
      G.blocks = blocks;
      hipLaunchKernelGGL(k_large_init, dim3(n_large), dim3(LBLOCK), 0, s, G);
      hipLaunchKernelGGL(k_large_bin, dim3(blocks), dim3(LBLOCK), 0, s, L, G);
      hipLaunchKernelGGL(k_large_sweep, dim3(n_large), dim3(WAVE), 0, s, L, G);
      hipLaunchKernelGGL(k_large_count, dim3(blocks), dim3(LBLOCK), 0, s, L, G);
      hipLaunchKernelGGL(k_large_scan, dim3(1), dim3(1), 0, s, G);
      hipLaunchKernelGGL(k_large_scatter, dim3(blocks), dim3(LBLOCK), 0, s, L, G);
      hipLaunchKernelGGL(k_large_final, dim3(n_large), dim3(1), 0, s, L, G);
    }
    uint32_t chunk = (n_active + NEXT_MAX_BLOCKS - 1) / NEXT_MAX_BLOCKS;
    chunk = std::max(NEXT_CHUNK, (chunk + NEXT_BLOCK - 1) / NEXT_BLOCK * NEXT_BLOCK);
    const uint32_t nb_next = (n_active + chunk - 1) / chunk;  // <= NEXT_MAX_BLOCKS
    hipLaunchKernelGGL(k_next_count, dim3(nb_next), dim3(NEXT_BLOCK), 0, s, L, chunk, blk_sum.p);
    hipLaunchKernelGGL(k_next_scan, dim3(1), dim3(SCAN_BLOCK), 0, s, blk_sum.p, nb_next, counter.p);
    hipLaunchKernelGGL(k_next_write, dim3(nb_next), dim3(NEXT_BLOCK), 0, s, L, chunk, blk_sum.p, node_base,
                       node_ids[cur ^ 1].p, range[cur ^ 1].p, box[cur ^ 1].p);
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
  switch (width) {
    case 2: return collapse_device<2>(nodes.p, node_base, order[ord].p, d_tri, n, s, out);
    case 4: return collapse_device<4>(nodes.p, node_base, order[ord].p, d_tri, n, s, out);
    default: return collapse_device<8>(nodes.p, node_base, order[ord].p, d_tri, n, s, out);
  }
}

}  // namespace apt
