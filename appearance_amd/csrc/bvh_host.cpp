// Host binned-SAH BVH build and the collapse of the binary tree into the
// 2/4/8-wide device layout.
//
// This replaces the driver BLAS build of the reference
// (src/scene_resources/scene_model.rs:53-88, PREFER_FAST_TRACE) with an
// explicit, deterministic builder: the split of every node depends only on the
// set of triangles in it, so the tree topology does not depend on thread
// timing, and the breadth-first collapse gives a canonical node order.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>

#include "apt_internal.h"

namespace apt {
namespace {

struct Prim {
  float lo[3], hi[3], c[3];
};

struct Box {
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  void grow(const float* l, const float* h) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], l[k]);
      hi[k] = std::max(hi[k], h[k]);
    }
  }
  void grow(const Box& b) { grow(b.lo, b.hi); }
  float area() const {
    float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0.0f || dy < 0.0f || dz < 0.0f) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
  }
};

struct Builder {
  const Prim* prims;
  uint32_t* order;
  Node2* nodes;
  std::atomic<uint32_t> next_node{1};
  uint32_t max_leaf;
  uint32_t bins;
  static constexpr uint32_t PAR_MIN = 1u << 15;  // spawn a thread above this size
  static constexpr int PAR_DEPTH = 4;            // ... and this shallow (<= 16 threads)
  static constexpr int MAX_DEPTH = 56;           // beyond this, median splits only

  void build(uint32_t node, uint32_t begin, uint32_t end, int depth) {
    Box bb, cb;
    for (uint32_t i = begin; i < end; ++i) {
      const Prim& p = prims[order[i]];
      bb.grow(p.lo, p.hi);
      cb.grow(p.c, p.c);
    }
    Node2& nd = nodes[node];
    std::memcpy(nd.lo, bb.lo, sizeof(nd.lo));
    std::memcpy(nd.hi, bb.hi, sizeof(nd.hi));
    const uint32_t n = end - begin;
    if (n <= max_leaf) {
      nd.first = begin;
      nd.count = n;
      nd.left = nd.right = 0;
      return;
    }

    uint32_t mid = begin + n / 2;
    int best_axis = -1;
    uint32_t best_split = 0;
    if (depth < MAX_DEPTH) {
      float best_cost = INFINITY;
      std::vector<Box> bin_box(bins);
      std::vector<uint32_t> bin_cnt(bins);
      std::vector<float> right_area(bins);
      std::vector<uint32_t> right_cnt(bins);
      for (int axis = 0; axis < 3; ++axis) {
        float ext = cb.hi[axis] - cb.lo[axis];
        if (!(ext > 0.0f)) continue;
        float scale = (float)bins / ext;
        std::fill(bin_box.begin(), bin_box.end(), Box());
        std::fill(bin_cnt.begin(), bin_cnt.end(), 0u);
        for (uint32_t i = begin; i < end; ++i) {
          const Prim& p = prims[order[i]];
          int b = (int)((p.c[axis] - cb.lo[axis]) * scale);
          b = std::min(std::max(b, 0), (int)bins - 1);
          bin_box[b].grow(p.lo, p.hi);
          bin_cnt[b]++;
        }
        Box acc;
        uint32_t cnt = 0;
        for (int b = (int)bins - 1; b > 0; --b) {
          acc.grow(bin_box[b]);
          cnt += bin_cnt[b];
          right_area[b] = acc.area();
          right_cnt[b] = cnt;
        }
        acc = Box();
        cnt = 0;
        for (uint32_t b = 1; b < bins; ++b) {
          acc.grow(bin_box[b - 1]);
          cnt += bin_cnt[b - 1];
          if (cnt == 0 || right_cnt[b] == 0) continue;
          float cost = acc.area() * (float)cnt + right_area[b] * (float)right_cnt[b];
          if (cost < best_cost) {
            best_cost = cost;
            best_axis = axis;
            best_split = b;
          }
        }
      }
    }
    if (best_axis >= 0) {
      const int axis = best_axis;
      const float lo = cb.lo[axis];
      const float scale = (float)bins / (cb.hi[axis] - cb.lo[axis]);
      const int bins_i = (int)bins;
      const uint32_t split = best_split;
      // stable: both halves keep the input order, so the tree is a function of
      // the input alone and the device builder reproduces it exactly
      uint32_t* m = std::stable_partition(order + begin, order + end, [&](uint32_t idx) {
        int b = (int)((prims[idx].c[axis] - lo) * scale);
        b = std::min(std::max(b, 0), bins_i - 1);
        return (uint32_t)b < split;
      });
      mid = (uint32_t)(m - order);
    } else {
      // Coincident centroids (or too deep): object median on the input order
      // (stable partitions keep every range sorted; the sort only restates it).
      std::sort(order + begin, order + end);
    }

    uint32_t l = next_node.fetch_add(2);
    nd.left = l;
    nd.right = l + 1;
    nd.first = 0;
    nd.count = 0;
    if (n >= PAR_MIN && depth < PAR_DEPTH) {
      std::thread t([this, l, begin, mid, depth] { build(l, begin, mid, depth + 1); });
      build(l + 1, mid, end, depth + 1);
      t.join();
    } else {
      build(l, begin, mid, depth + 1);
      build(l + 1, mid, end, depth + 1);
    }
  }
};

float box_area(const float* lo, const float* hi) {
  float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
  return 2.0f * (dx * dy + dy * dz + dz * dx);
}

}  // namespace

// Padding added to every node box: 2^-14 of the largest vertex coordinate
// magnitude.  It covers the rounding of the slab test and of the triangle t
// (tens of ulps for rays starting within ~100x the scene extent), so a
// triangle whose rounded hit lies on or just outside its exact box is never
// culled.
float scene_pad(const float* tri_xyz, uint32_t n) {
  float m = 0.0f;
  for (uint64_t i = 0; i < (uint64_t)n * 9; ++i) m = std::max(m, std::fabs(tri_xyz[i]));
  return std::max(m * (1.0f / 16384.0f), 1e-30f);
}

apt_status build_bvh2_host(const float* tri_xyz, uint32_t n, uint32_t max_leaf, uint32_t bins,
                           BvhHostResult& out) {
  std::vector<Prim> prims(n);
  for (uint32_t i = 0; i < n; ++i) {
    const float* t = tri_xyz + (uint64_t)i * 9;
    Prim& p = prims[i];
    for (int k = 0; k < 3; ++k) {
      p.lo[k] = std::min(t[k], std::min(t[3 + k], t[6 + k]));
      p.hi[k] = std::max(t[k], std::max(t[3 + k], t[6 + k]));
      p.c[k] = (p.lo[k] + p.hi[k]) * 0.5f;
      if (!std::isfinite(t[k]) || !std::isfinite(t[3 + k]) || !std::isfinite(t[6 + k])) {
        set_error("triangle %u has a non-finite coordinate", i);
        return APT_ERR_INVALID_ARG;
      }
    }
  }
  out.order.resize(n);
  for (uint32_t i = 0; i < n; ++i) out.order[i] = i;
  out.nodes.assign(std::max<size_t>(1, 2 * (size_t)n), Node2{});
  Builder b;
  b.prims = prims.data();
  b.order = out.order.data();
  b.nodes = out.nodes.data();
  b.max_leaf = max_leaf;
  b.bins = bins;
  b.build(0, 0, n, 0);
  out.nodes.resize(b.next_node.load());

  out.sah_cost = bvh2_sah_cost(out.nodes);
  return APT_OK;
}

// SAH cost (C_trav = 1, C_isect = 1) relative to the root area.
double bvh2_sah_cost(const std::vector<Node2>& nodes) {
  if (nodes.empty()) return 0.0;
  double root_area = std::max(1e-30f, box_area(nodes[0].lo, nodes[0].hi));
  double cost = 0.0;
  for (const Node2& nd : nodes) {
    double a = box_area(nd.lo, nd.hi) / root_area;
    cost += nd.count ? a * nd.count : a;
  }
  return cost;
}

namespace {

// Host worker threads that stay alive between calls: the collapse runs two
// parallel passes per tree level, and starting threads for each would cost
// more than the passes.  One job at a time (callers hold `call_mu`).
class HostPool {
 public:
  static HostPool& get() {
    static HostPool* p = new HostPool();  // never destroyed: the workers just wait at exit
    return *p;
  }
  uint32_t parts() const { return (uint32_t)workers_.size() + 1u; }
  // Runs job(t) for t in [0, parts()) and returns when all have finished.
  void run(const std::function<void(uint32_t)>& job) {
    std::lock_guard<std::mutex> call(call_mu_);
    {
      std::lock_guard<std::mutex> lock(mu_);
      job_ = &job;
      pending_ = (uint32_t)workers_.size();
      ++gen_;
    }
    cv_.notify_all();
    job(0);
    std::unique_lock<std::mutex> lock(mu_);
    done_.wait(lock, [&] { return pending_ == 0; });
    job_ = nullptr;
  }

 private:
  HostPool() {
    const uint32_t hw = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    for (uint32_t t = 1; t < hw; ++t) workers_.emplace_back([this, t] { loop(t); });
  }
  void loop(uint32_t t) {
    uint64_t seen = 0;
    while (true) {
      const std::function<void(uint32_t)>* job;
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return gen_ != seen; });
        seen = gen_;
        job = job_;
      }
      (*job)(t);
      std::lock_guard<std::mutex> lock(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex mu_, call_mu_;
  std::condition_variable cv_, done_;
  const std::function<void(uint32_t)>* job_ = nullptr;
  uint32_t pending_ = 0;
  uint64_t gen_ = 0;
};

// Runs f(begin, end) over [0, n) in contiguous chunks on the host's cores.
template <class F>
void parallel_for(uint32_t n, F f) {
  if (n < 4096) {
    f(0u, n);
    return;
  }
  HostPool& pool = HostPool::get();
  const uint32_t parts = pool.parts();
  pool.run([&](uint32_t t) {
    const uint32_t i0 = (uint32_t)((uint64_t)n * t / parts), i1 = (uint32_t)((uint64_t)n * (t + 1) / parts);
    if (i0 < i1) f(i0, i1);
  });
}

// Grid of one node: per axis the smallest power-of-two step that spans the
// node box in 255 steps, then each child box rounded outwards to the grid
// (checked in double, so origin + q * step never cuts into the padded box).
template <int W>
void quantize_node(WNode<W>& w, int nk, const float (&clo)[W][3], const float (&chi)[W][3],
                   const float (&nlo)[3], const float (&nhi)[3]) {
  float* org[3] = {&w.ox, &w.oy, &w.oz};
  uint8_t* ex[3] = {&w.ex, &w.ey, &w.ez};
  for (int a = 0; a < 3; ++a) {
    const double lo = nk ? (double)nlo[a] : 0.0, ext = nk ? (double)nhi[a] - lo : 0.0;
    int e = -126;
    if (ext > 0.0) {
      (void)std::frexp(ext / 255.0, &e);  // 2^(e-1) <= ext/255 < 2^e
      e = std::max(-126, std::min(127, e));
      while (e > -126 && ext <= 255.0 * std::ldexp(1.0, e - 1)) --e;
      while (e < 127 && ext > 255.0 * std::ldexp(1.0, e)) ++e;
    }
    *org[a] = (float)lo;
    *ex[a] = (uint8_t)(e + 127);
    const double step = std::ldexp(1.0, e);
    for (int k = 0; k < W; ++k) {
      if (k >= nk) {
        w.q[a][k] = 255;
        w.q[a + 3][k] = 0;
        continue;
      }
      double ql = std::floor(((double)clo[k][a] - lo) / step);
      double qh = std::ceil(((double)chi[k][a] - lo) / step);
      ql = std::min(255.0, std::max(0.0, ql));
      qh = std::min(255.0, std::max(0.0, qh));
      while (ql > 0.0 && lo + ql * step > (double)clo[k][a]) ql -= 1.0;
      while (qh < 255.0 && lo + qh * step < (double)chi[k][a]) qh += 1.0;
      w.q[a][k] = (uint8_t)ql;
      w.q[a + 3][k] = (uint8_t)qh;
    }
  }
}

// The wide tree of a binary one (breadth first), its root ref and statistics.
template <int W>
struct WideTree {
  std::vector<WNode<W>> wide;
  uint32_t root = 0, max_depth = 1, max_stack = 1, n_leaves = 0;
};

template <int W>
void collapse_tree(const BvhHostResult& b2, float pad, WideTree<W>& out) {
  const std::vector<Node2>& bn = b2.nodes;
  std::vector<WNode<W>>& wide = out.wide;
  std::vector<uint32_t> wkids;  // the W BVH2 nodes under each wide node (~0u: empty slot)
  uint32_t max_depth = 1, max_stack = 1, n_leaves = 0;

  auto leaf_ref = [&](const Node2& nd) { return make_leaf_ref(nd.first, nd.count); };

  // Breadth first, one level of wide nodes at a time: the nodes of a level
  // are opened in parallel, and a prefix sum over their internal children
  // numbers the next level in the order a FIFO queue would (parent order,
  // then child slot).
  uint32_t root;
  std::vector<uint32_t> front, fstack;  // BVH2 node and worst-case stack of each wide node of the level
  if (bn[0].count) {
    root = leaf_ref(bn[0]);
    n_leaves = 1;
  } else {
    root = 0;
    front.push_back(0);
    fstack.push_back(0);
  }
  for (uint32_t depth = 1; !front.empty(); ++depth) {
    const uint32_t m = (uint32_t)front.size(), base = (uint32_t)wide.size();
    wide.resize((size_t)base + m);
    wkids.resize(((size_t)base + m) * W);
    std::vector<uint8_t> nk_of(m), nint(m);
    parallel_for(m, [&](uint32_t i0, uint32_t i1) {
      for (uint32_t i = i0; i < i1; ++i) {
        // Open the binary node into up to W children, greedily expanding the
        // internal child of largest surface area.
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
            float a = box_area(c.lo, c.hi);
            if (a > best_a) {
              best_a = a;
              best = k;
            }
          }
          if (best < 0) break;
          uint32_t x = kids[best];
          kids[best] = bn[x].left;
          kids[nk++] = bn[x].right;
        }
        int ni = 0;
        for (int k = 0; k < W; ++k) {
          wkids[((size_t)base + i) * W + k] = k < nk ? kids[k] : 0xFFFFFFFFu;
          if (k < nk && !bn[kids[k]].count) ++ni;
        }
        nk_of[i] = (uint8_t)nk;
        nint[i] = (uint8_t)ni;
      }
    });
    std::vector<uint32_t> off((size_t)m + 1, 0u);
    for (uint32_t i = 0; i < m; ++i) {
      off[i + 1] = off[i] + nint[i];
      // Worst-case stack: every sibling but one may be pushed at each level.
      max_stack = std::max(max_stack, fstack[i] + (uint32_t)nk_of[i]);
      n_leaves += (uint32_t)(nk_of[i] - nint[i]);
    }
    const uint32_t next_base = base + m;
    std::vector<uint32_t> nfront(off[m]), nstack(off[m]);
    parallel_for(m, [&](uint32_t i0, uint32_t i1) {
      for (uint32_t i = i0; i < i1; ++i) {
        const uint32_t here = fstack[i] + (uint32_t)nk_of[i] - 1u;
        WNode<W> w;
        std::memset(&w, 0, sizeof(w));
        uint32_t j = off[i];
        for (int k = 0; k < W; ++k) {
          const uint32_t kid = wkids[((size_t)base + i) * W + k];
          if (kid == 0xFFFFFFFFu) {
            w.child[k] = REF_EMPTY;
          } else if (bn[kid].count) {
            w.child[k] = leaf_ref(bn[kid]);
          } else {
            w.child[k] = next_base + j;
            nfront[j] = kid;
            nstack[j] = here;
            ++j;
          }
        }
        wide[(size_t)base + i] = w;
      }
    });
    if (off[m]) max_depth = std::max(max_depth, depth + 1);
    front.swap(nfront);
    fstack.swap(nstack);
  }

  // Child grids of every wide node (independent per node: in parallel).
  parallel_for((uint32_t)wide.size(), [&](uint32_t i0, uint32_t i1) {
    for (uint32_t i = i0; i < i1; ++i) {
      const uint32_t* kids = &wkids[(size_t)i * W];
      int nk = 0;
      while (nk < W && kids[nk] != 0xFFFFFFFFu) ++nk;
      float clo[W][3], chi[W][3], nlo[3] = {INFINITY, INFINITY, INFINITY};
      float nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int k = 0; k < nk; ++k) {
        const Node2& c = bn[kids[k]];
        for (int a = 0; a < 3; ++a) {
          clo[k][a] = c.lo[a] - pad;
          chi[k][a] = c.hi[a] + pad;
          nlo[a] = std::min(nlo[a], clo[k][a]);
          nhi[a] = std::max(nhi[a], chi[k][a]);
        }
      }
      quantize_node<W>(wide[i], nk, clo, chi, nlo, nhi);
    }
  });
  out.root = root;
  out.max_depth = max_depth;
  out.max_stack = max_stack;
  out.n_leaves = n_leaves;
}

template <int W>
apt_status collapse_upload(const float* tri_xyz, uint32_t n, const BvhHostResult& b2, float pad, apt_bvh* out,
                           bool upload) {
  WideTree<W> wt;
  collapse_tree<W>(b2, pad, wt);
  const std::vector<WNode<W>>& wide = wt.wide;
  const uint32_t root = wt.root, max_depth = wt.max_depth, max_stack = wt.max_stack, n_leaves = wt.n_leaves;

  // Triangles in leaf order: v0 | prim, v1, v2 as float4.
  std::vector<float4> tris((size_t)n * 3);
  parallel_for(n, [&](uint32_t i0, uint32_t i1) {
    for (uint32_t i = i0; i < i1; ++i) {
      const uint32_t p = b2.order[i];
      const float* t = tri_xyz + (uint64_t)p * 9;
      tris[3 * (size_t)i + 0] = make_float4(t[0], t[1], t[2], u2f(p));
      tris[3 * (size_t)i + 1] = make_float4(t[3], t[4], t[5], 0.0f);
      tris[3 * (size_t)i + 2] = make_float4(t[6], t[7], t[8], 0.0f);
    }
  });

  size_t node_bytes = std::max<size_t>(1, wide.size()) * sizeof(WNode<W>);
  size_t tri_bytes = std::max<size_t>(1, (size_t)n * 3) * sizeof(float4);
  out->stats.n_tris = n;
  out->stats.n_nodes = (uint32_t)wide.size();
  out->stats.n_leaves = n_leaves;
  out->stats.width = W;
  out->stats.depth = wide.empty() ? 0 : max_depth;
  out->stats.max_stack = max_stack;
  out->stats.device_bytes = node_bytes + tri_bytes;
  out->stats.sah_cost = b2.sah_cost;
  out->width = W;
  out->n_tris = n;
  out->n_nodes = (uint32_t)wide.size();
  out->root = root;
  if (!upload) return APT_OK;
  void* d_nodes = nullptr;
  void* d_tris = nullptr;
  APT_HIP(hipMalloc(&d_nodes, node_bytes));
  hipError_t e = hipMalloc(&d_tris, tri_bytes);
  if (e != hipSuccess) {
    (void)hipFree(d_nodes);
    set_error("hipMalloc of %zu triangle bytes failed: %s", tri_bytes, hipGetErrorString(e));
    return APT_ERR_OOM;
  }
  if (!wide.empty()) APT_HIP(hipMemcpy(d_nodes, wide.data(), wide.size() * sizeof(WNode<W>), hipMemcpyHostToDevice));
  if (!tris.empty()) APT_HIP(hipMemcpy(d_tris, tris.data(), tris.size() * sizeof(float4), hipMemcpyHostToDevice));

  out->d_nodes = d_nodes;
  out->d_tris = (float4*)d_tris;
  e = hipMalloc(&out->d_counters, APT_BVH_COUNTERS * POOL_WORDS * sizeof(uint32_t));
  if (e != hipSuccess) {
    out->d_counters = nullptr;
    set_error("hipMalloc of the ray-pool counters failed: %s", hipGetErrorString(e));
    return APT_ERR_OOM;
  }
  return APT_OK;
}

template <int W>
void collapse_host_bytes(const BvhHostResult& b2, float pad, std::vector<uint8_t>& nodes, uint32_t& root,
                         apt_bvh_stats& stats) {
  WideTree<W> wt;
  collapse_tree<W>(b2, pad, wt);
  nodes.resize(wt.wide.size() * sizeof(WNode<W>));
  if (!wt.wide.empty()) std::memcpy(nodes.data(), wt.wide.data(), nodes.size());
  root = wt.root;
  stats.n_nodes = (uint32_t)wt.wide.size();
  stats.n_leaves = wt.n_leaves;
  stats.width = W;
  stats.depth = wt.wide.empty() ? 0 : wt.max_depth;
  stats.max_stack = wt.max_stack;
}

}  // namespace

apt_status collapse_wide_host(const BvhHostResult& b2, uint32_t width, float pad, std::vector<uint8_t>& nodes,
                              uint32_t& root, apt_bvh_stats& stats) {
  switch (width) {
    case 2: collapse_host_bytes<2>(b2, pad, nodes, root, stats); return APT_OK;
    case 4: collapse_host_bytes<4>(b2, pad, nodes, root, stats); return APT_OK;
    case 8: collapse_host_bytes<8>(b2, pad, nodes, root, stats); return APT_OK;
    default:
      set_error("BVH width %u is not one of 2, 4, 8", width);
      return APT_ERR_INVALID_ARG;
  }
}

apt_status upload_wide_bvh(const float* tri_xyz, uint32_t n, const BvhHostResult& b2, uint32_t width, float pad,
                           apt_bvh* out, bool upload) {
  switch (width) {
    case 2: return collapse_upload<2>(tri_xyz, n, b2, pad, out, upload);
    case 4: return collapse_upload<4>(tri_xyz, n, b2, pad, out, upload);
    case 8: return collapse_upload<8>(tri_xyz, n, b2, pad, out, upload);
    default:
      set_error("BVH width %u is not one of 2, 4, 8", width);
      return APT_ERR_INVALID_ARG;
  }
}

}  // namespace apt
