// Serving loop: continuous batching of render requests over resident scenes.
//
// The reference renders one job per render-node call
// (crates/appearance-render-loop/src/node.rs:44-105).  Here a scheduler
// thread owns the GPU: it admits queued requests while the accumulator budget
// and the active-slot table allow, packs tiles of every admitted request into
// one wavefront batch (requests that fit in one batch shortest-first, then the
// larger ones in arrival order), replays the batch from
// a captured hipGraph, keeps two batches in flight, and retires a request as
// soon as the batch holding its last paths completes.  Requests that arrive
// while others run join the next batch.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "apt_internal.h"
#include "scene_gpu.h"
#include "server.h"

namespace {

using namespace apt;
using Clock = std::chrono::steady_clock;

struct Req {
  uint64_t id = 0;
  apt_request r{};
  float* out = nullptr;
  uint32_t state = APT_REQ_QUEUED;
  apt_status status = APT_OK;
  double t_submit = 0, t_admit = 0, t_done = 0;
  uint64_t total = 0, issued = 0, finished = 0;
  int slot = -1;
  int inflight = 0;
  bool cancel = false;
  unsigned long long* d_accum = nullptr;
  float* d_out = nullptr;
  size_t footprint = 0;
};

struct Batch {
  int buf = 0;
  uint32_t depth = 0;  // deepest request of the batch
  std::vector<std::pair<Req*, uint32_t>> parts;
  hipEvent_t t0 = nullptr, t1 = nullptr;
};

}  // namespace

struct apt_server {
  apt_server_config cfg{};
  int device = 0;
  Clock::time_point epoch = Clock::now();
  std::vector<SceneGpu> scenes;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> queued;
  std::unordered_map<uint64_t, Req*> all;
  std::vector<Req*> slots;
  uint64_t next_id = 1;
  bool stop = false;
  uint64_t accum_used = 0;
  apt_server_stats stats{};
  std::thread worker;

  // Per-request accumulators and results live in one arena of accum_budget
  // bytes allocated at creation: admission and retirement take and return
  // ranges of it (first fit, neighbours merged) instead of calling hipMalloc
  // and hipFree, which can stall the device, for every request.
  unsigned char* arena = nullptr;
  std::map<size_t, size_t> arena_free;  // offset -> bytes

  // device
  hipStream_t stream = nullptr, side = nullptr;
  float4 *p_o = nullptr, *p_d = nullptr, *p_t = nullptr, *p_l = nullptr, *p_h = nullptr;
  float4 *p_so = nullptr, *p_sd = nullptr, *p_sc = nullptr;
  uint32_t *p_rng = nullptr, *p_q0 = nullptr, *p_q1 = nullptr, *p_qs = nullptr, *p_cnt = nullptr;
  uint32_t *path_req = nullptr, *path_pix = nullptr;
  SrvReq* d_reqs = nullptr;
  SrvReq* h_reqs = nullptr;           // pinned staging, one per slot
  SrvBatch* d_batch[2] = {nullptr, nullptr};
  SrvBatch* h_batch[2] = {nullptr, nullptr};  // pinned staging
  // per staging buffer, one captured batch per depth 1..max_bounces: a batch
  // runs as many bounces as its deepest request asks for
  std::vector<hipGraphExec_t> graph[2];
  SrvDevice dev[2];
  bool device_failed = false;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(Clock::now() - epoch).count();
  }
  ~apt_server();
};

namespace {

constexpr size_t ARENA_ALIGN = 256;

size_t arena_bytes(const Req* q) {
  const size_t npix = (size_t)q->r.width * q->r.height;
  const size_t acc = (npix * 3 * sizeof(unsigned long long) + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
  return acc + (npix * 3 * sizeof(float) + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
}

// Under the lock: a range of the arena for q's accumulator and result.
bool arena_take(apt_server* s, Req* q) {
  const size_t need = arena_bytes(q);
  for (auto it = s->arena_free.begin(); it != s->arena_free.end(); ++it) {
    if (it->second < need) continue;
    const size_t off = it->first, left = it->second - need;
    s->arena_free.erase(it);
    if (left) s->arena_free[off + need] = left;
    const size_t npix = (size_t)q->r.width * q->r.height;
    const size_t acc = (npix * 3 * sizeof(unsigned long long) + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN;
    q->d_accum = reinterpret_cast<unsigned long long*>(s->arena + off);
    q->d_out = reinterpret_cast<float*>(s->arena + off + acc);
    return true;
  }
  return false;
}

// Under the lock: return q's range to the arena, merged with free neighbours.
void release_request_memory(apt_server* s, Req* q) {
  if (!q->d_accum) return;
  size_t off = (size_t)(reinterpret_cast<unsigned char*>(q->d_accum) - s->arena), len = arena_bytes(q);
  auto next = s->arena_free.lower_bound(off);
  if (next != s->arena_free.end() && next->first == off + len) {
    len += next->second;
    next = s->arena_free.erase(next);
  }
  if (next != s->arena_free.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == off) {
      off = prev->first;
      len += prev->second;
      s->arena_free.erase(prev);
    }
  }
  s->arena_free[off] = len;
  q->d_accum = nullptr;
  q->d_out = nullptr;
}

// Under the lock: move queued requests into free slots while the footprint fits.
void admit(apt_server* s) {
  while (!s->queued.empty()) {
    Req* q = s->queued.front();
    int slot = -1;
    for (size_t i = 0; i < s->slots.size(); ++i)
      if (!s->slots[i]) { slot = (int)i; break; }
    if (slot < 0) return;
    // footprint admission: wait until the arena has a range for it
    if (!arena_take(s, q)) return;
    s->queued.pop_front();
    const size_t npix = (size_t)q->r.width * q->r.height;
    if (srv_zero(q->d_accum, npix * 3, s->stream) != APT_OK) {
      release_request_memory(s, q);
      q->state = APT_REQ_FAILED;
      q->status = APT_ERR_DEVICE;
      q->t_done = s->now_ms();
      s->cv_done.notify_all();
      continue;
    }
    SrvReq& h = s->h_reqs[slot];
    h.sc = scene_gpu_view(s->scenes[q->r.scene_id], q->r.camera, q->r.width, q->r.height);
    h.width = q->r.width;
    h.height = q->r.height;
    h.spp = q->r.spp;
    h.max_bounces = q->r.max_bounces;
    h.rr_start = q->r.rr_start_bounce;
    h.seed = q->r.seed;
    h.accum = q->d_accum;
    (void)hipMemcpyAsync(s->d_reqs + slot, &h, sizeof(SrvReq), hipMemcpyHostToDevice, s->stream);
    q->slot = slot;
    q->state = APT_REQ_ACTIVE;
    q->t_admit = s->now_ms();
    s->slots[slot] = q;
    s->accum_used += arena_bytes(q);
  }
}

// Under the lock: fill one batch with tiles of the active requests.
void build_batch(apt_server* s, Batch& b, SrvBatch& hb) {
  std::vector<Req*> act;
  for (Req* q : s->slots)
    if (q && !q->cancel && q->issued < q->total) act.push_back(q);
  if (s->cfg.policy == 0) {
    // Requests that fit in one batch go first, shortest remaining work
    // first; the larger ones follow in arrival order, so a stream that
    // outruns the device delays its large requests by the small requests'
    // share of the work only, not by every later request smaller than they.
    const uint64_t one = s->cfg.path_slots;
    std::stable_sort(act.begin(), act.end(), [one](const Req* a, const Req* c) {
      const bool sa = a->total <= one, sc = c->total <= one;
      if (sa != sc) return sa;
      if (sa) return a->total - a->issued < c->total - c->issued;
      return a->id < c->id;
    });
  } else if (s->cfg.policy == 2) {
    std::stable_sort(act.begin(), act.end(), [](const Req* a, const Req* c) {
      return a->total - a->issued < c->total - c->issued;
    });
  } else {
    std::stable_sort(act.begin(), act.end(), [](const Req* a, const Req* c) { return a->id < c->id; });
  }
  uint32_t room = s->cfg.path_slots;
  hb.n_tiles = 0;
  hb.n_paths = 0;
  for (Req* q : act) {
    if (room == 0 || hb.n_tiles >= SRV_MAX_TILES) break;
    uint32_t taken = 0;
    while (room > 0 && q->issued < q->total && hb.n_tiles < SRV_MAX_TILES) {
      const uint32_t n = (uint32_t)std::min<uint64_t>({(uint64_t)room, q->total - q->issued,
                                                       (uint64_t)s->cfg.tile_paths});
      SrvTile& t = hb.tiles[hb.n_tiles++];
      t.req_slot = (uint32_t)q->slot;
      t.first = (uint32_t)q->issued;
      t.count = n;
      t.offset = hb.n_paths;
      hb.n_paths += n;
      q->issued += n;
      room -= n;
      taken += n;
    }
    if (taken) {
      q->inflight++;
      b.parts.push_back({q, taken});
      b.depth = std::max(b.depth, q->r.max_bounces);
    }
  }
}

apt_status launch_batch(apt_server* s, Batch& b) {
  SrvBatch& hb = *s->h_batch[b.buf];
  const size_t bytes = offsetof(SrvBatch, tiles) + (size_t)hb.n_tiles * sizeof(SrvTile);
  APT_HIP(hipMemcpyAsync(s->d_batch[b.buf], &hb, bytes, hipMemcpyHostToDevice, s->stream));
  APT_HIP(hipEventRecord(b.t0, s->stream));
  const uint32_t depth = std::max<uint32_t>(1u, std::min(b.depth, s->cfg.max_bounces));
  hipGraphExec_t g = s->graph[b.buf].size() >= depth ? s->graph[b.buf][depth - 1] : nullptr;
  if (g) {
    APT_HIP(hipGraphLaunch(g, s->stream));
  } else {
    SrvDevice d = s->dev[b.buf];
    d.max_bounces = depth;
    apt_status st = srv_enqueue_batch(d, s->stream);
    if (st != APT_OK) return st;
  }
  APT_HIP(hipEventRecord(b.t1, s->stream));
  return APT_OK;
}

// After a batch's completion event: account its paths and retire requests.
void finalize(apt_server* s, Batch& b) {
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, b.t0, b.t1);
  std::vector<Req*> done, dropped;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stats.batches++;
    s->stats.gpu_busy_ms += ms;
    for (auto& pr : b.parts) {
      Req* q = pr.first;
      q->finished += pr.second;
      q->inflight--;
      s->stats.paths += pr.second;
      if (q->cancel && q->inflight == 0) dropped.push_back(q);
      else if (q->finished == q->total) done.push_back(q);
    }
  }
  for (Req* q : done) {
    const uint32_t npix = q->r.width * q->r.height;
    apt_status st = srv_resolve(q->d_accum, npix, q->r.spp, q->d_out, s->side);
    hipError_t e = hipSuccess;
    if (st == APT_OK) e = hipMemcpyAsync(q->out, q->d_out, (size_t)npix * 3 * sizeof(float),
                                         hipMemcpyDeviceToHost, s->side);
    if (st == APT_OK && e == hipSuccess) e = hipStreamSynchronize(s->side);
    std::lock_guard<std::mutex> lock(s->mu);
    q->state = (st == APT_OK && e == hipSuccess) ? APT_REQ_DONE : APT_REQ_FAILED;
    q->status = st != APT_OK ? st : (e == hipSuccess ? APT_OK : APT_ERR_DEVICE);
    q->t_done = s->now_ms();
    s->stats.requests_done++;
    dropped.push_back(q);
  }
  std::lock_guard<std::mutex> lock(s->mu);
  for (Req* q : dropped) {
    if (q->cancel && q->state != APT_REQ_DONE && q->state != APT_REQ_FAILED) {
      q->state = APT_REQ_CANCELLED;
      q->t_done = s->now_ms();
    }
    if (q->d_accum) s->accum_used -= arena_bytes(q);
    release_request_memory(s, q);
    if (q->slot >= 0) s->slots[q->slot] = nullptr;
    q->slot = -1;
  }
  if (!dropped.empty()) s->cv_done.notify_all();
  s->cv_work.notify_all();
}

// After a device failure: every queued and active request fails, and the
// active ones give back their slots and arena ranges (no batch that could
// still write them is awaited any more: the caller drops its in-flight list).
void fail_all(apt_server* s, apt_status st) {
  std::lock_guard<std::mutex> lock(s->mu);
  s->device_failed = true;
  for (auto& kv : s->all) {
    Req* q = kv.second;
    if (q->state == APT_REQ_QUEUED || q->state == APT_REQ_ACTIVE) {
      q->state = APT_REQ_FAILED;
      q->status = st;
      q->t_done = s->now_ms();
    }
  }
  for (Req*& q : s->slots) {
    if (!q) continue;
    if (q->d_accum) s->accum_used -= arena_bytes(q);
    release_request_memory(s, q);
    q->slot = -1;
    q->inflight = 0;
    q = nullptr;
  }
  s->queued.clear();
  s->cv_done.notify_all();
}

void worker_loop(apt_server* s) {
  (void)hipSetDevice(s->device);
  std::deque<Batch> inflight;
  int next_buf = 0;
  hipEvent_t ev_pool[4];
  for (int i = 0; i < 4; ++i) (void)hipEventCreate(&ev_pool[i]);
  int ev_next = 0;
  while (true) {
    Batch b;
    b.buf = next_buf;
    {
      std::unique_lock<std::mutex> lock(s->mu);
      // cancelled requests that never reached the GPU
      for (Req*& q : s->slots) {
        if (q && q->cancel && q->inflight == 0) {
          q->state = APT_REQ_CANCELLED;
          q->t_done = s->now_ms();
          if (q->d_accum) s->accum_used -= arena_bytes(q);
          release_request_memory(s, q);
          q->slot = -1;
          q = nullptr;
          s->cv_done.notify_all();
        }
      }
      if (s->stop && inflight.empty()) break;
      if (!s->device_failed) {
        admit(s);
        build_batch(s, b, *s->h_batch[b.buf]);
      }
      if (b.parts.empty() && inflight.empty()) {
        s->cv_work.wait_for(lock, std::chrono::milliseconds(2));
        continue;
      }
    }
    if (!b.parts.empty()) {
      b.t0 = ev_pool[ev_next];
      b.t1 = ev_pool[ev_next + 1];
      ev_next = (ev_next + 2) % 4;
      apt_status st = launch_batch(s, b);
      if (st != APT_OK) {
        (void)hipStreamSynchronize(s->stream);
        fail_all(s, st);
        inflight.clear();
        continue;
      }
      inflight.push_back(b);
      next_buf ^= 1;
    }
    if (inflight.size() >= 2 || (b.parts.empty() && !inflight.empty())) {
      Batch& f = inflight.front();
      hipError_t e = hipEventSynchronize(f.t1);
      if (e != hipSuccess) {
        set_error("server: batch failed: %s", hipGetErrorString(e));
        fail_all(s, APT_ERR_DEVICE);
        inflight.clear();
        continue;
      }
      finalize(s, f);
      inflight.pop_front();
    }
  }
  for (int i = 0; i < 4; ++i) (void)hipEventDestroy(ev_pool[i]);
}

template <typename T>
apt_status dmalloc(T*& p, size_t n) {
  hipError_t e = hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T));
  if (e != hipSuccess) {
    p = nullptr;
    set_error("server: hipMalloc of %zu B failed: %s", n * sizeof(T), hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? APT_ERR_OOM : APT_ERR_DEVICE;
  }
  return APT_OK;
}

apt_status create_device(apt_server* s) {
  const size_t P = s->cfg.path_slots;
  apt_status st;
  for (float4** b : {&s->p_o, &s->p_d, &s->p_t, &s->p_l, &s->p_h, &s->p_so, &s->p_sd, &s->p_sc})
    if ((st = dmalloc(*b, P)) != APT_OK) return st;
  for (uint32_t** b : {&s->p_rng, &s->p_q0, &s->p_q1, &s->p_qs, &s->path_req, &s->path_pix})
    if ((st = dmalloc(*b, P)) != APT_OK) return st;
  if ((st = dmalloc(s->arena, s->cfg.accum_budget)) != APT_OK) return st;
  s->arena_free[0] = s->cfg.accum_budget;
  if ((st = dmalloc(s->p_cnt, CNT_COUNT)) != APT_OK) return st;
  APT_HIP(hipMemset(s->p_cnt, 0, CNT_COUNT * sizeof(uint32_t)));
  if ((st = dmalloc(s->d_reqs, s->cfg.max_active)) != APT_OK) return st;
  APT_HIP(hipMemset(s->d_reqs, 0, s->cfg.max_active * sizeof(SrvReq)));
  APT_HIP(hipHostMalloc((void**)&s->h_reqs, s->cfg.max_active * sizeof(SrvReq), hipHostMallocDefault));
  std::memset(s->h_reqs, 0, s->cfg.max_active * sizeof(SrvReq));
  APT_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  APT_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
  for (int i = 0; i < 2; ++i) {
    if ((st = dmalloc(s->d_batch[i], 1)) != APT_OK) return st;
    APT_HIP(hipMemset(s->d_batch[i], 0, sizeof(SrvBatch)));
    APT_HIP(hipHostMalloc((void**)&s->h_batch[i], sizeof(SrvBatch), hipHostMallocDefault));
    std::memset(s->h_batch[i], 0, sizeof(SrvBatch));
    SrvDevice& d = s->dev[i];
    d.path_slots = s->cfg.path_slots;
    d.max_bounces = s->cfg.max_bounces;
    d.width = s->cfg.bvh_width;
    d.lds_stack = s->cfg.lds_stack_depth;
    d.persistent = s->cfg.persistent;
    d.batch = s->d_batch[i];
    d.reqs = s->d_reqs;
    d.ps.o_pdf = s->p_o; d.ps.d_flags = s->p_d; d.ps.thr = s->p_t; d.ps.rad = s->p_l;
    d.ps.hit = s->p_h; d.ps.sh_o = s->p_so; d.ps.sh_d = s->p_sd; d.ps.sh_c = s->p_sc;
    d.ps.rng = s->p_rng; d.ps.queue[0] = s->p_q0; d.ps.queue[1] = s->p_q1;
    d.ps.shadow_queue = s->p_qs; d.ps.counters = s->p_cnt;
    d.path_req = s->path_req;
    d.path_pix = s->path_pix;
    // Capture the batch once per staging buffer and depth; replays only read
    // device memory the host updates between batches.  If capture is
    // unavailable the same kernels are launched directly.
    s->graph[i].assign(s->cfg.max_bounces, nullptr);
    for (uint32_t depth = 1; depth <= s->cfg.max_bounces; ++depth) {
      SrvDevice dd = d;
      dd.max_bounces = depth;
      hipGraph_t g = nullptr;
      if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed) != hipSuccess) break;
      apt_status est = srv_enqueue_batch(dd, s->stream);
      hipError_t e = hipStreamEndCapture(s->stream, &g);
      if (est == APT_OK && e == hipSuccess && g) {
        if (hipGraphInstantiate(&s->graph[i][depth - 1], g, nullptr, nullptr, 0) != hipSuccess)
          s->graph[i][depth - 1] = nullptr;
      }
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
    }
  }
  APT_HIP(hipStreamSynchronize(s->stream));
  return APT_OK;
}

}  // namespace

apt_server::~apt_server() {
  {
    std::lock_guard<std::mutex> lock(mu);
    stop = true;
    for (auto& kv : all)
      if (kv.second->state == APT_REQ_QUEUED) {
        kv.second->state = APT_REQ_CANCELLED;
      }
    queued.clear();
    for (Req* q : slots)
      if (q) q->cancel = true;
  }
  cv_work.notify_all();
  if (worker.joinable()) worker.join();
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& kv : all) delete kv.second;
  if (arena) (void)hipFree(arena);
  for (SceneGpu& g : scenes) scene_gpu_free(g);
  for (int i = 0; i < 2; ++i) {
    for (hipGraphExec_t g : graph[i])
      if (g) (void)hipGraphExecDestroy(g);
    if (d_batch[i]) (void)hipFree(d_batch[i]);
    if (h_batch[i]) (void)hipHostFree(h_batch[i]);
  }
  for (void* p : {(void*)p_o, (void*)p_d, (void*)p_t, (void*)p_l, (void*)p_h, (void*)p_so, (void*)p_sd,
                  (void*)p_sc, (void*)p_rng, (void*)p_q0, (void*)p_q1, (void*)p_qs, (void*)p_cnt,
                  (void*)path_req, (void*)path_pix, (void*)d_reqs})
    if (p) (void)hipFree(p);
  if (h_reqs) (void)hipHostFree(h_reqs);
  if (stream) (void)hipStreamDestroy(stream);
  if (side) (void)hipStreamDestroy(side);
}

extern "C" {

apt_status apt_server_create(const apt_server_config* cfg, apt_server** out) {
  clear_error();
  APT_REQUIRE(out, APT_ERR_INVALID_ARG, "apt_server_create: out is NULL");
  apt_server_config c = cfg ? *cfg : apt_server_config{};
  if (!c.path_slots) c.path_slots = 1u << 24;  // 2.6 GB of path state
  if (!c.max_active) c.max_active = 64;
  if (!c.accum_budget) c.accum_budget = 4ull << 30;
  if (!c.max_bounces) c.max_bounces = 8;
  if (!c.tile_paths) c.tile_paths = 65536;
  if (!c.bvh_width) c.bvh_width = 4;
  if (!c.lds_stack_depth) c.lds_stack_depth = 8;
  APT_REQUIRE(c.path_slots >= 256 && c.path_slots <= (1u << 24), APT_ERR_INVALID_ARG,
              "server: path_slots %u outside 256..2^24", c.path_slots);
  APT_REQUIRE(c.max_active >= 1 && c.max_active <= 1024, APT_ERR_INVALID_ARG,
              "server: max_active %u outside 1..1024", c.max_active);
  APT_REQUIRE(c.max_bounces >= 1 && c.max_bounces <= 64, APT_ERR_INVALID_ARG,
              "server: max_bounces %u outside 1..64", c.max_bounces);
  APT_REQUIRE(c.bvh_width == 2 || c.bvh_width == 4 || c.bvh_width == 8, APT_ERR_INVALID_ARG,
              "server: bvh_width %u is not 2, 4 or 8", c.bvh_width);
  APT_REQUIRE(c.lds_stack_depth == 8 || c.lds_stack_depth == 12 || c.lds_stack_depth == 16,
              APT_ERR_INVALID_ARG, "server: lds_stack_depth %u is not 8, 12 or 16", c.lds_stack_depth);
  APT_REQUIRE(c.policy <= 2, APT_ERR_INVALID_ARG, "server: unknown policy %u", c.policy);
  APT_REQUIRE(c.persistent <= 1, APT_ERR_INVALID_ARG, "server: persistent must be 0 or 1, not %u", c.persistent);
  apt_server* s = new apt_server();
  s->cfg = c;
  (void)hipGetDevice(&s->device);
  s->slots.assign(c.max_active, nullptr);
  apt_status st = create_device(s);
  if (st != APT_OK) {
    delete s;
    return st;
  }
  s->worker = std::thread(worker_loop, s);
  *out = s;
  return APT_OK;
}

void apt_server_destroy(apt_server* s) { delete s; }

apt_status apt_server_add_scene(apt_server* s, const apt_mesh_desc* mesh, const apt_sky* sky,
                                uint32_t* scene_id) {
  clear_error();
  APT_REQUIRE(s && mesh && scene_id, APT_ERR_INVALID_ARG, "apt_server_add_scene: null argument");
  SceneGpu g;
  apt_status st = scene_gpu_upload(mesh->positions, mesh->normals, mesh->material_ids, mesh->n_tris,
                                   mesh->materials, mesh->n_materials, sky, s->cfg.bvh_width,
                                   s->cfg.lds_stack_depth, g);
  if (st != APT_OK) return st;
  std::lock_guard<std::mutex> lock(s->mu);
  s->scenes.push_back(g);
  *scene_id = (uint32_t)s->scenes.size() - 1;
  return APT_OK;
}

apt_status apt_submit(apt_server* s, const apt_request* req, float* out_rgb, uint64_t* id) {
  clear_error();
  APT_REQUIRE(s && req && out_rgb && id, APT_ERR_INVALID_ARG, "apt_submit: null argument");
  APT_REQUIRE(req->width > 0 && req->height > 0 && req->spp > 0, APT_ERR_INVALID_ARG,
              "apt_submit: empty request %ux%u x %u spp", req->width, req->height, req->spp);
  const uint64_t npix = (uint64_t)req->width * req->height;
  APT_REQUIRE(npix * req->spp < 0xFFFFFFFFull, APT_ERR_INVALID_ARG,
              "apt_submit: %llu paths exceed 2^32", (unsigned long long)(npix * req->spp));
  APT_REQUIRE(req->max_bounces >= 1 && req->max_bounces <= s->cfg.max_bounces, APT_ERR_INVALID_ARG,
              "apt_submit: max_bounces %u outside 1..%u", req->max_bounces, s->cfg.max_bounces);
  Req* q = new Req();
  q->r = *req;
  q->out = out_rgb;
  q->total = npix * req->spp;
  q->footprint = arena_bytes(q);
  {
    std::lock_guard<std::mutex> lock(s->mu);
    if (req->scene_id >= s->scenes.size()) {
      delete q;
      set_error("apt_submit: unknown scene %u", req->scene_id);
      return APT_ERR_NOT_FOUND;
    }
    if (s->device_failed) {
      delete q;
      set_error("apt_submit: the server's device failed");
      return APT_ERR_DEVICE;
    }
    if (q->footprint > s->cfg.accum_budget) {
      delete q;
      set_error("apt_submit: request footprint %zu B exceeds the budget of %llu B", q->footprint,
                (unsigned long long)s->cfg.accum_budget);
      return APT_ERR_OOM;
    }
    q->id = s->next_id++;
    q->t_submit = s->now_ms();
    s->all[q->id] = q;
    s->queued.push_back(q);
    *id = q->id;
  }
  s->cv_work.notify_all();
  return APT_OK;
}

static void fill_result(const Req* q, apt_result* out) {
  out->state = q->state;
  out->status = q->status;
  out->submit_ms = q->t_submit;
  out->admit_ms = q->t_admit;
  out->done_ms = q->t_done;
  out->paths = q->finished;
}

apt_status apt_poll(apt_server* s, uint64_t id, apt_result* out) {
  clear_error();
  APT_REQUIRE(s && out, APT_ERR_INVALID_ARG, "apt_poll: null argument");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->all.find(id);
  APT_REQUIRE(it != s->all.end(), APT_ERR_NOT_FOUND, "apt_poll: unknown request %llu",
              (unsigned long long)id);
  fill_result(it->second, out);
  return APT_OK;
}

apt_status apt_wait(apt_server* s, uint64_t id, uint32_t timeout_ms, apt_result* out) {
  clear_error();
  APT_REQUIRE(s, APT_ERR_INVALID_ARG, "apt_wait: null server");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->all.find(id);
  APT_REQUIRE(it != s->all.end(), APT_ERR_NOT_FOUND, "apt_wait: unknown request %llu",
              (unsigned long long)id);
  Req* q = it->second;
  const bool ok = s->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), [&] {
    return q->state >= APT_REQ_DONE;
  });
  if (out) fill_result(q, out);
  if (!ok) {
    set_error("apt_wait: request %llu still running", (unsigned long long)id);
    return APT_ERR_BUSY;
  }
  return APT_OK;
}

apt_status apt_release(apt_server* s, uint64_t id) {
  clear_error();
  APT_REQUIRE(s, APT_ERR_INVALID_ARG, "apt_release: null server");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->all.find(id);
  APT_REQUIRE(it != s->all.end(), APT_ERR_NOT_FOUND, "apt_release: unknown request %llu",
              (unsigned long long)id);
  Req* q = it->second;
  APT_REQUIRE(q->state >= APT_REQ_DONE && q->slot < 0 && q->inflight == 0, APT_ERR_BUSY,
              "apt_release: request %llu is still queued or running", (unsigned long long)id);
  delete q;
  s->all.erase(it);
  return APT_OK;
}

apt_status apt_cancel(apt_server* s, uint64_t id) {
  clear_error();
  APT_REQUIRE(s, APT_ERR_INVALID_ARG, "apt_cancel: null server");
  {
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->all.find(id);
    APT_REQUIRE(it != s->all.end(), APT_ERR_NOT_FOUND, "apt_cancel: unknown request %llu",
                (unsigned long long)id);
    Req* q = it->second;
    if (q->state == APT_REQ_QUEUED) {
      s->queued.erase(std::remove(s->queued.begin(), s->queued.end(), q), s->queued.end());
      q->state = APT_REQ_CANCELLED;
      q->t_done = s->now_ms();
      s->cv_done.notify_all();
    } else if (q->state == APT_REQ_ACTIVE) {
      q->cancel = true;
    }
  }
  s->cv_work.notify_all();
  return APT_OK;
}

apt_status apt_server_get_stats(apt_server* s, apt_server_stats* out) {
  APT_REQUIRE(s && out, APT_ERR_INVALID_ARG, "apt_server_get_stats: null argument");
  std::lock_guard<std::mutex> lock(s->mu);
  *out = s->stats;
  out->queued = (uint32_t)s->queued.size();
  uint32_t act = 0;
  for (Req* q : s->slots) act += q ? 1u : 0u;
  out->active = act;
  return APT_OK;
}

}  // extern "C"
