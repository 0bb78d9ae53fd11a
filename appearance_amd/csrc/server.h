// Device-side tables of the serving loop, shared by server.cpp and server.hip.
#pragma once

#include "integrator.h"

namespace apt {

// One admitted request as the kernels see it: its scene (with this request's
// camera folded in), frame, depth, seed and accumulator.
struct SrvReq {
  SceneView sc;
  uint32_t width, height, spp, max_bounces, rr_start, seed;
  unsigned long long* accum;  // 3 per pixel, fixed point 2^-24
};

// A run of a request's paths (request path index = sample * npix + pixel).
struct SrvTile {
  uint32_t req_slot;
  uint32_t first;   // first request path index
  uint32_t count;
  uint32_t offset;  // first batch path of this tile
};

constexpr uint32_t SRV_MAX_TILES = 4096;

struct SrvBatch {
  uint32_t n_paths;
  uint32_t n_tiles;
  uint32_t pad[2];
  SrvTile tiles[SRV_MAX_TILES];
};

// Everything a captured batch graph reads; pointers never change after set-up.
struct SrvDevice {
  uint32_t path_slots;
  uint32_t max_bounces;
  uint32_t width;      // BVH width of every resident scene
  uint32_t lds_stack;
  const SrvBatch* batch;
  const SrvReq* reqs;
  PathView ps;
  uint32_t* path_req;
  uint32_t* path_pix;
  uint32_t persistent;  // 1: one persistent kernel per batch, 0: wavefront stages
};

apt_status srv_enqueue_batch(const SrvDevice& d, hipStream_t s);
apt_status srv_resolve(const unsigned long long* accum, uint32_t npix, uint32_t spp, float* out,
                       hipStream_t s);
apt_status srv_zero(unsigned long long* a, size_t n, hipStream_t s);

}  // namespace apt
