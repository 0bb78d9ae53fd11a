// Device state of the ReSTIR integrator (reservoirs, G-buffer, path vertices,
// denoiser history), shared by renderer.cpp and restir.hip.
#pragma once

#include "integrator.h"

namespace apt {

constexpr uint32_t LIGHT_SUN = 0xFFFFFFFEu;    // LightSample::new_sun_sample (light_sample.wgsl:49)
constexpr uint32_t LIGHT_EMPTY = 0xFFFFFFFFu;  // LightSample::empty (light_sample.wgsl:53)

// G-buffer texel, 32 B (PackedGBufferTexel, gbuffer.wgsl:10-17, with FP16
// depth and normal): world position and depth in FP32, then half2(depth, nx),
// half2(ny, nz), albedo (rgb9e5), material id | back-face << 31.
// depth == 0 marks sky (PackedGBufferTexel::is_sky, gbuffer.wgsl:52-54).
struct GTexel {
  float4 pos_depth;
  uint4 packed;
};

// DI reservoir, 32 B (PackedDiReservoir, di_reservoir.wgsl:14-20): the light
// sample (emissive triangle + barycentrics, sun + cone coordinates, or empty),
// weight sum, candidate count M, contribution weight W and the target value of
// the selected sample.
struct DiRes {
  uint32_t light;
  float u, v;
  float wsum;
  float M;
  float W;
  float phat;
  float pad;
};

// GI reservoir, 32 B (PackedGiReservoir, gi_reservoir.wgsl:15-22): the sample
// point x2 and the RNG state that replays its target value, then M, W, weight
// sum and target value.
struct GiRes {
  float4 x2_rng;  // x2.xyz, phat_rng bits
  float4 mwsp;    // M, W, wsum, phat
};

struct CamView {
  v3 pos, right, up, back;
  float tan_half, aspect;
};

struct RestirParams {
  uint32_t width, rows, row_start, height;  // slab [row_start, row_start + rows) of width x height
  uint32_t seed;                            // per-sample seed (frame * spp + sample + base seed)
  uint32_t max_bounces;
  uint32_t di_candidates;                   // area candidates of the DI RIS (one BSDF candidate beside)
  uint32_t di_mcap;                         // temporal M cap, times the current M
  uint32_t neighbours;                      // spatial taps
  float radius;                             // spatial radius in pixels
  uint32_t gi_mcap;                         // temporal M cap (absolute)
  uint32_t gi_spatial_passes;               // GI spatial reuse passes (0: none)
  uint32_t restir_di, restir_gi;            // spatiotemporal reuse switches
  uint32_t cur;                             // parity of the frame: history buffers [cur] are written, [cur ^ 1] read
  uint32_t svgf_history;                    // max frames of SVGF moment history
  uint32_t atrous_passes;
  uint32_t first_frame;                     // no previous frame (TAA is skipped)
  CamView prev_cam;
};

// Per-pixel buffers of one slab.
struct RestirBuffers {
  // path state
  float4* ray_o;   // origin
  float4* ray_d;   // direction
  float4* thr;     // throughput, w = 1 while the path continues
  uint32_t* rng;
  float4* radiance;  // sum over the samples of this frame
  // vertex of the current bounce (LightSampleCtx, light_sample.wgsl:23-30)
  float4* vtx_p;   // hit point, t (< 0: no vertex)
  float4* vtx_n;   // front-facing shading normal, material | back << 31 (bits)
  float4* vtx_t;   // throughput at the vertex
  float4* vtx_m;   // textured scenes: uv and the clearcoat normal (oct30 bits) of the vertex
  // reservoirs
  DiRes* di;       // RIS output, then the reused reservoir that is shaded
  DiRes* di_tmp;   // temporal output
  DiRes* di_hist[2];
  float* di_self;       // target of each pixel's temporal DI sample at the pixel itself, before
                        // its shadow ray (queued in rq_*, traced before the spatial pass)
  GiRes* gi;
  GiRes* gi_tmp;   // the GI spatial passes' other buffer
  GiRes* gi_hist[2];
  GTexel* gbuf[2];
  // post chain
  float4* demod[2];     // demodulated radiance: [cur] this frame, [cur ^ 1] the previous final
  float4* work;         // scratch image (firefly / TAA output)
  float4* svgf_col[2];  // temporal / a-trous ping-pong: colour, variance in w
  float2* svgf_mom[2];
  uint2* svgf_guide;    // the a-trous edge stops of this frame's texels: the G-buffer's FP16
                        // depth and normal words, x = ~0u for sky
  uint32_t* svgf_count;
  float4* svgf_hist;    // a-trous pass-0 output kept for the next frame
  uint32_t* counters;   // closest-hit rays, then shadow rays traced: RS_CTR_SLOTS partial sums each
  // rays of one tracing stage, per pixel, traced by the persistent pool kernel:
  // origin | 1 when the pixel has a ray, direction | tmax; the result
  // (t, u, v, prim bits for a closest hit; w = 1 when occluded for any hit)
  float4* rq_o;
  float4* rq_d;
  float4* rq_hit;
  uint32_t* rq_ctr;     // [RQ_POOL ..] the pool's counters
};

constexpr uint32_t RS_CTR_SLOTS = 64;
constexpr uint32_t RQ_POOL = 0;
constexpr uint32_t RQ_WORDS = RQ_POOL + POOL_WORDS;

// One sample of the ReSTIR integrator over the slab: camera rays, bounces with
// DI/GI resampling and reuse at the first vertex; radiance accumulates into
// b.radiance.
apt_status restir_sample(const RestirParams& p, const SceneView& sc, const RestirBuffers& b,
                         hipStream_t s);
// restir_sample's four variants: two-level scene (sc.insts set) or not,
// textured scene (sc.tex set) or not, each compiled in a translation unit of
// its own (restir.hip, restir_two_level.hip, restir_textured.hip,
// restir_textured_two_level.hip) so they build in parallel.
template <bool INST, bool TEX>
apt_status restir_sample_variant(const RestirParams& p, const SceneView& sc, const RestirBuffers& b,
                                 hipStream_t s);
// Post chain of the frame: demodulation, firefly filter, SVGF, TAA,
// remodulation; adds the result to frame_sum.
apt_status restir_post(const RestirParams& p, const RestirBuffers& b, bool firefly, bool svgf,
                       bool taa, float4* frame_sum, hipStream_t s);

}  // namespace apt
