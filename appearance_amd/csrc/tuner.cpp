// Per-scene autotuner: picks the BVH width, the traversal stack depth kept in
// LDS and the wavefront path budget (queue chunk) for a scene by timing a probe
// frame of the real integrator on the device for every candidate.  The
// candidates only change how the same work is scheduled: every candidate must
// produce the identical image (closest hits do not depend on the BVH width, the
// film adds samples in sample order whatever the pass grouping), and the tuner
// checks that it does.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "apt_internal.h"

extern "C" {

apt_status apt_autotune(const apt_mesh_desc* mesh, const apt_sky* sky, const apt_camera* camera,
                        const apt_tune_opts* opts, apt_tune_result* out, char* json,
                        size_t json_len) {
  apt::clear_error();
  APT_REQUIRE(mesh && camera && out, APT_ERR_INVALID_ARG, "apt_autotune: null argument");
  apt_tune_opts o = opts ? *opts : apt_tune_opts{};
  if (!o.probe_width) o.probe_width = 512;
  if (!o.probe_height) o.probe_height = 256;
  if (!o.probe_spp) o.probe_spp = 2;
  if (!o.probe_bounces) o.probe_bounces = 4;
  if (!o.repeats) o.repeats = 2;
  const uint32_t widths[3] = {2, 4, 8};
  const uint32_t stacks[3] = {8, 12, 16};
  const uint32_t npix = o.probe_width * o.probe_height;
  // whole frame per pass (0 = default budget) or one sample per pass
  const uint32_t chunks[2] = {0, npix};
  std::string js = "{\"candidates\": [";
  std::vector<float> ref, img((size_t)npix * 3);
  double best = 1e30;
  apt_tune_result res{};
  uint32_t trials = 0, rejected = 0;
  for (uint32_t w : widths) {
    apt_config cfg;
    apt_config_default(&cfg);
    // the wavefront path tracer: the candidates schedule its passes
    cfg.integrator = APT_INTEGRATOR_PLAIN;
    cfg.restir_di = cfg.restir_gi = cfg.firefly_filter = cfg.svgf = cfg.taa = 0;
    cfg.bvh_width = w;
    cfg.sample_count = o.probe_spp;
    cfg.max_bounces = o.probe_bounces;
    apt_renderer* r = nullptr;
    apt_status st = apt_create(&cfg, &r);
    if (st != APT_OK) return st;
    st = apt_scene_set(r, mesh, sky);
    if (st == APT_OK) st = apt_world_action(r, 0, camera, sizeof(apt_camera));
    // The scene is built at the first render: a tree deeper than the
    // traversal stack at this width is refused there and is not a candidate.
    if (st == APT_OK) st = apt_render(r, o.probe_width, o.probe_height, 0, o.probe_height, nullptr, 0, img.data(),
                                      img.size());
    if (st == APT_ERR_UNSUPPORTED) {
      char buf[96];
      std::snprintf(buf, sizeof(buf), "%s{\"bvh_width\": %u, \"rejected\": \"traversal stack\"}",
                    trials || rejected ? ", " : "", w);
      js += buf;
      rejected++;
      apt_destroy(r);
      continue;
    }
    for (uint32_t s : stacks) {
      for (uint32_t q : chunks) {
        if (st != APT_OK) break;
        cfg.lds_stack_depth = s;
        cfg.queue_chunk = q;
        cfg.frame_idx = 0;
        st = apt_set_config(r, &cfg);
        double ms_best = 1e30;
        for (uint32_t k = 0; k < o.repeats + 1 && st == APT_OK; ++k) {
          // frame 0 of a fresh sequence every time, so every candidate renders
          // the same samples
          st = apt_set_frame_index(r, 0);
          if (st != APT_OK) break;
          st = apt_render(r, o.probe_width, o.probe_height, 0, o.probe_height, nullptr, 0, img.data(),
                          img.size());
          if (st != APT_OK) break;
          apt_render_stats rs;
          apt_get_render_stats(r, &rs);
          if (k > 0) ms_best = std::min(ms_best, rs.ms_total);  // first run warms up
        }
        if (st != APT_OK) break;
        if (ref.empty()) ref = img;
        const bool same = std::memcmp(ref.data(), img.data(), img.size() * sizeof(float)) == 0;
        char buf[160];
        std::snprintf(buf, sizeof(buf), "%s{\"bvh_width\": %u, \"lds_stack_depth\": %u, \"queue_chunk\": %u, \"ms\": %.4f, \"identical\": %s}",
                      trials || rejected ? ", " : "", w, s, q, ms_best, same ? "true" : "false");
        js += buf;
        trials++;
        if (!same) {
          size_t nd = 0, first = 0;
          for (size_t i = 0; i < img.size(); ++i) {
            if (std::memcmp(&ref[i], &img[i], sizeof(float)) != 0) {
              if (nd++ == 0) first = i;
            }
          }
          apt_destroy(r);
          apt::set_error("apt_autotune: candidate width %u stack %u chunk %u changed the image: %zu of %zu "
                         "values differ, first at %zu: %.9g vs %.9g",
                         w, s, q, nd, img.size(), first, (double)img[first], (double)ref[first]);
          return APT_ERR_DEVICE;
        }
        if (ms_best < best) {
          best = ms_best;
          res.bvh_width = w;
          res.lds_stack_depth = s;
          res.queue_chunk = q;
        }
      }
    }
    apt_destroy(r);
    if (st != APT_OK) return st;
  }
  APT_REQUIRE(trials > 0, APT_ERR_UNSUPPORTED,
              "apt_autotune: the scene's tree needs a deeper traversal stack than %u entries at every width",
              apt::trace_stack_capacity());
  res.best_ms = best;
  res.trials = trials;
  char tail[200];
  std::snprintf(tail, sizeof(tail), "], \"best\": {\"bvh_width\": %u, \"lds_stack_depth\": %u, \"queue_chunk\": %u, \"ms\": %.4f}}",
                res.bvh_width, res.lds_stack_depth, res.queue_chunk, res.best_ms);
  js += tail;
  if (json && json_len) {
    std::snprintf(json, json_len, "%s", js.c_str());
    APT_REQUIRE(js.size() < json_len, APT_ERR_INVALID_ARG, "apt_autotune: json buffer needs %zu bytes",
                js.size() + 1);
  }
  *out = res;
  return APT_OK;
}

}  // extern "C"
