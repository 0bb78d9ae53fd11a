// ORACLE - test infrastructure only (see oracle_core.cpp).  C entry points
// loaded by tests/ and bench.py through ctypes.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint32_t or_pcg_hash(uint32_t x);
uint32_t or_xor_shift(uint32_t x);
uint32_t or_hash_combine(uint32_t x, uint32_t y);
uint32_t or_wangh_hash(uint32_t x);
void or_uniform_floats(uint32_t state, uint32_t n, float* out);
void or_rgb9e5_pack(const float* rgb, uint32_t n, uint32_t* out);
void or_rgb9e5_unpack(const uint32_t* d, uint32_t n, float* rgb);
void or_oct30_pack(const float* v, uint32_t n, uint32_t* out);
void or_oct30_unpack(const uint32_t* p, uint32_t n, float* v);
int or_tri_test(const float* ray, const float* tri, float* out_tuv);
void* or_bvh_build(const float* tris, uint32_t n);
void or_bvh_destroy(void* b);
uint32_t or_bvh_node_count(void* b);
void or_trace(void* b, const float* rays, uint32_t n, int any, float* out, uint64_t* counts,
              int threads);
void or_brute(const float* tris, uint32_t ntris, const float* rays, uint32_t n, int any,
              float* out);
void* or_inst_create(const float* tris, const uint32_t* blas_first, uint32_t n_blas);
void or_inst_destroy(void* sc);
void or_inst_trace(void* sc, const void* instances, uint32_t n_inst, const float* rays, uint32_t n, int any,
                   int brute, float* out, int threads);
int or_affine_inverse(const float* o2w, float* w2o);
uint32_t or_block_index(uint32_t x, uint32_t y, uint32_t width);
float or_aces(float x);
int or_render(const float* pos, const float* nrm, const uint32_t* mat_ids, const void* mats,
              uint32_t n_tris, void* bvh, const void* sky, const void* cam, const uint32_t* cfg10,
              const uint32_t* pixels, uint32_t n_pixels, int precision, int threads, float* out,
              uint64_t* counts, const void* attribs);
int or_render_counted(const float* pos, const float* nrm, const uint32_t* mat_ids, const void* mats,
                      uint32_t n_tris, void* bvh, const void* sky, const void* cam, const uint32_t* cfg10,
                      const uint32_t* pixels, uint32_t n_pixels, int precision, int threads, float* out,
                      uint64_t* counts, uint32_t n_counts, const void* attribs);
void* or_restir_create(uint32_t width, uint32_t rows);
void or_restir_destroy(void* st);
int or_restir_frame(void* st, const float* pos, const float* nrm, const uint32_t* mat_ids, const void* mats,
                    uint32_t n_tris, void* bvh, const void* sky, const void* cam, const uint32_t* cfg20,
                    int threads, float* out, uint64_t* rays, const void* attribs);
void or_set_plant(int kind);
int or_reservoir_probe(uint32_t kind, const float* a, const float* b, const float* extra, uint32_t* rng,
                       float* out);
void or_f16_round(const float* in, uint32_t n, float* out);
void or_atrous_probe(uint32_t width, uint32_t height, uint32_t pass, const float* col, const float* var,
                     const float* depth, const float* nrm, float* out);
int or_post_probe(uint32_t kind, uint32_t width, uint32_t height, uint32_t frames, const float* imgs,
                  const uint8_t* sky, uint32_t svgf_history, float* out, float* out_var);
double or_disney_eval(const void* mat, const double* n, const double* wo, const double* wi, int back,
                      double* f, const double* ncc);
double or_disney_sample(const void* mat, const double* n, const double* wo, int back, double r0, double r1,
                        double r2, double* wi, double* f, const double* ncc);
void or_tex_sample(const void* tex, float u, float v, float* out4);
void or_sky_probe(const void* sky, const void* attribs, const float* d, float* out);
void or_shading_probe(const void* mat, const void* attribs, uint32_t prim, float bu, float bv, const float* n,
                      float* out);
size_t or_jpeg_block(const uint8_t* rgbx, int quality, uint8_t* out, size_t cap);

#ifdef __cplusplus
}
#endif
