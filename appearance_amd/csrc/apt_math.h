// Device/host maths shared by every kernel of libapt: vectors, the hash-chain
// RNG, the HDR / unit-vector packing codecs, the watertight ray-triangle test
// and the padded slab test.  Everything here is __host__ __device__ so the host
// side of the library (BVH build, scene set-up) and the gfx950 kernels run the
// same arithmetic.  All sources are compiled with -ffp-contract=off: no fused
// multiply-add is formed unless written as fmaf().
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#define APT_HD __host__ __device__ __forceinline__

namespace apt {

constexpr float PI = 3.14159265358979323846f;
constexpr float INV_PI = 1.0f / PI;
constexpr float TWO_PI = 2.0f * PI;
constexpr uint32_t INVALID = 0xFFFFFFFFu;

struct v3 {
  float x, y, z;
};

APT_HD v3 mk(float x, float y, float z) { return v3{x, y, z}; }
APT_HD v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
APT_HD v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
APT_HD v3 operator-(v3 a) { return {-a.x, -a.y, -a.z}; }
APT_HD v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
APT_HD v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
APT_HD v3 operator*(float s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
APT_HD v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
APT_HD v3& operator+=(v3& a, v3 b) { a = a + b; return a; }
APT_HD v3& operator*=(v3& a, v3 b) { a = a * b; return a; }
APT_HD v3& operator*=(v3& a, float s) { a = a * s; return a; }
APT_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
APT_HD v3 cross(v3 a, v3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
APT_HD float length(v3 a) { return sqrtf(dot(a, a)); }
APT_HD v3 normalize(v3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
APT_HD float comp(v3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
APT_HD float maxc(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
APT_HD float sqr(float x) { return x * x; }
APT_HD v3 reflect(v3 i, v3 n) { return i - n * (2.0f * dot(n, i)); }
APT_HD float lerpf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
APT_HD float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
APT_HD float luma(v3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }

// ---------------------------------------------------------------------------
// RNG: the integer hash chain of assets/shaders/random.wgsl:1-68.
// ---------------------------------------------------------------------------
APT_HD uint32_t pcg_hash(uint32_t x) {  // random.wgsl:10-14
  uint32_t state = x * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
APT_HD uint32_t xor_shift_u32(uint32_t s) {  // random.wgsl:49-54
  s ^= s << 13;
  s ^= s >> 17;
  s ^= s << 5;
  return s;
}
APT_HD uint32_t hash_combine(uint32_t x, uint32_t y) {  // random.wgsl:57-68
  uint32_t seed = (x * 1664525u + y + 1013904223u) * 1664525u;
  seed ^= seed >> 11u;
  seed ^= (seed << 7u) & 0x9d2c5680u;
  seed ^= (seed << 15u) & 0xefc60000u;
  seed ^= seed >> 18u;
  return seed;
}
// random.wgsl:16-19: f32(state) / f32(0xFFFFFFFF).  f32(0xFFFFFFFF) rounds to
// 2^32, and f32(state) rounds to nearest, so the result lies in [0, 1] with 1.0
// reachable for states within 128 of 2^32.
APT_HD float rng_next(uint32_t& state) {
  state = pcg_hash(state);
  return (float)state / 4294967296.0f;
}
// Per-path seed: pcg_hash(pixel ^ xor_shift_u32(frame_seed)) as trace.wgsl:136
// does with seed = frame*spp + sample (appearance-path-tracer-gpu/src/lib.rs:374).
APT_HD uint32_t path_seed(uint32_t pixel, uint32_t sample_seed) {
  return pcg_hash(pixel ^ xor_shift_u32(sample_seed));
}

// ---------------------------------------------------------------------------
// Packing codecs (appearance-packing/assets/shaders/shared/packing.wgsl).
// ---------------------------------------------------------------------------
APT_HD uint32_t f2u(float f) {
#ifdef __HIP_DEVICE_COMPILE__
  return __float_as_uint(f);
#else
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
#endif
}
APT_HD float u2f(uint32_t u) {
#ifdef __HIP_DEVICE_COMPILE__
  return __uint_as_float(u);
#else
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
#endif
}

// Shared-exponent RGB (packing.wgsl:55-67).
APT_HD uint32_t pack_rgb9e5(v3 rgb) {
  const float max_val = u2f(0x477F8000u);
  const float min_val = u2f(0x37800000u);
  float r = fminf(fmaxf(rgb.x, 0.0f), max_val);
  float g = fminf(fmaxf(rgb.y, 0.0f), max_val);
  float b = fminf(fmaxf(rgb.z, 0.0f), max_val);
  float max_channel = fmaxf(fmaxf(min_val, r), fmaxf(g, b));
  float bias = u2f((f2u(max_channel) + 0x07804000u) & 0x7F800000u);
  uint32_t ri = f2u(r + bias), gi = f2u(g + bias), bi = f2u(b + bias);
  uint32_t e = (f2u(bias) << 4u) + 0x10000000u;
  return e | bi << 18 | gi << 9 | (ri & 0x1FFu);
}
// packing.wgsl:69-76: ldexp(mantissa, exponent - 24).
APT_HD v3 unpack_rgb9e5(uint32_t d) {
  int e = (int)(d >> 27) - 24;
  float scale = u2f((uint32_t)(e + 127) << 23);  // 2^e, e in [-24, 7]
  return {(float)(d & 0x1FFu) * scale, (float)((d >> 9) & 0x1FFu) * scale,
          (float)((d >> 18) & 0x1FFu) * scale};
}

// Octahedral unit vector in 2 x 15 bit (packing.wgsl:11-50).  WGSL round() is
// ties-to-even, so rintf is used.
APT_HD uint32_t pack_oct30(v3 d) {
  float s = fabsf(d.x) + fabsf(d.y) + fabsf(d.z);
  float ex = d.x / s, ey = d.y / s;
  if (d.z < 0.0f) {
    float sx = ex >= 0.0f ? 1.0f : -1.0f;
    float sy = ey >= 0.0f ? 1.0f : -1.0f;
    float nx = (1.0f - fabsf(ey)) * sx;
    float ny = (1.0f - fabsf(ex)) * sy;
    ex = nx;
    ey = ny;
  }
  ex = ex * 0.5f + 0.5f;
  ey = ey * 0.5f + 0.5f;
  return ((uint32_t)rintf(ey * 32767.0f) << 15) | (uint32_t)rintf(ex * 32767.0f);
}
APT_HD v3 unpack_oct30(uint32_t p) {
  float ex = (float)(p & 0x7fffu) / 32767.0f;
  float ey = (float)((p >> 15) & 0x7fffu) / 32767.0f;
  ex = ex * 2.0f - 1.0f;
  ey = ey * 2.0f - 1.0f;
  v3 n = {ex, ey, 1.0f - fabsf(ex) - fabsf(ey)};
  float t = clampf(-n.z, 0.0f, 1.0f);
  n.x += n.x >= 0.0f ? -t : t;
  n.y += n.y >= 0.0f ? -t : t;
  return normalize(n);
}

// ---------------------------------------------------------------------------
// Watertight ray-triangle test (Woop, Benthin, Wald, JCGT 2013).
// Hit contract shared with the oracle: t = T * (1/det), u = V * (1/det) is the
// weight of v1, v = W * (1/det) the weight of v2 (the reference's
// intersection.barycentrics, trace.wgsl:168).  A hit needs tmin < t < tmax.
// ---------------------------------------------------------------------------
struct RayPre {
  v3 o;
  int kz;  // dominant axis of d; the other two follow it cyclically
  float Sx, Sy, Sz;
  float tmin, tmax;
};

// Components of a in the ray's frame: (a[kx], a[ky], a[kz]) with
// kx = kz+1, ky = kz+2 (mod 3).
APT_HD v3 ray_frame(v3 a, int kz) {
  const bool z0 = kz == 0, z1 = kz == 1;
  return mk(z0 ? a.y : (z1 ? a.z : a.x), z0 ? a.z : (z1 ? a.x : a.y), z0 ? a.x : (z1 ? a.y : a.z));
}

// Woop et al. swap kx and ky when d[kz] < 0 to keep the triangles' winding.
// That swap negates U, V, W, det and T exactly (every product and sum changes
// sign, and round-to-nearest is symmetric), so t = T/det, u and v are the
// same without it; it is left out, which keeps kx and ky implied by kz.
APT_HD RayPre ray_pre(v3 o, v3 d, float tmin, float tmax) {
  RayPre r;
  r.o = o;
  float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
  const v3 f = ray_frame(d, kz);
  r.kz = kz;
  r.Sx = f.x / f.z;
  r.Sy = f.y / f.z;
  r.Sz = 1.0f / f.z;
  r.tmin = tmin;
  r.tmax = tmax;
  return r;
}

struct TriHit {
  float t, u, v;
};

APT_HD bool intersect_tri(const RayPre& r, v3 p0, v3 p1, v3 p2, TriHit& h) {
  const v3 A = ray_frame(p0 - r.o, r.kz), B = ray_frame(p1 - r.o, r.kz), C = ray_frame(p2 - r.o, r.kz);
  const float Akz = A.z, Bkz = B.z, Ckz = C.z;
  float Ax = A.x - r.Sx * Akz;
  float Ay = A.y - r.Sy * Akz;
  float Bx = B.x - r.Sx * Bkz;
  float By = B.y - r.Sy * Bkz;
  float Cx = C.x - r.Sx * Ckz;
  float Cy = C.y - r.Sy * Ckz;
  float U = Cx * By - Cy * Bx;
  float V = Ax * Cy - Ay * Cx;
  float W = Bx * Ay - By * Ax;
  if (U == 0.0f || V == 0.0f || W == 0.0f) {
    // Exact-zero edge functions are re-evaluated in double (Woop 2013, sec. 3).
    U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
    V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
    W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
  }
  if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
  float det = U + V + W;
  if (det == 0.0f) return false;
  float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
  float T = U * Az + V * Bz + W * Cz;
  float rcp = 1.0f / det;
  float t = T * rcp;
  if (!(t > r.tmin && t < r.tmax)) return false;
  h.t = t;
  h.u = V * rcp;
  h.v = W * rcp;
  return true;
}

// Lexicographic (t, prim) order used for closest hits: equal t resolves to the
// lower primitive index, so the answer does not depend on traversal order.
APT_HD bool hit_less(float t, uint32_t prim, float tb, uint32_t pb) {
  return t < tb || (t == tb && prim < pb);
}

// Slab test against a padded box.  Boxes are widened at build time by
// box_pad(), so a ray that lies on a face, and a triangle hit whose rounded t
// sits a few ulps outside the exact slab interval, still enter the node.
APT_HD float safe_rcp(float d) {
  return fabsf(d) < 1e-30f ? (d < 0.0f ? -1e30f : 1e30f) : 1.0f / d;
}

}  // namespace apt
