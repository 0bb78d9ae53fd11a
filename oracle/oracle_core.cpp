// ORACLE - test infrastructure only.  Nothing in the product (libapt) links,
// loads or calls this code; it is used by tests/, by the smoke check and by
// the CPU-baseline leg of bench.py as the independent CPU restatement that
// the HIP path is compared against.
//
// This file: integer RNG, packing codecs, the watertight ray-triangle test,
// the canonical binary BVH (binned SAH, 32 bins, leaf <= 4) with visit
// counters, brute force, and the output layout (ACES fit + 64x64 blocks).
//
// Parity pins (see tests/test_oracle_*.py): closed-form RNG values checked
// against an independent numpy uint32 evaluation, codec round-trip and
// special-case bounds, FP64 brute force and edge-sharing (watertightness) known
// answers for the triangle test, brute force for the BVH, an exact
// permutation test for the block layout.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"

namespace ors {
extern int g_plant;  // planted mistakes, oracle_restir.cpp (tests only)
}

namespace orc {

// ---- RNG: assets/shaders/random.wgsl:1-68 ---------------------------------
uint32_t pcg_hash(uint32_t x) {  // random.wgsl:10-14
  uint32_t st = x * 747796405u + 2891336453u;
  uint32_t w = ((st >> ((st >> 28u) + 4u)) ^ st) * 277803737u;
  return (w >> 22u) ^ w;
}
uint32_t xor_shift(uint32_t x) {  // random.wgsl:49-54
  uint32_t s = x ^ (x << 13);
  s ^= s >> 17;
  s ^= s << 5;
  return s;
}
uint32_t hash_combine(uint32_t x, uint32_t y) {  // random.wgsl:57-68 (Kajiya, Matsumoto tempering)
  uint32_t v = (x * 1664525u + y + 1013904223u) * 1664525u;
  v ^= (v >> 11u);
  v ^= (v << 7u) & 0x9d2c5680u;
  v ^= (v << 15u) & 0xefc60000u;
  v ^= (v >> 18u);
  return v;
}
uint32_t wangh_hash(uint32_t x) {  // random.wgsl:1-8
  uint32_t s = (x ^ 61u) ^ (x >> 16);
  s *= 9u;
  s ^= s >> 4;
  s *= 0x27d4eb2du;
  s ^= s >> 15;
  return s;
}
float uniform(uint32_t& st) {  // random.wgsl:16-19; f32(0xFFFFFFFF) == 2^32
  st = pcg_hash(st);
  return (float)st / (float)4294967295u;
}

// ---- packing.wgsl ---------------------------------------------------------
static inline uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static inline float fbits(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

uint32_t rgb9e5_pack(const float* rgb) {  // packing.wgsl:55-67 (MiniEngine)
  const float hi = fbits(0x477F8000u), lo = fbits(0x37800000u);
  float c[3];
  for (int k = 0; k < 3; ++k) c[k] = std::min(std::max(rgb[k], 0.0f), hi);
  float mx = std::max(std::max(lo, c[0]), std::max(c[1], c[2]));
  float bias = fbits((bits(mx) + 0x07804000u) & 0x7F800000u);
  uint32_t r = bits(c[0] + bias), g = bits(c[1] + bias), b = bits(c[2] + bias);
  uint32_t e = (bits(bias) << 4u) + 0x10000000u;
  return e | (b << 18) | (g << 9) | (r & 0x1FFu);
}
void rgb9e5_unpack(uint32_t d, float* rgb) {  // packing.wgsl:69-76
  int e = (int)(d >> 27) - 24;
  rgb[0] = std::ldexp((float)(d & 0x1FFu), e);
  rgb[1] = std::ldexp((float)((d >> 9) & 0x1FFu), e);
  rgb[2] = std::ldexp((float)((d >> 18) & 0x1FFu), e);
}
uint32_t oct30_pack(const float* n) {  // packing.wgsl:11-39; round() is ties-to-even in WGSL
  float s = std::fabs(n[0]) + std::fabs(n[1]) + std::fabs(n[2]);
  float x = n[0] / s, y = n[1] / s;
  if (n[2] < 0.0f) {
    float ox = x, oy = y;
    x = (1.0f - std::fabs(oy)) * (ox >= 0.0f ? 1.0f : -1.0f);
    y = (1.0f - std::fabs(ox)) * (oy >= 0.0f ? 1.0f : -1.0f);
  }
  x = x * 0.5f + 0.5f;
  y = y * 0.5f + 0.5f;
  uint32_t qx = (uint32_t)std::nearbyint(x * 32767.0f);
  uint32_t qy = (uint32_t)std::nearbyint(y * 32767.0f);
  return (qy << 15) | qx;
}
void oct30_unpack(uint32_t p, float* n) {  // packing.wgsl:41-50
  float x = (float)(p & 0x7fffu) / 32767.0f * 2.0f - 1.0f;
  float y = (float)((p >> 15) & 0x7fffu) / 32767.0f * 2.0f - 1.0f;
  float z = 1.0f - std::fabs(x) - std::fabs(y);
  float t = std::min(std::max(-z, 0.0f), 1.0f);
  x += (x >= 0.0f) ? -t : t;
  y += (y >= 0.0f) ? -t : t;
  float l = std::sqrt(x * x + y * y + z * z);
  n[0] = x * (1.0f / l);
  n[1] = y * (1.0f / l);
  n[2] = z * (1.0f / l);
}

// ---- watertight ray-triangle (Woop, Benthin, Wald 2013) -------------------
struct Shear {
  float o[3];
  int ax[3];  // kx, ky, kz
  float S[3]; // Sx, Sy, Sz
  float tmin, tmax;
};

Shear shear_of(const float* o, const float* d, float tmin, float tmax) {
  Shear s;
  int kz = 0;
  float a0 = std::fabs(d[0]), a1 = std::fabs(d[1]), a2 = std::fabs(d[2]);
  if (a0 > a1) kz = a0 > a2 ? 0 : 2;
  else kz = a1 > a2 ? 1 : 2;
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (d[kz] < 0.0f) std::swap(kx, ky);
  s.ax[0] = kx;
  s.ax[1] = ky;
  s.ax[2] = kz;
  s.S[0] = d[kx] / d[kz];
  s.S[1] = d[ky] / d[kz];
  s.S[2] = 1.0f / d[kz];
  for (int k = 0; k < 3; ++k) s.o[k] = o[k];
  s.tmin = tmin;
  s.tmax = tmax;
  return s;
}

// Returns true and (t, u, v) for tmin < t < tmax.  u, v weight vertices 1, 2.
bool tri_test(const Shear& s, const float* v, float& t, float& u, float& w2) {
  float rel[3][3];  // vertex j relative to the origin, axis k
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) rel[j][k] = v[3 * j + k] - s.o[k];
  const int kx = s.ax[0], ky = s.ax[1], kz = s.ax[2];
  float X[3], Y[3];
  for (int j = 0; j < 3; ++j) {
    X[j] = rel[j][kx] - s.S[0] * rel[j][kz];
    Y[j] = rel[j][ky] - s.S[1] * rel[j][kz];
  }
  // Edge functions: e0 opposite vertex 0 (weight of v0), e1 of v1, e2 of v2.
  float e0 = X[2] * Y[1] - Y[2] * X[1];
  float e1 = X[0] * Y[2] - Y[0] * X[2];
  float e2 = X[1] * Y[0] - Y[1] * X[0];
  if (e0 == 0.0f || e1 == 0.0f || e2 == 0.0f) {
    e0 = (float)((double)X[2] * (double)Y[1] - (double)Y[2] * (double)X[1]);
    e1 = (float)((double)X[0] * (double)Y[2] - (double)Y[0] * (double)X[2]);
    e2 = (float)((double)X[1] * (double)Y[0] - (double)Y[1] * (double)X[0]);
  }
  bool any_neg = e0 < 0.0f || e1 < 0.0f || e2 < 0.0f;
  bool any_pos = e0 > 0.0f || e1 > 0.0f || e2 > 0.0f;
  if (any_neg && any_pos) return false;
  float det = e0 + e1 + e2;
  if (det == 0.0f) return false;
  float Z0 = s.S[2] * rel[0][kz], Z1 = s.S[2] * rel[1][kz], Z2 = s.S[2] * rel[2][kz];
  float T = e0 * Z0 + e1 * Z1 + e2 * Z2;
  float inv = 1.0f / det;
  float tt = T * inv;
  if (!(tt > s.tmin && tt < s.tmax)) return false;
  t = tt;
  u = e1 * inv;
  w2 = e2 * inv;
  return true;
}

static inline bool before(float t, uint32_t p, float tb, uint32_t pb) {
  return t < tb || (t == tb && p < pb);
}

// ---- canonical BVH2 -------------------------------------------------------
struct BNode {  // 32 B: box + (left child | first tri), count (0 = internal)
  float lo[3], hi[3];
  uint32_t a, count;
};

struct Bvh {
  std::vector<BNode> nodes;
  std::vector<uint32_t> idx;
  std::vector<float> tris;  // original order, 9 floats each
};

static void build_rec(Bvh& b, const std::vector<float>& cen, const std::vector<float>& blo,
                      const std::vector<float>& bhi, uint32_t ni, uint32_t begin, uint32_t end) {
  const int NB = 32;
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t i = begin; i < end; ++i) {
    uint32_t p = b.idx[i];
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], blo[3 * p + k]);
      hi[k] = std::max(hi[k], bhi[3 * p + k]);
      clo[k] = std::min(clo[k], cen[3 * p + k]);
      chi[k] = std::max(chi[k], cen[3 * p + k]);
    }
  }
  for (int k = 0; k < 3; ++k) {
    b.nodes[ni].lo[k] = lo[k];
    b.nodes[ni].hi[k] = hi[k];
  }
  uint32_t n = end - begin;
  if (n <= 4) {
    b.nodes[ni].a = begin;
    b.nodes[ni].count = n;
    return;
  }
  auto area = [](const float* l, const float* h) {
    float dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
    return (dx < 0 || dy < 0 || dz < 0) ? 0.0f : 2.0f * (dx * dy + dy * dz + dz * dx);
  };
  float best = INFINITY;
  int bax = -1, bsplit = 0;
  for (int ax = 0; ax < 3; ++ax) {
    float ext = chi[ax] - clo[ax];
    if (!(ext > 0.0f)) continue;
    float bl[NB][3], bh[NB][3];
    uint32_t cnt[NB] = {0};
    for (int q = 0; q < NB; ++q)
      for (int k = 0; k < 3; ++k) {
        bl[q][k] = INFINITY;
        bh[q][k] = -INFINITY;
      }
    for (uint32_t i = begin; i < end; ++i) {
      uint32_t p = b.idx[i];
      int q = std::min(NB - 1, std::max(0, (int)((cen[3 * p + ax] - clo[ax]) * ((float)NB / ext))));
      cnt[q]++;
      for (int k = 0; k < 3; ++k) {
        bl[q][k] = std::min(bl[q][k], blo[3 * p + k]);
        bh[q][k] = std::max(bh[q][k], bhi[3 * p + k]);
      }
    }
    for (int s = 1; s < NB; ++s) {
      float l0[3] = {INFINITY, INFINITY, INFINITY}, h0[3] = {-INFINITY, -INFINITY, -INFINITY};
      float l1[3] = {INFINITY, INFINITY, INFINITY}, h1[3] = {-INFINITY, -INFINITY, -INFINITY};
      uint32_t c0 = 0, c1 = 0;
      for (int q = 0; q < NB; ++q) {
        float* L = q < s ? l0 : l1;
        float* H = q < s ? h0 : h1;
        (q < s ? c0 : c1) += cnt[q];
        for (int k = 0; k < 3; ++k) {
          L[k] = std::min(L[k], bl[q][k]);
          H[k] = std::max(H[k], bh[q][k]);
        }
      }
      if (!c0 || !c1) continue;
      float c = area(l0, h0) * c0 + area(l1, h1) * c1;
      if (c < best) {
        best = c;
        bax = ax;
        bsplit = s;
      }
    }
  }
  uint32_t mid;
  if (bax < 0) {
    std::sort(b.idx.begin() + begin, b.idx.begin() + end);
    mid = begin + n / 2;
  } else {
    float ext = chi[bax] - clo[bax];
    auto it = std::partition(b.idx.begin() + begin, b.idx.begin() + end, [&](uint32_t p) {
      int q = std::min(NB - 1, std::max(0, (int)((cen[3 * p + bax] - clo[bax]) * ((float)NB / ext))));
      return q < bsplit;
    });
    mid = (uint32_t)(it - b.idx.begin());
  }
  uint32_t l = (uint32_t)b.nodes.size();
  b.nodes.push_back(BNode{});
  b.nodes.push_back(BNode{});
  b.nodes[ni].a = l;
  b.nodes[ni].count = 0;
  build_rec(b, cen, blo, bhi, l, begin, mid);
  build_rec(b, cen, blo, bhi, l + 1, mid, end);
}

Bvh* build(const float* tris, uint32_t n) {
  Bvh* b = new Bvh();
  b->tris.assign(tris, tris + (size_t)n * 9);
  std::vector<float> cen(3 * (size_t)n), blo(3 * (size_t)n), bhi(3 * (size_t)n);
  float m = 0.0f;
  for (uint32_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      float a = tris[9 * (size_t)i + k], c = tris[9 * (size_t)i + 3 + k], d = tris[9 * (size_t)i + 6 + k];
      blo[3 * i + k] = std::min(a, std::min(c, d));
      bhi[3 * i + k] = std::max(a, std::max(c, d));
      cen[3 * i + k] = 0.5f * (blo[3 * i + k] + bhi[3 * i + k]);
      m = std::max(m, std::max(std::fabs(a), std::max(std::fabs(c), std::fabs(d))));
    }
  }
  b->idx.resize(n);
  for (uint32_t i = 0; i < n; ++i) b->idx[i] = i;
  b->nodes.reserve(2 * (size_t)n);
  b->nodes.push_back(BNode{});
  build_rec(*b, cen, blo, bhi, 0, 0, n);
  // Conservative padding: 2^-13 of the largest coordinate magnitude.
  float pad = std::max(m / 8192.0f, 1e-30f);
  for (BNode& nd : b->nodes)
    for (int k = 0; k < 3; ++k) {
      nd.lo[k] -= pad;
      nd.hi[k] += pad;
    }
  return b;
}

// With WIDE (the object-space rays of instances) the box's entry and exit
// distances move out by 2^-18 of their magnitude before the clamps, so a box
// far from the ray origin is still entered when rounding would just miss it.
template <bool WIDE = false>
static inline bool box_enter(const BNode& nd, const float* o, const float* inv, float tmin,
                             float tcap, float& tnear) {
  float bn = -INFINITY, bf = INFINITY;
  for (int k = 0; k < 3; ++k) {
    float a = (nd.lo[k] - o[k]) * inv[k], c = (nd.hi[k] - o[k]) * inv[k];
    bn = std::max(bn, std::min(a, c));
    bf = std::min(bf, std::max(a, c));
  }
  if (WIDE) {
    bn -= std::fabs(bn) * 0x1p-18f;
    bf += std::fabs(bf) * 0x1p-18f;
  }
  const float tn = std::max(tmin, bn), tf = std::min(tcap, bf);
  tnear = tn;
  return tn <= tf;
}

static inline float rcp_safe(float d) {
  return std::fabs(d) < 1e-30f ? (d < 0.0f ? -1e30f : 1e30f) : 1.0f / d;
}

// Closest (any == false) or any hit of one ray; counts box and triangle tests.
template <bool WIDE = false>
void trace1(const Bvh& b, const float* ray, bool any, float* hit, uint64_t& nbox, uint64_t& ntri) {
  const float* o = ray;
  const float* d = ray + 4;
  Shear s = shear_of(o, d, ray[3], ray[7]);
  float inv[3] = {rcp_safe(d[0]), rcp_safe(d[1]), rcp_safe(d[2])};
  float bt = ray[7], bu = 0.0f, bv = 0.0f;
  uint32_t bp = 0xFFFFFFFFu;
  uint32_t stack[256];
  int sp = 0;
  uint32_t cur = 0;
  float tn0;
  nbox++;
  if (!box_enter<WIDE>(b.nodes[0], o, inv, s.tmin, bt, tn0)) goto done;
  while (true) {
    const BNode& nd = b.nodes[cur];
    if (nd.count) {
      for (uint32_t i = nd.a; i < nd.a + nd.count; ++i) {
        uint32_t p = b.idx[i];
        float t, u, v;
        ntri++;
        if (tri_test(s, &b.tris[9 * (size_t)p], t, u, v)) {
          if (any) {
            bp = 1;
            goto done;
          }
          if (before(t, p, bt, bp)) {
            bt = t;
            bu = u;
            bv = v;
            bp = p;
          }
        }
      }
    } else {
      float t0, t1;
      nbox += 2;
      bool h0 = box_enter<WIDE>(b.nodes[nd.a], o, inv, s.tmin, bt, t0);
      bool h1 = box_enter<WIDE>(b.nodes[nd.a + 1], o, inv, s.tmin, bt, t1);
      if (h0 && h1) {
        uint32_t nearer = (t1 < t0) ? nd.a + 1 : nd.a, farther = (t1 < t0) ? nd.a : nd.a + 1;
        if (sp < 256) stack[sp++] = farther;
        cur = nearer;
        continue;
      }
      if (h0 || h1) {
        cur = h0 ? nd.a : nd.a + 1;
        continue;
      }
    }
    // pop, skipping nodes that the shrinking best t now excludes
    bool found = false;
    while (sp > 0) {
      uint32_t c = stack[--sp];
      float tn;
      if (any || box_enter<WIDE>(b.nodes[c], o, inv, s.tmin, bt, tn)) {
        cur = c;
        found = true;
        break;
      }
    }
    if (!found) break;
  }
done:
  if (any) {
    hit[0] = bp == 1 ? 1.0f : 0.0f;
    return;
  }
  hit[0] = bp == 0xFFFFFFFFu ? INFINITY : bt;
  hit[1] = bu;
  hit[2] = bv;
  std::memcpy(&hit[3], &bp, 4);
}


// ---- hooks for the integrator (oracle_render.cpp) --------------------------
// Per-thread tally of the FP32 queries' box and triangle tests (closest hit
// at [0], [1], any hit at [2], [3]) while a caller has set it.
thread_local uint64_t* g_tally = nullptr;
void oracle_set_tally(uint64_t* t4) { g_tally = t4; }

bool oracle_closest_f(const Bvh* b, const float* ray8, float* hit4) {
  uint64_t nb = 0, nt = 0;
  trace1(*b, ray8, false, hit4, nb, nt);
  if (g_tally) {
    g_tally[0] += nb;
    g_tally[1] += nt;
  }
  uint32_t p;
  std::memcpy(&p, &hit4[3], 4);
  return p != 0xFFFFFFFFu;
}

bool oracle_any_f(const Bvh* b, const float* ray8) {
  uint64_t nb = 0, nt = 0;
  float h[4];
  trace1(*b, ray8, true, h, nb, nt);
  if (g_tally) {
    g_tally[2] += nb;
    g_tally[3] += nt;
  }
  return h[0] != 0.0f;
}

// FP64 watertight test (the same algorithm evaluated in double).
static bool tri_test_d(const double* o, const int* ax, const double* S, const float* v, double tmin,
                       double tmax, double& t, double& u, double& w) {
  double X[3], Y[3], Z[3];
  for (int j = 0; j < 3; ++j) {
    double r[3] = {(double)v[3 * j] - o[0], (double)v[3 * j + 1] - o[1], (double)v[3 * j + 2] - o[2]};
    X[j] = r[ax[0]] - S[0] * r[ax[2]];
    Y[j] = r[ax[1]] - S[1] * r[ax[2]];
    Z[j] = S[2] * r[ax[2]];
  }
  double e0 = X[2] * Y[1] - Y[2] * X[1], e1 = X[0] * Y[2] - Y[0] * X[2], e2 = X[1] * Y[0] - Y[1] * X[0];
  if ((e0 < 0 || e1 < 0 || e2 < 0) && (e0 > 0 || e1 > 0 || e2 > 0)) return false;
  double det = e0 + e1 + e2;
  if (det == 0.0) return false;
  double tt = (e0 * Z[0] + e1 * Z[1] + e2 * Z[2]) / det;
  if (!(tt > tmin && tt < tmax)) return false;
  t = tt;
  u = e1 / det;
  w = e2 / det;
  return true;
}

static bool trace_d(const Bvh* b, const double* o, const double* d, double tmin, double tmax, bool any,
                    double* ot, double* ou, double* ov, uint32_t* op) {
  int kz = 0;
  double a0 = std::fabs(d[0]), a1 = std::fabs(d[1]), a2 = std::fabs(d[2]);
  if (a0 > a1) kz = a0 > a2 ? 0 : 2;
  else kz = a1 > a2 ? 1 : 2;
  int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
  if (d[kz] < 0.0) std::swap(kx, ky);
  int ax[3] = {kx, ky, kz};
  double S[3] = {d[kx] / d[kz], d[ky] / d[kz], 1.0 / d[kz]};
  double inv[3];
  for (int k = 0; k < 3; ++k) inv[k] = std::fabs(d[k]) < 1e-300 ? (d[k] < 0 ? -1e300 : 1e300) : 1.0 / d[k];
  auto enter = [&](const BNode& nd, double cap) {
    double tn = tmin, tf = cap;
    for (int k = 0; k < 3; ++k) {
      double p = ((double)nd.lo[k] - o[k]) * inv[k], q = ((double)nd.hi[k] - o[k]) * inv[k];
      tn = std::max(tn, std::min(p, q));
      tf = std::min(tf, std::max(p, q));
    }
    return tn <= tf;
  };
  double bt = tmax, bu = 0, bv = 0;
  uint32_t bp = 0xFFFFFFFFu;
  uint32_t stack[256];
  int sp = 0;
  if (!enter(b->nodes[0], bt)) return false;
  stack[sp++] = 0;
  while (sp > 0) {
    const BNode& nd = b->nodes[stack[--sp]];
    if (!enter(nd, bt)) continue;
    if (nd.count) {
      for (uint32_t i = nd.a; i < nd.a + nd.count; ++i) {
        uint32_t p = b->idx[i];
        double t, u, w;
        if (tri_test_d(o, ax, S, &b->tris[9 * (size_t)p], tmin, tmax, t, u, w)) {
          if (any) return true;
          if (t < bt || (t == bt && p < bp)) {
            bt = t;
            bu = u;
            bv = w;
            bp = p;
          }
        }
      }
    } else if (sp + 2 <= 256) {
      stack[sp++] = nd.a + 1;
      stack[sp++] = nd.a;
    }
  }
  if (bp == 0xFFFFFFFFu) return false;
  *ot = bt;
  *ou = bu;
  *ov = bv;
  *op = bp;
  return true;
}

bool oracle_closest_d(const Bvh* b, const double* o, const double* d, double tmin, double tmax,
                      double* t, double* u, double* v, uint32_t* prim) {
  return trace_d(b, o, d, tmin, tmax, false, t, u, v, prim);
}

bool oracle_any_d(const Bvh* b, const double* o, const double* d, double tmin, double tmax) {
  double t, u, v;
  uint32_t p;
  return trace_d(b, o, d, tmin, tmax, true, &t, &u, &v, &p);
}

// ---- two-level scenes (instances of BLASes) ------------------------------
// The reference traces rays against a TLAS of BLAS instances
// (scene_resources/mod.rs:159-276); the driver maps each ray into the
// instance's object space.  The restatement: the world-to-object matrix is
// the inverse of the instance's object-to-world 3x4, evaluated in double by
// cofactors and rounded to float; an instance whose transform is exactly the
// identity keeps the ray; otherwise o' = ((m0 x + m1 y) + m2 z) + m3 and
// d' = (m0 x + m1 y) + m2 z per row in float.  The closest hit is the
// lexicographic minimum of (t, custom + BLAS prim) over all instances.

struct Inst {  // apt_instance: 12 floats, blas, custom, 2 pad (64 B)
  float xf[12];
  uint32_t blas, custom, pad[2];
};

bool affine_inverse(const float* o2w, float* w2o) {
  const double m00 = o2w[0], m01 = o2w[1], m02 = o2w[2], t0 = o2w[3];
  const double m10 = o2w[4], m11 = o2w[5], m12 = o2w[6], t1 = o2w[7];
  const double m20 = o2w[8], m21 = o2w[9], m22 = o2w[10], t2 = o2w[11];
  // cofactors of the first row, then the determinant
  const double c00 = m11 * m22 - m12 * m21, c01 = m12 * m20 - m10 * m22, c02 = m10 * m21 - m11 * m20;
  const double det = m00 * c00 + m01 * c01 + m02 * c02;
  if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return false;
  const double r[3][3] = {{c00 / det, (m02 * m21 - m01 * m22) / det, (m01 * m12 - m02 * m11) / det},
                          {c01 / det, (m00 * m22 - m02 * m20) / det, (m02 * m10 - m00 * m12) / det},
                          {c02 / det, (m01 * m20 - m00 * m21) / det, (m00 * m11 - m01 * m10) / det}};
  for (int i = 0; i < 3; ++i) {
    for (int k = 0; k < 3; ++k) w2o[4 * i + k] = (float)r[i][k];
    w2o[4 * i + 3] = (float)(-((r[i][0] * t0 + r[i][1] * t1) + r[i][2] * t2));
  }
  for (int k = 0; k < 12; ++k)
    if (!std::isfinite(w2o[k])) return false;
  return true;
}

static bool is_identity(const float* m) {
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 4; ++k) {
      const float want = (k == i) ? 1.0f : 0.0f;
      if (m[4 * i + k] != want || std::signbit(m[4 * i + k])) return false;
    }
  return true;
}

// The ray (8 floats) of instance `in` in its object space.
static bool inst_ray(const Inst& in, const float* ray, float* out) {
  std::memcpy(out, ray, 8 * sizeof(float));
  if (is_identity(in.xf)) return true;
  float w[12];
  if (!affine_inverse(in.xf, w)) return false;
  if (ors::g_plant == 7) {
    std::swap(w[1], w[4]);
    std::swap(w[2], w[8]);
    std::swap(w[6], w[9]);
  } else if (ors::g_plant == 8) {
    w[3] = -w[3];
    w[7] = -w[7];
    w[11] = -w[11];
  }
  for (int i = 0; i < 3; ++i) {
    const float* m = w + 4 * i;
    out[i] = ((m[0] * ray[0] + m[1] * ray[1]) + m[2] * ray[2]) + m[3];
    out[4 + i] = (m[0] * ray[4] + m[1] * ray[5]) + m[2] * ray[6];
  }
  return true;
}

struct InstScene {
  std::vector<float> tris;          // all BLASes, 9 floats per triangle
  std::vector<uint32_t> first;      // n_blas + 1 triangle offsets
  std::vector<Bvh*> bvh;            // per BLAS
  ~InstScene() {
    for (Bvh* b : bvh) delete b;
  }
};

// One ray against all instances: brute force over every BLAS triangle, or
// each BLAS's BVH2 with the widened box test.
static void inst_trace1(const InstScene& sc, const Inst* inst, uint32_t n_inst, const float* ray, bool any,
                        bool brute, float* hit) {
  float bt = ray[7], bu = 0.0f, bv = 0.0f;
  uint32_t bp = 0xFFFFFFFFu;
  bool occ = false;
  for (uint32_t k = 0; k < n_inst && !occ; ++k) {
    const Inst& in = inst[k];
    float r[8];
    if (!inst_ray(in, ray, r)) continue;
    const uint32_t f0 = sc.first[in.blas], f1 = sc.first[in.blas + 1];
    if (brute) {
      Shear s = shear_of(r, r + 4, r[3], r[7]);
      for (uint32_t p = f0; p < f1 && !occ; ++p) {
        float t, u, v;
        if (!tri_test(s, &sc.tris[9 * (size_t)p], t, u, v)) continue;
        if (any) {
          occ = true;
          break;
        }
        const uint32_t g = in.custom + (p - f0);
        if (before(t, g, bt, bp)) {
          bt = t;
          bu = u;
          bv = v;
          bp = g;
        }
      }
    } else {
      float h[4];
      uint64_t nb = 0, nt = 0;
      trace1<true>(*sc.bvh[in.blas], r, any, h, nb, nt);
      if (any) {
        occ = h[0] != 0.0f;
        continue;
      }
      uint32_t lp;
      std::memcpy(&lp, &h[3], 4);
      if (lp == 0xFFFFFFFFu) continue;
      const uint32_t g = in.custom + lp;
      if (before(h[0], g, bt, bp)) {
        bt = h[0];
        bu = h[1];
        bv = h[2];
        bp = g;
      }
    }
  }
  if (any) {
    hit[0] = occ ? 1.0f : 0.0f;
    return;
  }
  hit[0] = bp == 0xFFFFFFFFu ? INFINITY : bt;
  hit[1] = bu;
  hit[2] = bv;
  std::memcpy(&hit[3], &bp, 4);
}

}  // namespace orc

using namespace orc;

extern "C" {

uint32_t or_pcg_hash(uint32_t x) { return pcg_hash(x); }
uint32_t or_xor_shift(uint32_t x) { return xor_shift(x); }
uint32_t or_hash_combine(uint32_t x, uint32_t y) { return hash_combine(x, y); }
uint32_t or_wangh_hash(uint32_t x) { return wangh_hash(x); }
void or_uniform_floats(uint32_t state, uint32_t n, float* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = uniform(state);
}

void or_rgb9e5_pack(const float* rgb, uint32_t n, uint32_t* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = rgb9e5_pack(rgb + 3 * (size_t)i);
}
void or_rgb9e5_unpack(const uint32_t* d, uint32_t n, float* rgb) {
  for (uint32_t i = 0; i < n; ++i) rgb9e5_unpack(d[i], rgb + 3 * (size_t)i);
}
void or_oct30_pack(const float* v, uint32_t n, uint32_t* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = oct30_pack(v + 3 * (size_t)i);
}
void or_oct30_unpack(const uint32_t* p, uint32_t n, float* v) {
  for (uint32_t i = 0; i < n; ++i) oct30_unpack(p[i], v + 3 * (size_t)i);
}

int or_tri_test(const float* ray, const float* tri, float* out_tuv) {
  Shear s = shear_of(ray, ray + 4, ray[3], ray[7]);
  float t, u, v;
  if (!tri_test(s, tri, t, u, v)) return 0;
  out_tuv[0] = t;
  out_tuv[1] = u;
  out_tuv[2] = v;
  return 1;
}

void* or_bvh_build(const float* tris, uint32_t n) { return build(tris, n); }
void or_bvh_destroy(void* b) { delete static_cast<Bvh*>(b); }
uint32_t or_bvh_node_count(void* b) { return (uint32_t)static_cast<Bvh*>(b)->nodes.size(); }

// rays: n x 8 floats (apt_ray); out: n x 4 floats (closest: t,u,v,prim bits)
// or n floats (any: 0/1).  counts[0..1] receive total box and triangle tests.
void or_trace(void* bh, const float* rays, uint32_t n, int any, float* out, uint64_t* counts,
              int threads) {
  const Bvh& b = *static_cast<Bvh*>(bh);
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<uint32_t> next{0};
  std::vector<uint64_t> nb(threads, 0), nt(threads, 0);
  auto work = [&](int tid) {
    uint64_t lb = 0, lt = 0;
    while (true) {
      uint32_t s = next.fetch_add(1024);
      if (s >= n) break;
      uint32_t e = std::min(n, s + 1024);
      for (uint32_t i = s; i < e; ++i) {
        float h[4];
        trace1(b, rays + 8 * (size_t)i, any != 0, h, lb, lt);
        if (any) out[i] = h[0];
        else std::memcpy(out + 4 * (size_t)i, h, 16);
      }
    }
    nb[tid] = lb;
    nt[tid] = lt;
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto& th : pool) th.join();
  if (counts) {
    counts[0] = counts[1] = 0;
    for (int t = 0; t < threads; ++t) {
      counts[0] += nb[t];
      counts[1] += nt[t];
    }
  }
}

// Brute force over all triangles (same hit contract), for small inputs.
void or_brute(const float* tris, uint32_t ntris, const float* rays, uint32_t n, int any,
              float* out) {
  for (uint32_t i = 0; i < n; ++i) {
    const float* r = rays + 8 * (size_t)i;
    Shear s = shear_of(r, r + 4, r[3], r[7]);
    float bt = r[7], bu = 0, bv = 0;
    uint32_t bp = 0xFFFFFFFFu;
    for (uint32_t p = 0; p < ntris; ++p) {
      float t, u, v;
      if (tri_test(s, tris + 9 * (size_t)p, t, u, v) && before(t, p, bt, bp)) {
        bt = t;
        bu = u;
        bv = v;
        bp = p;
      }
    }
    if (any) {
      out[i] = bp != 0xFFFFFFFFu ? 1.0f : 0.0f;
    } else {
      out[4 * (size_t)i] = bp == 0xFFFFFFFFu ? INFINITY : bt;
      out[4 * (size_t)i + 1] = bu;
      out[4 * (size_t)i + 2] = bv;
      std::memcpy(out + 4 * (size_t)i + 3, &bp, 4);
    }
  }
}

// block.wgsl:3-15: linear pixel -> 64x64 block-major pixel index.
uint32_t or_block_index(uint32_t x, uint32_t y, uint32_t width) {
  const uint32_t B = 64;
  uint32_t bx = x / B, by = y / B, ox = x % B, oy = y % B;
  uint32_t per_row = (width + B - 1) / B;
  return (by * per_row + bx) * (B * B) + oy * B + ox;
}

// Two-level scenes.  tris: the BLASes' triangles one after another (9 floats
// each); blas_first: n_blas + 1 offsets into them (in triangles).
void* or_inst_create(const float* tris, const uint32_t* blas_first, uint32_t n_blas) {
  InstScene* sc = new InstScene();
  sc->first.assign(blas_first, blas_first + n_blas + 1);
  sc->tris.assign(tris, tris + 9 * (size_t)blas_first[n_blas]);
  for (uint32_t b = 0; b < n_blas; ++b)
    sc->bvh.push_back(build(tris + 9 * (size_t)blas_first[b], blas_first[b + 1] - blas_first[b]));
  return sc;
}
void or_inst_destroy(void* sc) { delete static_cast<InstScene*>(sc); }

// instances: n_inst apt_instance records (64 B each); rays: n x 8 floats;
// out: n x 4 floats (closest: t, u, v, prim bits) or n floats (any: 0/1).
// brute != 0: every triangle of every instance; else the BLASes' BVH2s.
void or_inst_trace(void* scp, const void* instances, uint32_t n_inst, const float* rays, uint32_t n, int any,
                   int brute, float* out, int threads) {
  const InstScene& sc = *static_cast<InstScene*>(scp);
  const Inst* inst = static_cast<const Inst*>(instances);
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  std::atomic<uint32_t> next{0};
  auto work = [&]() {
    while (true) {
      const uint32_t s0 = next.fetch_add(256);
      if (s0 >= n) break;
      const uint32_t e = std::min(n, s0 + 256);
      for (uint32_t i = s0; i < e; ++i) {
        float h[4];
        inst_trace1(sc, inst, n_inst, rays + 8 * (size_t)i, any != 0, brute != 0, h);
        if (any) out[i] = h[0];
        else std::memcpy(out + 4 * (size_t)i, h, 16);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto& th : pool) th.join();
}

int or_affine_inverse(const float* o2w, float* w2o) { return affine_inverse(o2w, w2o) ? 1 : 0; }

// resolve.wgsl:29-38 ACES fit, clamped to [0, 1].
float or_aces(float x) {
  float s = (x * (2.51f * x + 0.03f)) / (x * (2.43f * x + 0.59f) + 0.14f);
  return std::min(std::max(s, 0.0f), 1.0f);
}

}  // extern "C"
