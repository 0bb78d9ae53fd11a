// Baseline JPEG decoding for glTF textures (ITU-T T.81 sequential DCT,
// Huffman, 8-bit), into RGBA8.
//
// The reference decodes texture images with the `image` crate
// (crates/appearance-model/src/asset.rs:382-420 through gltf's importer);
// ClearCoatTest.glb's clearcoat normal map is a 4:2:0 JPEG.  The samples
// this writes are those of libjpeg-turbo's default decompression: the
// accurate integer IDCT (Loeffler-Ligtenberg-Moschytz) with its range
// limiting, "fancy" triangle-filter upsampling of subsampled chroma (2x1,
// 1x2 and 2x2; plain replication for other ratios and for chroma planes two
// samples wide or less), and JFIF YCbCr -> RGB in 16-bit fixed point.
// Progressive, arithmetic-coded, lossless and 12-bit files are refused.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "apt_internal.h"

namespace apt {
namespace {

constexpr uint8_t ZIGZAG[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// A Huffman table in the form of T.81 F.2.2.3: per code length, the largest
// code (or -1) and the index of its first symbol.
struct Huff {
  bool present = false;
  int32_t maxcode[18];
  int32_t valptr[17];
  int32_t mincode[17];
  uint8_t vals[256];
  // 9-bit lookahead: (length << 8) | symbol, 0 when longer
  uint16_t fast[512];
};

bool build_huff(const uint8_t* bits, const uint8_t* vals, int nvals, Huff& h) {
  int k = 0;
  int32_t code = 0;
  std::memset(h.fast, 0, sizeof(h.fast));
  for (int len = 1; len <= 16; ++len) {
    h.valptr[len] = k;
    h.mincode[len] = code;
    code += bits[len - 1];
    k += bits[len - 1];
    if (k > nvals || code > (1 << len)) return false;
    h.maxcode[len] = bits[len - 1] ? code - 1 : -1;
    code <<= 1;
  }
  h.maxcode[17] = 0x7FFFFFFF;
  std::memcpy(h.vals, vals, (size_t)nvals);
  // lookahead for codes of at most 9 bits
  for (int len = 1; len <= 9; ++len)
    for (int32_t c = h.mincode[len]; c <= h.maxcode[len]; ++c) {
      const int sym = h.vals[h.valptr[len] + c - h.mincode[len]];
      const int shift = 9 - len;
      for (int f = 0; f < (1 << shift); ++f) h.fast[(c << shift) | f] = (uint16_t)((len << 8) | sym);
    }
  h.present = true;
  return true;
}

// Entropy-coded bits: 0xFF 0x00 is a data 0xFF; a marker ends the data and
// zeros are read past it.  `over` tells whether any such zero was used: a
// scan that needs bits past its data is truncated.
struct Bits {
  const uint8_t* p;
  size_t n, pos;
  uint64_t acc = 0;
  int cnt = 0;
  bool marker = false;
  int64_t real = 0, used = 0;  // bits read from the data / consumed since the last restart
  bool over() const { return used > real; }
  void fill() {
    while (cnt <= 56) {
      uint32_t b = 0;
      if (!marker && pos < n) {
        b = p[pos];
        if (b == 0xFF) {
          const uint8_t nx = pos + 1 < n ? p[pos + 1] : 0;
          if (nx == 0x00) {
            pos += 2;
          } else {
            marker = true;
            b = 0;
          }
        } else {
          ++pos;
        }
        if (!marker) real += 8;
      }
      acc |= (uint64_t)b << (56 - cnt);
      cnt += 8;
    }
  }
  uint32_t peek(int k) {
    if (cnt < k) fill();
    return (uint32_t)(acc >> (64 - k));
  }
  void skip(int k) {
    acc <<= k;
    cnt -= k;
    used += k;
  }
  uint32_t get(int k) {
    if (k == 0) return 0;
    const uint32_t v = peek(k);
    skip(k);
    return v;
  }
  // Restart: drop the partial byte, step over the RSTn marker.
  bool restart() {
    if (over()) return false;
    acc = 0;
    cnt = 0;
    marker = false;
    real = used = 0;
    while (pos + 1 < n && !(p[pos] == 0xFF && p[pos + 1] >= 0xD0 && p[pos + 1] <= 0xD7)) ++pos;
    if (pos + 1 >= n) return false;
    pos += 2;
    return true;
  }
};

int decode_sym(Bits& b, const Huff& h) {
  const uint32_t look = b.peek(9);
  const uint16_t f = h.fast[look];
  if (f) {
    b.skip(f >> 8);
    return f & 0xFF;
  }
  int32_t code = (int32_t)b.peek(16);
  for (int len = 10; len <= 16; ++len) {
    const int32_t c = code >> (16 - len);
    if (c <= h.maxcode[len]) {
      b.skip(len);
      return h.vals[h.valptr[len] + c - h.mincode[len]];
    }
  }
  return -1;
}

inline int extend(uint32_t v, int s) { return s == 0 ? 0 : (v < (1u << (s - 1)) ? (int)v - (1 << s) + 1 : (int)v); }

// Output sample of the IDCT: x + 128 for x in [-128, 127], 255 above up to
// 511, 0 below down to -512, the pattern repeating every 1024 (libjpeg's
// post-IDCT range-limit table indexed by x & 1023).
inline uint8_t idct_limit(int32_t x) {
  const int32_t i = x & 1023;
  if (i < 128) return (uint8_t)(i + 128);
  if (i < 512) return 255;
  if (i < 896) return 0;
  return (uint8_t)(i - 896);
}

inline int32_t descale(int32_t x, int n) { return (x + (1 << (n - 1))) >> n; }

// Accurate integer IDCT of dequantised coefficients (natural order) into an
// 8x8 block of samples.  Constants: round(x * 2^13).
void idct_islow(const int32_t* in, uint8_t* out, size_t pitch) {
  constexpr int CB = 13, P1 = 2;
  int32_t ws[64];
  for (int c = 0; c < 8; ++c) {  // columns
    const int32_t* s = in + c;
    if (!s[8] && !s[16] && !s[24] && !s[32] && !s[40] && !s[48] && !s[56]) {
      const int32_t dc = s[0] * (1 << P1);
      for (int r = 0; r < 8; ++r) ws[r * 8 + c] = dc;
      continue;
    }
    int32_t z2 = s[16], z3 = s[48];
    int32_t z1 = (z2 + z3) * 4433;
    const int32_t t2 = z1 - z3 * 15137, t3 = z1 + z2 * 6270;
    z2 = s[0];
    z3 = s[32];
    const int32_t t0 = (z2 + z3) * (1 << CB), t1 = (z2 - z3) * (1 << CB);
    const int32_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
    int32_t a0 = s[56], a1 = s[40], a2 = s[24], a3 = s[8];
    const int32_t y1 = a0 + a3, y2 = a1 + a2, y3 = a0 + a2, y4 = a1 + a3;
    const int32_t z5 = (y3 + y4) * 9633;
    a0 *= 2446;
    a1 *= 16819;
    a2 *= 25172;
    a3 *= 12299;
    const int32_t m1 = -y1 * 7373, m2 = -y2 * 20995, m3 = -y3 * 16069 + z5, m4 = -y4 * 3196 + z5;
    a0 += m1 + m3;
    a1 += m2 + m4;
    a2 += m2 + m3;
    a3 += m1 + m4;
    const int sh = CB - P1;
    ws[0 * 8 + c] = descale(t10 + a3, sh);
    ws[7 * 8 + c] = descale(t10 - a3, sh);
    ws[1 * 8 + c] = descale(t11 + a2, sh);
    ws[6 * 8 + c] = descale(t11 - a2, sh);
    ws[2 * 8 + c] = descale(t12 + a1, sh);
    ws[5 * 8 + c] = descale(t12 - a1, sh);
    ws[3 * 8 + c] = descale(t13 + a0, sh);
    ws[4 * 8 + c] = descale(t13 - a0, sh);
  }
  for (int r = 0; r < 8; ++r) {  // rows
    const int32_t* w = ws + r * 8;
    uint8_t* o = out + (size_t)r * pitch;
    if (!w[1] && !w[2] && !w[3] && !w[4] && !w[5] && !w[6] && !w[7]) {
      const uint8_t v = idct_limit(descale(w[0], P1 + 3));
      for (int k = 0; k < 8; ++k) o[k] = v;
      continue;
    }
    int32_t z2 = w[2], z3 = w[6];
    int32_t z1 = (z2 + z3) * 4433;
    const int32_t t2 = z1 - z3 * 15137, t3 = z1 + z2 * 6270;
    const int32_t t0 = (w[0] + w[4]) * (1 << CB), t1 = (w[0] - w[4]) * (1 << CB);
    const int32_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
    int32_t a0 = w[7], a1 = w[5], a2 = w[3], a3 = w[1];
    const int32_t y1 = a0 + a3, y2 = a1 + a2, y3 = a0 + a2, y4 = a1 + a3;
    const int32_t z5 = (y3 + y4) * 9633;
    a0 *= 2446;
    a1 *= 16819;
    a2 *= 25172;
    a3 *= 12299;
    const int32_t m1 = -y1 * 7373, m2 = -y2 * 20995, m3 = -y3 * 16069 + z5, m4 = -y4 * 3196 + z5;
    a0 += m1 + m3;
    a1 += m2 + m4;
    a2 += m2 + m3;
    a3 += m1 + m4;
    const int sh = CB + P1 + 3;
    o[0] = idct_limit(descale(t10 + a3, sh));
    o[7] = idct_limit(descale(t10 - a3, sh));
    o[1] = idct_limit(descale(t11 + a2, sh));
    o[6] = idct_limit(descale(t11 - a2, sh));
    o[2] = idct_limit(descale(t12 + a1, sh));
    o[5] = idct_limit(descale(t12 - a1, sh));
    o[3] = idct_limit(descale(t13 + a0, sh));
    o[4] = idct_limit(descale(t13 - a0, sh));
  }
}

struct Comp {
  int id = 0, h = 1, v = 1, tq = 0;
  int td = 0, ta = 0;
  uint32_t bw = 0, bh = 0;  // blocks per row / column, padded to whole MCUs
  uint32_t dw = 0, dh = 0;  // downsampled size: ceil(W * h / hmax), ceil(H * v / vmax)
  std::vector<uint8_t> plane;  // bw * 8 x bh * 8 samples
  int pred = 0;
};

// One component plane upsampled to the frame's full sampling (width
// w_out = dw * fx, height dh * fy), libjpeg's methods per ratio.
std::vector<uint8_t> upsample(const Comp& c, int fx, int fy, uint32_t& w_out, uint32_t& h_out) {
  const uint32_t dw = c.dw, dh = c.dh, pitch = c.bw * 8;
  w_out = dw * (uint32_t)fx;
  h_out = dh * (uint32_t)fy;
  std::vector<uint8_t> out((size_t)w_out * h_out);
  const auto row = [&](int64_t y) {  // edge rows replicate
    y = y < 0 ? 0 : (y >= (int64_t)dh ? (int64_t)dh - 1 : y);
    return c.plane.data() + (size_t)y * pitch;
  };
  if (fx == 1 && fy == 1) {
    for (uint32_t y = 0; y < dh; ++y) std::memcpy(&out[(size_t)y * w_out], row(y), dw);
  } else if (fx == 2 && fy == 1 && dw > 2) {
    for (uint32_t y = 0; y < dh; ++y) {
      const uint8_t* in = row(y);
      uint8_t* o = &out[(size_t)y * w_out];
      o[0] = in[0];
      o[1] = (uint8_t)((in[0] * 3 + in[1] + 2) >> 2);
      for (uint32_t x = 1; x + 1 < dw; ++x) {
        const int v = in[x] * 3;
        o[2 * x] = (uint8_t)((v + in[x - 1] + 1) >> 2);
        o[2 * x + 1] = (uint8_t)((v + in[x + 1] + 2) >> 2);
      }
      o[2 * dw - 2] = (uint8_t)((in[dw - 1] * 3 + in[dw - 2] + 1) >> 2);
      o[2 * dw - 1] = in[dw - 1];
    }
  } else if (fx == 1 && fy == 2) {
    for (uint32_t y = 0; y < dh; ++y)
      for (int v = 0; v < 2; ++v) {
        const uint8_t* i0 = row(y);
        const uint8_t* i1 = row(v == 0 ? (int64_t)y - 1 : (int64_t)y + 1);
        const int bias = v == 0 ? 1 : 2;
        uint8_t* o = &out[(size_t)(2 * y + v) * w_out];
        for (uint32_t x = 0; x < dw; ++x) o[x] = (uint8_t)((i0[x] * 3 + i1[x] + bias) >> 2);
      }
  } else if (fx == 2 && fy == 2 && dw > 2) {
    for (uint32_t y = 0; y < dh; ++y)
      for (int v = 0; v < 2; ++v) {
        const uint8_t* i0 = row(y);
        const uint8_t* i1 = row(v == 0 ? (int64_t)y - 1 : (int64_t)y + 1);
        uint8_t* o = &out[(size_t)(2 * y + v) * w_out];
        int last = 0, cur = i0[0] * 3 + i1[0], next = i0[1] * 3 + i1[1];
        o[0] = (uint8_t)((cur * 4 + 8) >> 4);
        o[1] = (uint8_t)((cur * 3 + next + 7) >> 4);
        last = cur;
        cur = next;
        for (uint32_t x = 1; x + 1 < dw; ++x) {
          next = i0[x + 1] * 3 + i1[x + 1];
          o[2 * x] = (uint8_t)((cur * 3 + last + 8) >> 4);
          o[2 * x + 1] = (uint8_t)((cur * 3 + next + 7) >> 4);
          last = cur;
          cur = next;
        }
        o[2 * dw - 2] = (uint8_t)((cur * 3 + last + 8) >> 4);
        o[2 * dw - 1] = (uint8_t)((cur * 4 + 7) >> 4);
      }
  } else {  // replication
    for (uint32_t y = 0; y < h_out; ++y) {
      const uint8_t* in = row(y / (uint32_t)fy);
      for (uint32_t x = 0; x < w_out; ++x) out[(size_t)y * w_out + x] = in[x / (uint32_t)fx];
    }
  }
  return out;
}

}  // namespace

bool jpeg_decode(const uint8_t* p, size_t n, uint32_t& W, uint32_t& H, std::vector<uint32_t>& rgba, std::string& err) {
  if (n < 4 || p[0] != 0xFF || p[1] != 0xD8) return err = "not a JPEG", false;
  uint16_t qt[4][64] = {};
  bool qt_ok[4] = {};
  Huff dc[4], ac[4];
  std::vector<Comp> comps;
  int hmax = 1, vmax = 1;
  uint32_t restart = 0;
  bool frame = false, jfif = false, adobe = false, done = false;
  int adobe_transform = -1;
  W = H = 0;
  size_t off = 2;
  while (off + 4 <= n && !done) {
    if (p[off] != 0xFF) return err = "JPEG marker expected", false;
    const uint8_t m = p[off + 1];
    if (m == 0xFF) {  // fill byte
      ++off;
      continue;
    }
    if (m == 0xD9) break;
    if ((m >= 0xD0 && m <= 0xD7) || m == 0x01) {  // markers without a length
      off += 2;
      continue;
    }
    const size_t len = ((size_t)p[off + 2] << 8) | p[off + 3];
    if (len < 2 || off + 2 + len > n) return err = "truncated JPEG segment", false;
    const uint8_t* d = p + off + 4;
    const size_t dl = len - 2;
    if (m == 0xDB) {  // DQT
      size_t k = 0;
      while (k < dl) {
        const int pq = d[k] >> 4, tq = d[k] & 15;
        if (tq > 3 || k + 1 + 64 * (pq + 1) > dl) return err = "bad DQT", false;
        for (int i = 0; i < 64; ++i)
          qt[tq][ZIGZAG[i]] = pq ? (uint16_t)((d[k + 1 + 2 * i] << 8) | d[k + 2 + 2 * i]) : d[k + 1 + i];
        qt_ok[tq] = true;
        k += 1 + 64 * (pq + 1);
      }
    } else if (m == 0xC4) {  // DHT
      size_t k = 0;
      while (k + 17 <= dl) {
        const int tc = d[k] >> 4, th = d[k] & 15;
        int cnt = 0;
        for (int i = 0; i < 16; ++i) cnt += d[k + 1 + i];
        if (tc > 1 || th > 3 || cnt > 256 || k + 17 + cnt > dl) return err = "bad DHT", false;
        if (!build_huff(d + k + 1, d + k + 17, cnt, tc ? ac[th] : dc[th])) return err = "bad Huffman table", false;
        k += 17 + cnt;
      }
    } else if (m == 0xC0 || m == 0xC1) {  // SOF0 / SOF1: sequential, Huffman
      if (dl < 6 || d[0] != 8) return err = "only 8-bit JPEG is decoded", false;
      H = ((uint32_t)d[1] << 8) | d[2];
      W = ((uint32_t)d[3] << 8) | d[4];
      const int nf = d[5];
      if (W == 0 || H == 0 || (nf != 1 && nf != 3) || dl < 6 + 3 * (size_t)nf)
        return err = "unsupported JPEG frame", false;
      if ((uint64_t)W * H > (1ull << 28)) return err = "JPEG too large", false;
      comps.resize((size_t)nf);
      for (int i = 0; i < nf; ++i) {
        comps[i].id = d[6 + 3 * i];
        comps[i].h = d[7 + 3 * i] >> 4;
        comps[i].v = d[7 + 3 * i] & 15;
        comps[i].tq = d[8 + 3 * i];
        if (comps[i].h < 1 || comps[i].h > 4 || comps[i].v < 1 || comps[i].v > 4 || comps[i].tq > 3)
          return err = "bad JPEG component", false;
        hmax = std::max(hmax, comps[i].h);
        vmax = std::max(vmax, comps[i].v);
      }
      const uint32_t mcux = (W + 8 * hmax - 1) / (8 * hmax), mcuy = (H + 8 * vmax - 1) / (8 * vmax);
      for (Comp& c : comps) {
        c.bw = mcux * c.h;
        c.bh = mcuy * c.v;
        c.dw = (W * c.h + hmax - 1) / hmax;
        c.dh = (H * c.v + vmax - 1) / vmax;
        c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
      }
      frame = true;
    } else if ((m >= 0xC2 && m <= 0xC3) || (m >= 0xC5 && m <= 0xCF && m != 0xC8 && m != 0xCC)) {
      return err = "progressive, lossless and arithmetic-coded JPEG are not decoded", false;
    } else if (m == 0xDD) {  // DRI
      if (dl < 2) return err = "bad DRI", false;
      restart = ((uint32_t)d[0] << 8) | d[1];
    } else if (m == 0xE0) {
      if (dl >= 5 && !std::memcmp(d, "JFIF", 5)) jfif = true;
    } else if (m == 0xEE) {  // Adobe: colour transform flag
      if (dl >= 12 && !std::memcmp(d, "Adobe", 5)) {
        adobe = true;
        adobe_transform = d[11];
      }
    } else if (m == 0xDA) {  // SOS
      if (!frame) return err = "JPEG scan before its frame", false;
      const int ns = d[0];
      if (ns < 1 || ns > (int)comps.size() || dl < 4 + 2 * (size_t)ns) return err = "bad SOS", false;
      std::vector<Comp*> sc;
      for (int i = 0; i < ns; ++i) {
        Comp* c = nullptr;
        for (Comp& x : comps)
          if (x.id == d[1 + 2 * i]) c = &x;
        if (!c) return err = "SOS names an unknown component", false;
        c->td = d[2 + 2 * i] >> 4;
        c->ta = d[2 + 2 * i] & 15;
        if (c->td > 3 || c->ta > 3 || !dc[c->td].present || !ac[c->ta].present || !qt_ok[c->tq])
          return err = "JPEG scan without its tables", false;
        c->pred = 0;
        sc.push_back(c);
      }
      if (d[1 + 2 * ns] != 0 || d[2 + 2 * ns] != 63 || d[3 + 2 * ns] != 0) return err = "not a sequential scan", false;
      Bits bits{p, n, off + 2 + len};
      uint32_t mx_n, my_n;
      if (ns == 1) {  // non-interleaved: the component's own blocks in raster order
        mx_n = (sc[0]->dw + 7) / 8;
        my_n = (sc[0]->dh + 7) / 8;
      } else {
        mx_n = (W + 8 * hmax - 1) / (8 * hmax);
        my_n = (H + 8 * vmax - 1) / (8 * vmax);
      }
      int32_t coef[64];
      const auto block = [&](Comp& c, uint32_t bx, uint32_t by) -> bool {
        std::memset(coef, 0, sizeof(coef));
        const int s = decode_sym(bits, dc[c.td]);
        if (s < 0 || s > 11) return false;
        c.pred += extend(bits.get(s), s);
        const uint16_t* q = qt[c.tq];
        coef[0] = c.pred * q[0];
        for (int k = 1; k < 64;) {
          const int rs = decode_sym(bits, ac[c.ta]);
          if (rs < 0) return false;
          const int r = rs >> 4, sz = rs & 15;
          if (sz == 0) {
            if (r != 15) break;  // EOB
            k += 16;
            continue;
          }
          k += r;
          if (k > 63) return false;
          coef[ZIGZAG[k]] = extend(bits.get(sz), sz) * q[ZIGZAG[k]];
          ++k;
        }
        const size_t pitch = (size_t)c.bw * 8;
        idct_islow(coef, c.plane.data() + (size_t)by * 8 * pitch + (size_t)bx * 8, pitch);
        return true;
      };
      uint32_t mcu = 0;
      for (uint32_t my = 0; my < my_n; ++my)
        for (uint32_t mx = 0; mx < mx_n; ++mx, ++mcu) {
          if (restart && mcu > 0 && mcu % restart == 0) {
            if (!bits.restart()) return err = "missing JPEG restart marker", false;
            for (Comp* c : sc) c->pred = 0;
          }
          bool ok = true;
          if (ns == 1) {
            ok = block(*sc[0], mx, my);
          } else {
            for (Comp* c : sc)
              for (int v = 0; v < c->v && ok; ++v)
                for (int h = 0; h < c->h && ok; ++h) ok = block(*c, mx * c->h + h, my * c->v + v);
          }
          if (!ok) return err = "corrupt JPEG scan data", false;
        }
      if (bits.over()) return err = "truncated JPEG scan data", false;
      // next marker after the scan's data
      size_t k = bits.pos;
      while (k + 1 < n && !(p[k] == 0xFF && p[k + 1] != 0x00 && !(p[k + 1] >= 0xD0 && p[k + 1] <= 0xD7))) ++k;
      off = k;
      continue;
    }
    off += 2 + len;
  }
  if (!frame) return err = "JPEG without a frame", false;
  // upsample to full resolution and convert
  std::vector<std::vector<uint8_t>> full;
  std::vector<uint32_t> fw;
  for (Comp& c : comps) {
    uint32_t wo, ho;
    if (hmax % c.h || vmax % c.v) return err = "unsupported JPEG sampling factors", false;
    full.push_back(upsample(c, hmax / c.h, vmax / c.v, wo, ho));
    fw.push_back(wo);
  }
  rgba.assign((size_t)W * H, 0);
  const bool rgb_direct = comps.size() == 3 && !jfif &&
                          ((adobe && adobe_transform == 0) ||
                           (!adobe && comps[0].id == 'R' && comps[1].id == 'G' && comps[2].id == 'B'));
  constexpr int SB = 16;
  const auto fix = [](double x) { return (int32_t)(x * (1 << SB) + 0.5); };
  int32_t cr_r[256], cb_b[256], cr_g[256], cb_g[256];
  for (int i = 0; i < 256; ++i) {
    const int32_t x = i - 128;
    cr_r[i] = (fix(1.40200) * x + (1 << (SB - 1))) >> SB;
    cb_b[i] = (fix(1.77200) * x + (1 << (SB - 1))) >> SB;
    cr_g[i] = -fix(0.71414) * x;
    cb_g[i] = -fix(0.34414) * x + (1 << (SB - 1));
  }
  const auto clamp8 = [](int v) { return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); };
  for (uint32_t y = 0; y < H; ++y)
    for (uint32_t x = 0; x < W; ++x) {
      uint32_t r, g, b;
      const int c0 = full[0][(size_t)y * fw[0] + x];
      if (comps.size() == 1) {
        r = g = b = (uint32_t)c0;
      } else {
        const int c1 = full[1][(size_t)y * fw[1] + x], c2 = full[2][(size_t)y * fw[2] + x];
        if (rgb_direct) {
          r = (uint32_t)c0;
          g = (uint32_t)c1;
          b = (uint32_t)c2;
        } else {
          r = clamp8(c0 + cr_r[c2]);
          g = clamp8(c0 + ((cb_g[c1] + cr_g[c2]) >> SB));
          b = clamp8(c0 + cb_b[c1]);
        }
      }
      rgba[(size_t)y * W + x] = r | (g << 8) | (b << 16) | 0xFF000000u;
    }
  return true;
}

}  // namespace apt
