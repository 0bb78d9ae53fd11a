// Oracle (test infrastructure only): baseline JPEG compression of one 64x64
// RGBX block, the step the reference node runs on every block it sends
// (turbojpeg::compress(image, 95, Subsamp::Sub2x2),
// crates/appearance-render-loop/src/node.rs:65-79, with
// NODE_PIXEL_FORMAT = RGBX, host.rs:22).
//
// The turbojpeg crate (Cargo.toml:57, version 1.2.1) wraps libjpeg-turbo,
// which is not part of the reference.  This restates the published
// algorithm of libjpeg-turbo's default compression path (TurboJPEG 3 with
// default parameters: accurate integer DCT, no Huffman optimisation, JFIF
// header) so that the bytes match the library's:
//   - colour conversion (ITU-T T.871 / JFIF) in 16-bit fixed point, Y rounded
//     to nearest, Cb and Cr with a rounding offset of one half minus one ulp;
//   - 2x2 box downsampling of Cb and Cr with a bias alternating 1, 2 along
//     each output row;
//   - the 8x8 forward DCT of Loeffler, Ligtenberg and Moschytz (ICASSP 1989)
//     in 13-bit fixed point with 2 extra bits between the passes, outputs
//     scaled by 8;
//   - quantisation by the Annex K tables scaled for quality, rounding half
//     away from zero (divisor = 8 x table entry);
//   - Huffman coding with the Annex K tables (ITU-T T.81 F.1.2), the last
//     byte padded with 1 bits, 0xFF bytes stuffed with 0x00;
//   - the markers SOI, APP0 (JFIF 1.01, 1:1 density), one DQT per table, SOF0,
//     one DHT per table (luma DC, luma AC, chroma DC, chroma AC), SOS, EOI.
// Parity of the restatement is pinned by comparing whole files with Pillow's
// encoder, which bundles libjpeg-turbo (tests/test_oracle_jpeg.py).
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                             12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                             35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                             58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// T.81 Annex K: K.1 luminance and K.2 chrominance quantisation tables
// (row-major natural order) and the K.3 / K.5 Huffman tables (counts of
// codes of length 1..16, then the symbols in code order).
const uint8_t kQLuma[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                            14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                            18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                            49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
const uint8_t kQChroma[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                              24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                              99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                              99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
const uint8_t kDcLumaBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcChromaBits[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcLumaBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcChromaBits[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};

// The AC symbols (run << 4 | size) of K.5 in code order.
std::vector<uint8_t> ac_symbols(bool chroma) {
  static const uint8_t luma_prefix[] = {0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06,
                                        0x13, 0x51, 0x61, 0x07, 0x22, 0x71, 0x14, 0x32, 0x81, 0x91, 0xa1, 0x08,
                                        0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
                                        0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28,
                                        0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45,
                                        0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
                                        0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75,
                                        0x76, 0x77, 0x78, 0x79, 0x7a, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89,
                                        0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
                                        0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6,
                                        0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9,
                                        0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
                                        0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4,
                                        0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
  static const uint8_t chroma_prefix[] = {
      0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13,
      0x22, 0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62,
      0x72, 0xd1, 0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29,
      0x2a, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54,
      0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75,
      0x76, 0x77, 0x78, 0x79, 0x7a, 0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94,
      0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3,
      0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2,
      0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea,
      0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
  return chroma ? std::vector<uint8_t>(chroma_prefix, chroma_prefix + 162)
                : std::vector<uint8_t>(luma_prefix, luma_prefix + 162);
}

struct Code {
  uint16_t code = 0;
  uint8_t len = 0;
};

// Canonical codes from a (counts, symbols) table (T.81 C.2).
void canonical(const uint8_t* bits, const std::vector<uint8_t>& vals, Code* table) {
  uint32_t code = 0;
  size_t k = 0;
  for (int len = 1; len <= 16; ++len) {
    for (int i = 0; i < bits[len - 1]; ++i, ++k) table[vals[k]] = Code{(uint16_t)code++, (uint8_t)len};
    code <<= 1;
  }
}

int scaled_q(int base, int quality) {
  const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
  long t = ((long)base * scale + 50) / 100;
  if (t < 1) t = 1;
  if (t > 255) t = 255;  // baseline: 8-bit tables
  return (int)t;
}

struct BitWriter {
  std::vector<uint8_t>& out;
  uint32_t acc = 0;
  int n = 0;
  void put(uint32_t v, int len) {
    for (int i = len - 1; i >= 0; --i) {
      acc = (acc << 1) | ((v >> i) & 1u);
      if (++n == 8) {
        out.push_back((uint8_t)acc);
        if (acc == 0xFF) out.push_back(0);
        acc = 0;
        n = 0;
      }
    }
  }
  void flush() {
    if (n) put(0x7F, 8 - n);
  }
};

int nbits(int v) {
  int a = v < 0 ? -v : v, n = 0;
  while (a) {
    ++n;
    a >>= 1;
  }
  return n;
}

// Arithmetic right shift with rounding (x + 2^(n-1)) >> n.
inline int32_t descale(int32_t x, int n) { return (x + (1 << (n - 1))) >> n; }

// The accurate integer forward DCT (Loeffler-Ligtenberg-Moschytz), in place
// on 64 level-shifted samples in natural order; outputs are 8x the
// orthonormal DCT.
void fdct_islow(int32_t* d) {
  constexpr int CB = 13, P1 = 2;
  const auto fix = [](double x) { return (int32_t)(x * (1 << CB) + 0.5); };
  const int32_t c0_298 = fix(0.298631336), c0_390 = fix(0.390180644), c0_541 = fix(0.541196100),
                c0_765 = fix(0.765366865), c0_899 = fix(0.899976223), c1_175 = fix(1.175875602),
                c1_501 = fix(1.501321110), c1_847 = fix(1.847759065), c1_961 = fix(1.961570560),
                c2_053 = fix(2.053119869), c2_562 = fix(2.562915447), c3_072 = fix(3.072711026);
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 0; i < 8; ++i) {
      // pass 0: rows (stride 1), pass 1: columns (stride 8)
      int32_t* p = pass == 0 ? d + i * 8 : d + i;
      const int st = pass == 0 ? 1 : 8;
      const int32_t t0 = p[0] + p[7 * st], t7 = p[0] - p[7 * st];
      const int32_t t1 = p[st] + p[6 * st], t6 = p[st] - p[6 * st];
      const int32_t t2 = p[2 * st] + p[5 * st], t5 = p[2 * st] - p[5 * st];
      const int32_t t3 = p[3 * st] + p[4 * st], t4 = p[3 * st] - p[4 * st];
      const int32_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
      const int sh = pass == 0 ? CB - P1 : CB + P1;
      int32_t o[8];
      if (pass == 0) {
        o[0] = (t10 + t11) * (1 << P1);
        o[4] = (t10 - t11) * (1 << P1);
      } else {
        o[0] = descale(t10 + t11, P1);
        o[4] = descale(t10 - t11, P1);
      }
      const int32_t z1 = (t12 + t13) * c0_541;
      o[2] = descale(z1 + t13 * c0_765, sh);
      o[6] = descale(z1 - t12 * c1_847, sh);
      // odd part
      const int32_t y1 = t4 + t7, y2 = t5 + t6, y3 = t4 + t6, y4 = t5 + t7;
      const int32_t z5 = (y3 + y4) * c1_175;
      const int32_t a4 = t4 * c0_298, a5 = t5 * c2_053, a6 = t6 * c3_072, a7 = t7 * c1_501;
      const int32_t m1 = -y1 * c0_899, m2 = -y2 * c2_562;
      const int32_t m3 = -y3 * c1_961 + z5, m4 = -y4 * c0_390 + z5;
      o[7] = descale(a4 + m1 + m3, sh);
      o[5] = descale(a5 + m2 + m4, sh);
      o[3] = descale(a6 + m2 + m3, sh);
      o[1] = descale(a7 + m1 + m4, sh);
      for (int k = 0; k < 8; ++k) p[k * st] = o[k];
    }
  }
}

void put16(std::vector<uint8_t>& h, uint32_t v) {
  h.push_back((uint8_t)(v >> 8));
  h.push_back((uint8_t)v);
}

}  // namespace

extern "C" size_t or_jpeg_block(const uint8_t* rgbx, int quality, uint8_t* out, size_t cap) {
  if (quality < 1 || quality > 100) return 0;
  // colour conversion, 16-bit fixed point
  constexpr int SB = 16;
  const auto fix = [](double x) { return (int32_t)(x * (1 << SB) + 0.5); };
  const int32_t half = 1 << (SB - 1), cbcr_off = 128 << SB;
  int Y[64 * 64], Cb[64 * 64], Cr[64 * 64];
  for (int i = 0; i < 64 * 64; ++i) {
    const int32_t r = rgbx[4 * i], g = rgbx[4 * i + 1], b = rgbx[4 * i + 2];
    Y[i] = (fix(0.29900) * r + fix(0.58700) * g + fix(0.11400) * b + half) >> SB;
    Cb[i] = (-fix(0.16874) * r - fix(0.33126) * g + fix(0.50000) * b + cbcr_off + half - 1) >> SB;
    Cr[i] = (fix(0.50000) * r - fix(0.41869) * g - fix(0.08131) * b + cbcr_off + half - 1) >> SB;
  }
  int Cb2[32 * 32], Cr2[32 * 32];
  for (int y = 0; y < 32; ++y)
    for (int x = 0; x < 32; ++x) {
      const int bias = (x & 1) ? 2 : 1;
      const int i = 2 * y * 64 + 2 * x;
      Cb2[y * 32 + x] = (Cb[i] + Cb[i + 1] + Cb[i + 64] + Cb[i + 65] + bias) >> 2;
      Cr2[y * 32 + x] = (Cr[i] + Cr[i + 1] + Cr[i + 64] + Cr[i + 65] + bias) >> 2;
    }
  int q[2][64];
  for (int i = 0; i < 64; ++i) {
    q[0][i] = scaled_q(kQLuma[i], quality);
    q[1][i] = scaled_q(kQChroma[i], quality);
  }
  Code dc[2][12], ac[2][256];
  canonical(kDcLumaBits, std::vector<uint8_t>(kDcVals, kDcVals + 12), dc[0]);
  canonical(kDcChromaBits, std::vector<uint8_t>(kDcVals, kDcVals + 12), dc[1]);
  canonical(kAcLumaBits, ac_symbols(false), ac[0]);
  canonical(kAcChromaBits, ac_symbols(true), ac[1]);

  std::vector<uint8_t> f = {0xFF, 0xD8, 0xFF, 0xE0, 0x00, 0x10, 'J',  'F',  'I',  'F',
                            0x00, 0x01, 0x01, 0x00, 0x00, 0x01, 0x00, 0x01, 0x00, 0x00};
  for (int t = 0; t < 2; ++t) {
    f.insert(f.end(), {0xFF, 0xDB, 0x00, 0x43, (uint8_t)t});
    for (int k = 0; k < 64; ++k) f.push_back((uint8_t)q[t][kZigzag[k]]);
  }
  f.insert(f.end(), {0xFF, 0xC0, 0x00, 0x11, 0x08, 0x00, 0x40, 0x00, 0x40, 0x03, 0x01, 0x22, 0x00, 0x02, 0x11,
                     0x01, 0x03, 0x11, 0x01});
  const auto dht = [&](uint8_t id, const uint8_t* bits, const std::vector<uint8_t>& vals) {
    f.insert(f.end(), {0xFF, 0xC4});
    put16(f, (uint32_t)(3 + 16 + vals.size()));
    f.push_back(id);
    f.insert(f.end(), bits, bits + 16);
    f.insert(f.end(), vals.begin(), vals.end());
  };
  const std::vector<uint8_t> dcv(kDcVals, kDcVals + 12);
  dht(0x00, kDcLumaBits, dcv);
  dht(0x10, kAcLumaBits, ac_symbols(false));
  dht(0x01, kDcChromaBits, dcv);
  dht(0x11, kAcChromaBits, ac_symbols(true));
  f.insert(f.end(), {0xFF, 0xDA, 0x00, 0x0C, 0x03, 0x01, 0x00, 0x02, 0x11, 0x03, 0x11, 0x00, 0x3F, 0x00});

  BitWriter bw{f};
  int last_dc[3] = {0, 0, 0};
  const auto code_block = [&](const int* plane, int pitch, int x0, int y0, int comp) {
    int32_t d[64];
    for (int y = 0; y < 8; ++y)
      for (int x = 0; x < 8; ++x) d[y * 8 + x] = plane[(y0 + y) * pitch + x0 + x] - 128;
    fdct_islow(d);
    const int t = comp ? 1 : 0;
    int c[64];
    for (int i = 0; i < 64; ++i) {
      const int32_t div = q[t][i] * 8, v = d[i];
      c[i] = v < 0 ? -((-v + (div >> 1)) / div) : (v + (div >> 1)) / div;
    }
    const int diff = c[0] - last_dc[comp];
    last_dc[comp] = c[0];
    const int s = nbits(diff);
    bw.put(dc[t][s].code, dc[t][s].len);
    if (s) bw.put((uint32_t)(diff < 0 ? diff - 1 : diff), s);
    int run = 0;
    for (int k = 1; k < 64; ++k) {
      const int v = c[kZigzag[k]];
      if (v == 0) {
        ++run;
        continue;
      }
      for (; run > 15; run -= 16) bw.put(ac[t][0xF0].code, ac[t][0xF0].len);
      const int sz = nbits(v);
      const Code cd = ac[t][(run << 4) | sz];
      bw.put(cd.code, cd.len);
      bw.put((uint32_t)(v < 0 ? v - 1 : v), sz);
      run = 0;
    }
    if (run) bw.put(ac[t][0].code, ac[t][0].len);
  };
  for (int my = 0; my < 4; ++my)
    for (int mx = 0; mx < 4; ++mx) {
      for (int v = 0; v < 2; ++v)
        for (int h = 0; h < 2; ++h) code_block(Y, 64, 16 * mx + 8 * h, 16 * my + 8 * v, 0);
      code_block(Cb2, 32, 8 * mx, 8 * my, 1);
      code_block(Cr2, 32, 8 * mx, 8 * my, 2);
    }
  bw.flush();
  f.push_back(0xFF);
  f.push_back(0xD9);
  if (f.size() > cap) return 0;
  std::memcpy(out, f.data(), f.size());
  return f.size();
}
