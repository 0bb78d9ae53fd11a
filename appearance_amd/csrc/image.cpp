// Image decoding for glTF textures: PNG (zlib inflate, RFC 1950/1951; the
// five scanline filters; grey, grey + alpha, RGB, RGBA and palette images at
// 1-16 bits; Adam7 interlacing) into RGBA8 (baseline JPEG is in
// jpeg_decode.cpp); and Radiance HDR (RGBE) sky
// panoramas into linear RGB floats.  The reference decodes images with the
// `image` crate and compresses them to BC formats for the GPU
// (crates/appearance-texture/src/lib.rs:207-342); here texels stay as
// decoded so the samplers see the decoded values.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "apt_internal.h"

namespace apt {
namespace {

// ---- inflate -----------------------------------------------------------------
struct BitIn {
  const uint8_t* p;
  size_t n, pos = 0;
  uint32_t acc = 0;
  int cnt = 0;
  bool over = false;
  uint32_t bits(int k) {  // LSB first
    while (cnt < k) {
      uint32_t b = 0;
      if (pos < n) b = p[pos++];
      else over = true;
      acc |= b << cnt;
      cnt += 8;
    }
    const uint32_t v = acc & ((1u << k) - 1u);
    acc >>= k;
    cnt -= k;
    return v;
  }
  void align() {
    acc = 0;
    cnt = 0;
  }
};

// Canonical Huffman code: counts per length and symbols in code order.
struct Huff {
  uint16_t count[16] = {0};
  std::vector<uint16_t> sym;
  bool build(const uint8_t* lens, int n) {
    std::memset(count, 0, sizeof(count));
    for (int i = 0; i < n; ++i) count[lens[i]]++;
    count[0] = 0;
    int left = 1;
    for (int l = 1; l < 16; ++l) {
      left <<= 1;
      left -= count[l];
      if (left < 0) return false;  // over-subscribed
    }
    uint16_t offs[16];
    offs[1] = 0;
    for (int l = 1; l < 15; ++l) offs[l + 1] = (uint16_t)(offs[l] + count[l]);
    sym.assign((size_t)n, 0);
    for (int i = 0; i < n; ++i)
      if (lens[i]) sym[offs[lens[i]]++] = (uint16_t)i;
    return true;
  }
  int decode(BitIn& in) const {
    int code = 0, first = 0, index = 0;
    for (int l = 1; l < 16; ++l) {
      code |= (int)in.bits(1);
      const int c = count[l];
      if (code - c < first) return sym[(size_t)(index + (code - first))];
      index += c;
      first += c;
      first <<= 1;
      code <<= 1;
    }
    return -1;
  }
};

const uint16_t LEN_BASE[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                               35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t DIST_BASE[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193,
                                257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

bool inflate_block(BitIn& in, const Huff& lit, const Huff& dist, std::vector<uint8_t>& out, size_t limit) {
  while (true) {
    const int s = lit.decode(in);
    if (s < 0 || in.over) return false;
    if (s < 256) {
      out.push_back((uint8_t)s);
    } else if (s == 256) {
      return true;
    } else {
      const int li = s - 257;
      if (li >= 29) return false;
      const size_t len = LEN_BASE[li] + in.bits(LEN_EXTRA[li]);
      const int ds = dist.decode(in);
      if (ds < 0 || ds >= 30) return false;
      const size_t d = DIST_BASE[ds] + in.bits(DIST_EXTRA[ds]);
      if (d > out.size()) return false;
      const size_t from = out.size() - d;
      for (size_t k = 0; k < len; ++k) out.push_back(out[from + k]);
    }
    if (out.size() > limit) return false;
  }
}

// zlib stream (RFC 1950) -> bytes; at most `limit` bytes are produced.
bool zlib_inflate(const uint8_t* p, size_t n, std::vector<uint8_t>& out, size_t limit) {
  if (n < 2 || (p[0] & 0x0F) != 8 || ((p[0] << 8) | p[1]) % 31 != 0 || (p[1] & 0x20)) return false;
  BitIn in{p + 2, n - 2};
  bool last = false;
  while (!last) {
    last = in.bits(1) != 0;
    const uint32_t type = in.bits(2);
    if (type == 0) {
      in.align();
      if (in.pos + 4 > in.n) return false;
      const uint32_t len = in.p[in.pos] | (in.p[in.pos + 1] << 8);
      const uint32_t nlen = in.p[in.pos + 2] | (in.p[in.pos + 3] << 8);
      in.pos += 4;
      if ((len ^ 0xFFFFu) != nlen || in.pos + len > in.n) return false;
      out.insert(out.end(), in.p + in.pos, in.p + in.pos + len);
      in.pos += len;
    } else if (type == 1) {
      uint8_t l[288], d[30];
      for (int i = 0; i < 144; ++i) l[i] = 8;
      for (int i = 144; i < 256; ++i) l[i] = 9;
      for (int i = 256; i < 280; ++i) l[i] = 7;
      for (int i = 280; i < 288; ++i) l[i] = 8;
      for (int i = 0; i < 30; ++i) d[i] = 5;
      Huff lit, dist;
      lit.build(l, 288);
      dist.build(d, 30);
      if (!inflate_block(in, lit, dist, out, limit)) return false;
    } else if (type == 2) {
      const int hlit = (int)in.bits(5) + 257, hdist = (int)in.bits(5) + 1, hclen = (int)in.bits(4) + 4;
      static const uint8_t order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
      uint8_t cl[19] = {0};
      for (int i = 0; i < hclen; ++i) cl[order[i]] = (uint8_t)in.bits(3);
      Huff clh;
      if (!clh.build(cl, 19)) return false;
      uint8_t lens[320] = {0};
      int k = 0;
      while (k < hlit + hdist) {
        const int s = clh.decode(in);
        if (s < 0 || in.over) return false;
        if (s < 16) {
          lens[k++] = (uint8_t)s;
        } else {
          int rep = 0;
          uint8_t v = 0;
          if (s == 16) {
            if (k == 0) return false;
            v = lens[k - 1];
            rep = 3 + (int)in.bits(2);
          } else if (s == 17) {
            rep = 3 + (int)in.bits(3);
          } else {
            rep = 11 + (int)in.bits(7);
          }
          if (k + rep > hlit + hdist) return false;
          while (rep--) lens[k++] = v;
        }
      }
      Huff lit, dist;
      if (!lit.build(lens, hlit) || !dist.build(lens + hlit, hdist)) return false;
      if (!inflate_block(in, lit, dist, out, limit)) return false;
    } else {
      return false;
    }
    if (in.over || out.size() > limit) return false;
  }
  return true;
}

// ---- PNG ---------------------------------------------------------------------
uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

uint8_t paeth(int a, int b, int c) {
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  return (uint8_t)(pa <= pb && pa <= pc ? a : pb <= pc ? b : c);
}

// Undoes the scanline filters of a w x h (sub)image in place; bpp = bytes per
// pixel (at least 1), stride = bytes per row.  Returns false on a bad filter.
bool unfilter(uint8_t* data, uint32_t h, size_t stride, size_t bpp, std::vector<uint8_t>& rows) {
  rows.assign(stride * h, 0);
  const uint8_t* src = data;
  for (uint32_t y = 0; y < h; ++y) {
    const uint8_t f = *src++;
    uint8_t* cur = rows.data() + y * stride;
    const uint8_t* prev = y ? cur - stride : nullptr;
    for (size_t x = 0; x < stride; ++x) {
      const int a = x >= bpp ? cur[x - bpp] : 0, b = prev ? prev[x] : 0, c = (prev && x >= bpp) ? prev[x - bpp] : 0;
      int v = src[x];
      switch (f) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += paeth(a, b, c); break;
        default: return false;
      }
      cur[x] = (uint8_t)v;
    }
    src += stride;
  }
  return true;
}

}  // namespace

bool png_decode(const uint8_t* p, size_t n, uint32_t& w, uint32_t& h, std::vector<uint32_t>& rgba, std::string& err) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (n < 8 || std::memcmp(p, sig, 8) != 0) return err = "not a PNG", false;
  size_t off = 8;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte, trns;
  w = h = 0;
  while (off + 12 <= n) {
    const uint32_t len = be32(p + off);
    const uint8_t* type = p + off + 4;
    const uint8_t* d = p + off + 8;
    if (off + 12 + (size_t)len > n) return err = "truncated PNG chunk", false;
    if (!std::memcmp(type, "IHDR", 4) && len >= 13) {
      w = be32(d);
      h = be32(d + 4);
      depth = d[8];
      ctype = d[9];
      interlace = d[12];
      if (d[10] != 0 || d[11] != 0) return err = "unsupported PNG compression or filter method", false;
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(d, d + len);
    } else if (!std::memcmp(type, "tRNS", 4)) {
      trns.assign(d, d + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), d, d + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      break;
    }
    off += 12 + (size_t)len;
  }
  if (w == 0 || h == 0 || w > 16384 || h > 16384) return err = "bad PNG size", false;
  const int chans = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
  if (!chans) return err = "unsupported PNG colour type", false;
  const bool depth_ok = depth == 8 || depth == 16 || ((ctype == 0 || ctype == 3) && (depth == 1 || depth == 2 || depth == 4));
  if (!depth_ok || (ctype == 3 && depth == 16)) return err = "unsupported PNG bit depth", false;
  if (ctype == 3 && plte.size() < 3) return err = "palette PNG without PLTE", false;
  const size_t bits_px = (size_t)chans * depth;
  const size_t bpp = std::max<size_t>(1, bits_px / 8);
  // decompressed size bound: every pass's rows plus their filter bytes
  const size_t limit = ((size_t)w * bits_px + 7) / 8 * h + h + 64 + (interlace ? 7 * (size_t)h + 64 : 0);
  std::vector<uint8_t> raw;
  if (!zlib_inflate(idat.data(), idat.size(), raw, limit * 2 + 1024)) return err = "corrupt PNG image data", false;
  rgba.assign((size_t)w * h, 0u);
  // sample k of a row -> 8-bit value (and the 16-bit value for tRNS checks)
  const auto sample = [&](const uint8_t* row, size_t k, uint32_t& raw16) -> uint8_t {
    if (depth == 16) {
      raw16 = (uint32_t)row[2 * k] << 8 | row[2 * k + 1];
      return row[2 * k];
    }
    if (depth == 8) {
      raw16 = row[k];
      return row[k];
    }
    const size_t bit = k * depth;
    const uint32_t v = (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1u);
    raw16 = v;
    return ctype == 3 ? (uint8_t)v : (uint8_t)(v * 255u / ((1u << depth) - 1u));
  };
  const auto put = [&](uint32_t x, uint32_t y, const uint8_t* row, uint32_t i) {
    uint32_t r16 = 0, g16 = 0, b16 = 0, a16 = 0;
    uint8_t r, g, b, a = 255;
    if (ctype == 3) {
      const uint8_t idx = sample(row, i, r16);
      if ((size_t)idx * 3 + 2 >= plte.size()) return;
      r = plte[3 * idx];
      g = plte[3 * idx + 1];
      b = plte[3 * idx + 2];
      if (idx < trns.size()) a = trns[idx];
    } else if (ctype == 0 || ctype == 4) {
      r = g = b = sample(row, (size_t)i * chans, r16);
      if (ctype == 4) a = sample(row, (size_t)i * chans + 1, a16);
      else if (trns.size() >= 2 && r16 == ((uint32_t)trns[0] << 8 | trns[1])) a = 0;
    } else {
      r = sample(row, (size_t)i * chans, r16);
      g = sample(row, (size_t)i * chans + 1, g16);
      b = sample(row, (size_t)i * chans + 2, b16);
      if (ctype == 6) a = sample(row, (size_t)i * chans + 3, a16);
      else if (trns.size() >= 6 && r16 == ((uint32_t)trns[0] << 8 | trns[1]) &&
               g16 == ((uint32_t)trns[2] << 8 | trns[3]) && b16 == ((uint32_t)trns[4] << 8 | trns[5]))
        a = 0;
    }
    rgba[(size_t)y * w + x] = (uint32_t)r | (uint32_t)g << 8 | (uint32_t)b << 16 | (uint32_t)a << 24;
  };
  std::vector<uint8_t> rows;
  size_t pos = 0;
  static const int X0[7] = {0, 4, 0, 2, 0, 1, 0}, Y0[7] = {0, 0, 4, 0, 2, 0, 1};
  static const int DX[7] = {8, 8, 4, 4, 2, 2, 1}, DY[7] = {8, 8, 8, 4, 4, 2, 2};
  const int passes = interlace ? 7 : 1;
  for (int ps = 0; ps < passes; ++ps) {
    const uint32_t x0 = interlace ? X0[ps] : 0, y0 = interlace ? Y0[ps] : 0;
    const uint32_t dx = interlace ? DX[ps] : 1, dy = interlace ? DY[ps] : 1;
    const uint32_t pw = w > x0 ? (w - x0 + dx - 1) / dx : 0, ph = h > y0 ? (h - y0 + dy - 1) / dy : 0;
    if (!pw || !ph) continue;
    const size_t stride = ((size_t)pw * bits_px + 7) / 8;
    if (pos + (stride + 1) * ph > raw.size()) return err = "PNG image data too short", false;
    if (!unfilter(raw.data() + pos, ph, stride, bpp, rows)) return err = "bad PNG filter", false;
    pos += (stride + 1) * ph;
    for (uint32_t y = 0; y < ph; ++y)
      for (uint32_t x = 0; x < pw; ++x) put(x0 + x * dx, y0 + y * dy, rows.data() + y * stride, x);
  }
  return true;
}

// ---- Radiance HDR (RGBE) ---------------------------------------------------------
// Header lines up to an empty line (the magic "#?RADIANCE" or "#?RGBE", an
// optional FORMAT=32-bit_rle_rgbe), then the resolution line; scanlines are
// flat 4-byte texels, or, for widths 8..32767, either run-length encoded per
// channel (a 2, 2, width-hi, width-lo marker, then for each of the four
// channels runs of count > 128 repeating one byte count - 128 times and
// literal runs of count bytes) or, in the old style, flat with (1, 1, 1, n)
// markers repeating the previous texel n << shift times (shift grows by 8
// over consecutive markers).
bool hdr_decode(const uint8_t* p, size_t n, uint32_t& w, uint32_t& h, std::vector<float>& rgb, std::string& err) {
  size_t pos = 0;
  auto line = [&](std::string& out) {
    out.clear();
    while (pos < n && p[pos] != '\n') out.push_back((char)p[pos++]);
    if (pos >= n) return false;
    ++pos;
    return true;
  };
  std::string l;
  if (!line(l) || (l.rfind("#?RADIANCE", 0) != 0 && l.rfind("#?RGBE", 0) != 0))
    return err = "not a Radiance HDR file", false;
  while (true) {
    if (!line(l)) return err = "HDR header not terminated", false;
    if (l.empty()) break;
    if (l.rfind("FORMAT=", 0) == 0 && l != "FORMAT=32-bit_rle_rgbe")
      return err = "unsupported HDR pixel format " + l.substr(7), false;
  }
  if (!line(l)) return err = "HDR resolution line missing", false;
  long hh = 0, ww = 0;
  char ya[3] = {0}, xa[3] = {0};
  if (std::sscanf(l.c_str(), "%2s %ld %2s %ld", ya, &hh, xa, &ww) != 4 || std::string(ya) != "-Y" ||
      std::string(xa) != "+X")
    return err = "unsupported HDR orientation '" + l + "' (only -Y h +X w)", false;
  if (hh <= 0 || ww <= 0 || hh > 65536 || ww > 65536) return err = "bad HDR size", false;
  w = (uint32_t)ww;
  h = (uint32_t)hh;
  rgb.assign((size_t)w * h * 3, 0.0f);
  std::vector<uint8_t> sl((size_t)w * 4);
  auto to_float = [](const uint8_t* e, float* out) {
    if (e[3] == 0) {
      out[0] = out[1] = out[2] = 0.0f;
      return;
    }
    const float f = std::ldexp(1.0f, (int)e[3] - (128 + 8));
    for (int c = 0; c < 3; ++c) out[c] = (float)e[c] * f;
  };
  for (uint32_t y = 0; y < h; ++y) {
    const bool rle = w >= 8 && w <= 0x7fff && pos + 4 <= n && p[pos] == 2 && p[pos + 1] == 2 && !(p[pos + 2] & 0x80);
    if (rle) {
      if ((((uint32_t)p[pos + 2] << 8) | p[pos + 3]) != w) return err = "HDR scanline width mismatch", false;
      pos += 4;
      for (int c = 0; c < 4; ++c) {
        uint32_t x = 0;
        while (x < w) {
          if (pos >= n) return err = "HDR data truncated", false;
          uint32_t cnt = p[pos++];
          if (cnt > 128) {
            cnt -= 128;
            if (cnt > w - x || pos >= n) return err = "bad HDR run", false;
            const uint8_t v = p[pos++];
            for (uint32_t k = 0; k < cnt; ++k) sl[(size_t)(x++) * 4 + c] = v;
          } else {
            if (cnt == 0 || cnt > w - x || pos + cnt > n) return err = "bad HDR run", false;
            for (uint32_t k = 0; k < cnt; ++k) sl[(size_t)(x++) * 4 + c] = p[pos++];
          }
        }
      }
    } else {
      uint32_t x = 0;
      int shift = 0;
      while (x < w) {
        if (pos + 4 > n) return err = "HDR data truncated", false;
        const uint8_t* e = p + pos;
        pos += 4;
        if (e[0] == 1 && e[1] == 1 && e[2] == 1) {  // old-style run of the previous texel
          if (x == 0) return err = "HDR run at the start of a scanline", false;
          if (shift > 24) return err = "bad HDR run", false;
          const uint64_t cnt = (uint64_t)e[3] << shift;
          if (cnt > w - x) return err = "bad HDR run", false;
          for (uint64_t k = 0; k < cnt; ++k, ++x) std::memcpy(&sl[(size_t)x * 4], &sl[(size_t)(x - 1) * 4], 4);
          shift += 8;
        } else {
          std::memcpy(&sl[(size_t)(x++) * 4], e, 4);
          shift = 0;
        }
      }
    }
    for (uint32_t x = 0; x < w; ++x) to_float(&sl[(size_t)x * 4], &rgb[((size_t)y * w + x) * 3]);
  }
  return true;
}

}  // namespace apt

using namespace apt;

apt_status apt_image_decode_hdr(const uint8_t* bytes, size_t len, uint32_t* width, uint32_t* height,
                                float** out_rgb) {
  clear_error();
  APT_REQUIRE(bytes && width && height && out_rgb, APT_ERR_INVALID_ARG, "apt_image_decode_hdr: null argument");
  std::vector<float> rgb;
  std::string err;
  uint32_t w = 0, h = 0;
  APT_REQUIRE(hdr_decode(bytes, len, w, h, rgb, err), APT_ERR_INVALID_ARG, "apt_image_decode_hdr: %s",
              err.c_str());
  float* out = static_cast<float*>(std::malloc(rgb.size() * sizeof(float)));
  APT_REQUIRE(out, APT_ERR_OOM, "apt_image_decode_hdr: out of host memory");
  std::memcpy(out, rgb.data(), rgb.size() * sizeof(float));
  *width = w;
  *height = h;
  *out_rgb = out;
  return APT_OK;
}

void apt_image_free(float* rgb) { std::free(rgb); }

apt_status apt_image_decode_rgba8(const uint8_t* bytes, size_t len, uint32_t* width, uint32_t* height,
                                  uint8_t* out_rgba, size_t out_len) {
  clear_error();
  APT_REQUIRE(bytes && width && height, APT_ERR_INVALID_ARG, "apt_image_decode_rgba8: null argument");
  std::vector<uint32_t> rgba;
  std::string err_png, err_jpeg;
  uint32_t w = 0, h = 0;
  const bool ok = png_decode(bytes, len, w, h, rgba, err_png) || jpeg_decode(bytes, len, w, h, rgba, err_jpeg);
  APT_REQUIRE(ok, APT_ERR_INVALID_ARG, "apt_image_decode_rgba8: %s",
              (len >= 2 && bytes[0] == 0xFF && bytes[1] == 0xD8 ? err_jpeg : err_png).c_str());
  *width = w;
  *height = h;
  if (out_rgba) {
    APT_REQUIRE(out_len >= rgba.size() * 4, APT_ERR_INVALID_ARG,
                "apt_image_decode_rgba8: out_rgba holds %zu bytes, needs %zu", out_len, rgba.size() * 4);
    std::memcpy(out_rgba, rgba.data(), rgba.size() * 4);
  }
  return APT_OK;
}
