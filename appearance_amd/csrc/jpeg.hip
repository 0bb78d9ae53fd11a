// JPEG compression of the render node's 64x64 blocks on the device.
//
// The reference node compresses every 64x64 RGBX block of a rendered slab on
// the CPU before it sends it (turbojpeg::compress(image, 95, Sub2x2),
// crates/appearance-render-loop/src/node.rs:65-79; ENABLE_COMPRESSION and
// NODE_PIXEL_FORMAT = RGBX, host.rs:21-22), and the host decompresses it
// (host.rs:341-350).  Here the frame is already in HBM in block-major order,
// so one 1024-thread work-group encodes one block, and the bytes are those of
// libjpeg-turbo's default compression (the library the turbojpeg crate
// wraps): integer JFIF colour conversion, 2x2 box downsampling with
// alternating bias, the accurate integer DCT (Loeffler-Ligtenberg-Moschytz),
// rounding division by 8 x the quality-scaled Annex K tables, Huffman coding
// with the Annex K tables.
//
// Work-group phases, all in LDS:
//   1. 4096 pixels -> Y (64x64) and 2x2-averaged Cb, Cr (32x32), level-shifted
//   2. row DCT in place (96 8x8 blocks x 8 rows, one row per thread)
//   3. column DCT and quantisation to zig-zag order (one column per thread)
//   4. Huffman bit counts: a wave per 8x8 block, a lane per coefficient; the
//      run lengths come from a ballot of the nonzero lanes
//   5. scan-order offsets of the 96 blocks, then every lane ORs its codes
//      into the bit buffer at its own offset (kept from step 4)
//   6. byte stuffing (0xFF -> 0xFF 0x00) by 1024 chunks with a prefix sum
// A second pass packs header + segment of every block back to back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "apt_internal.h"
#include "jpeg.h"

namespace apt {

namespace {

// Threads per 64x64 block: 16 waves share the 96 8x8 blocks of the Huffman
// phases (6 each); one frame's 510 blocks took 71 / 51 / 46 us at 256 / 512 /
// 1024 threads (a 2^26-pixel slab 1.10 / 0.89 / 0.99 ms).
constexpr uint32_t JB = 1024;
constexpr uint32_t JP = 256;  // threads of the offset scan and of a block's copy
constexpr uint32_t NB8 = 96;           // 8x8 blocks: 64 Y, 16 Cb, 16 Cr
constexpr uint32_t PIX_WORDS = 6144;   // Y | Cb | Cr samples; then the bit buffer (<= 4980 words)

// Natural index of the k-th coefficient in zig-zag order (T.81 figure A.6).
constexpr uint8_t ZIGZAG[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// T.81 Annex K: K.1 / K.2 quantisation tables (natural order) and the K.3 /
// K.5 Huffman tables (number of codes of each length 1..16, then the symbols
// in code order).
constexpr uint8_t QLUMA[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                               14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                               18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                               49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
constexpr uint8_t QCHROMA[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                                 24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                                 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                                 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
constexpr uint8_t DC_LUMA_BITS[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
constexpr uint8_t DC_CHROMA_BITS[16] = {0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0};
constexpr uint8_t DC_VALS[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
constexpr uint8_t AC_LUMA_BITS[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
constexpr uint8_t AC_LUMA_VALS[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71,
    0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72,
    0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37,
    0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59,
    0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3,
    0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
    0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2,
    0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};
constexpr uint8_t AC_CHROMA_BITS[16] = {0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77};
constexpr uint8_t AC_CHROMA_VALS[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22,
    0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1,
    0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36,
    0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58,
    0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a,
    0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a,
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
    0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda,
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

// What the block kernel reads besides the pixels, copied to LDS at its start.
struct JpegTab {
  uint32_t dc[2][12];     // Huffman code << 8 | length, luma (0) and chroma (1)
  uint32_t ac[2][256];
  uint32_t magic[2][64];  // ceil(2^32 / d) for the divisor d = 8 x quantiser (natural order)
  uint32_t half[2][64];   // d / 2
  uint8_t unzig[64];      // natural index -> zig-zag position
};
constexpr uint32_t TAB_WORDS = sizeof(JpegTab) / 4;
static_assert(sizeof(JpegTab) % 4 == 0, "JpegTab is copied as words");

void huff_codes(const uint8_t* bits, const uint8_t* vals, uint32_t* table) {
  uint32_t code = 0;
  int k = 0;
  for (int len = 1; len <= 16; ++len) {
    for (int i = 0; i < bits[len - 1]; ++i, ++k) table[vals[k]] = (code++ << 8) | (uint32_t)len;
    code <<= 1;
  }
}

// jpeg_quality_scaling with the baseline clamp to 8-bit entries.
int scaled_q(int base, int quality) {
  const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
  return std::min(255, std::max(1, (base * scale + 50) / 100));
}

JpegTab make_tab(int quality) {
  JpegTab t{};
  huff_codes(DC_LUMA_BITS, DC_VALS, t.dc[0]);
  huff_codes(DC_CHROMA_BITS, DC_VALS, t.dc[1]);
  huff_codes(AC_LUMA_BITS, AC_LUMA_VALS, t.ac[0]);
  huff_codes(AC_CHROMA_BITS, AC_CHROMA_VALS, t.ac[1]);
  for (int i = 0; i < 64; ++i) {
    for (int c = 0; c < 2; ++c) {
      const uint64_t d = 8u * (uint32_t)scaled_q((c ? QCHROMA : QLUMA)[i], quality);
      t.magic[c][i] = (uint32_t)(((1ull << 32) + d - 1) / d);
      t.half[c][i] = (uint32_t)(d >> 1);
    }
    t.unzig[ZIGZAG[i]] = (uint8_t)i;
  }
  return t;
}

void put16(std::vector<uint8_t>& h, uint32_t v) {
  h.push_back((uint8_t)(v >> 8));
  h.push_back((uint8_t)v);
}

__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
  const uint32_t lane = __lane_id();
#pragma unroll
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o);
    if (lane >= o) v += t;
  }
  return v;
}

// Exclusive prefix sum over the work-group's T threads; `total` gets the sum.
template <uint32_t T>
__device__ __forceinline__ uint32_t group_scan_excl(uint32_t v, uint32_t* s_wsum, uint32_t& total) {
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
  const uint32_t incl = wave_scan_incl(v);
  if (lane == 63) s_wsum[wave] = incl;
  __syncthreads();
  uint32_t base = 0;
  total = 0;
#pragma unroll
  for (uint32_t w = 0; w < T / 64; ++w) {
    const uint32_t x = s_wsum[w];
    base += w < wave ? x : 0u;
    total += x;
  }
  __syncthreads();
  return base + incl - v;
}

__device__ __forceinline__ uint32_t nbits(int v) {
  const uint32_t a = (uint32_t)(v < 0 ? -v : v);
  return a ? 32u - (uint32_t)__clz((int)a) : 0u;
}

// ORs `len` (1..16) bits of `value`, most significant first, into the bit
// buffer at bit `pos` (bit 0 = the most significant bit of word 0).
__device__ __forceinline__ void put_bits(uint32_t* words, uint32_t& pos, uint32_t value, uint32_t len) {
  if (len == 0) return;
  value &= (1u << len) - 1u;
  const uint32_t w = pos >> 5, off = pos & 31u;
  const uint64_t v = (uint64_t)value << (64u - len - off);
  atomicOr(&words[w], (uint32_t)(v >> 32));
  if (off + len > 32u) atomicOr(&words[w + 1], (uint32_t)v);
  pos += len;
}

// Storage index of 8x8 block j of MCU m.  An MCU (16x16 pixels) is
// Y(0,0) Y(1,0) Y(0,1) Y(1,1) Cb Cr; the 16 MCUs run in raster order.  Y
// blocks are stored by 8x8 grid row * 8 + column, Cb at 64 + MCU index, Cr
// at 80 + it.
__device__ __forceinline__ int store_of(uint32_t m, uint32_t j) {
  const uint32_t my = m >> 2, mx = m & 3u;
  if (j < 4) return (int)((2 * my + (j >> 1)) * 8 + 2 * mx + (j & 1u));
  return (int)((j == 4 ? 64u : 80u) + m);
}

// Scan-order block k: its storage index and that of the previous block of
// its component (-1 for the first), whose DC predicts its DC.
__device__ __forceinline__ void scan_block(uint32_t k, int& store, int& pred) {
  const uint32_t m = k / 6, j = k % 6;
  store = store_of(m, j);
  if (j > 0 && j < 4)
    pred = store_of(m, j - 1);
  else if (m == 0)
    pred = -1;
  else
    pred = store_of(m - 1, j == 0 ? 3u : j);
}

// One lane's share of an 8x8 block's Huffman code: lane 0 the DC difference,
// lane k >= 1 its nonzero coefficient (with the ZRL codes of the zero run
// before it), and the lane of the last nonzero coefficient (lane 0 when all
// AC terms are zero) the end-of-block code.  Returns the bit count; WRITE
// also ORs the bits into `words` from `pos`.
template <bool WRITE>
__device__ __forceinline__ uint32_t code_lane(int c, int pred_dc, uint64_t nz, uint32_t lane, const uint32_t* dc,
                                              const uint32_t* ac, uint32_t* words, uint32_t pos) {
  uint32_t bits = 0;
  const auto emit = [&](uint32_t code_len) {
    const uint32_t len = code_len & 0xFFu;
    bits += len;
    if (WRITE) put_bits(words, pos, code_len >> 8, len);
  };
  const auto emit_value = [&](int v, uint32_t size) {
    bits += size;
    if (WRITE) put_bits(words, pos, (uint32_t)(v < 0 ? v - 1 : v), size);
  };
  if (lane == 0) {
    const int diff = c - pred_dc;
    const uint32_t s = nbits(diff);
    emit(dc[s]);
    emit_value(diff, s);
  } else if (c != 0) {
    const uint64_t below = nz & ((1ull << lane) - 1ull);
    const uint32_t prev = below ? 63u - (uint32_t)__clzll((long long)below) : 0u;
    uint32_t run = lane - prev - 1u;
    for (; run > 15u; run -= 16u) emit(ac[0xF0]);
    const uint32_t sz = nbits(c);
    emit(ac[(run << 4) | sz]);
    emit_value(c, sz);
  }
  const uint32_t last = nz ? 63u - (uint32_t)__clzll((long long)nz) : 0u;
  if (last < 63u && lane == last) emit(ac[0x00]);
  return bits;
}

// Rounding descale (x + 2^(n-1)) >> n with an arithmetic shift.
__device__ __forceinline__ int32_t descale(int32_t x, int n) { return (x + (1 << (n - 1))) >> n; }

// One 8-point pass of the accurate integer DCT.  Pass 1 (rows) keeps 2 extra
// bits; pass 2 (columns) removes them.  Constants: round(x * 2^13).
template <int PASS>
__device__ __forceinline__ void fdct8(const int32_t* d, int32_t* o) {
  constexpr int CB = 13, P1 = 2, SH = PASS == 1 ? CB - P1 : CB + P1;
  const int32_t t0 = d[0] + d[7], t7 = d[0] - d[7];
  const int32_t t1 = d[1] + d[6], t6 = d[1] - d[6];
  const int32_t t2 = d[2] + d[5], t5 = d[2] - d[5];
  const int32_t t3 = d[3] + d[4], t4 = d[3] - d[4];
  const int32_t t10 = t0 + t3, t13 = t0 - t3, t11 = t1 + t2, t12 = t1 - t2;
  if (PASS == 1) {
    o[0] = (t10 + t11) * (1 << P1);
    o[4] = (t10 - t11) * (1 << P1);
  } else {
    o[0] = descale(t10 + t11, P1);
    o[4] = descale(t10 - t11, P1);
  }
  const int32_t z1 = (t12 + t13) * 4433;  // 0.541196100
  o[2] = descale(z1 + t13 * 6270, SH);    // 0.765366865
  o[6] = descale(z1 - t12 * 15137, SH);   // 1.847759065
  const int32_t y1 = t4 + t7, y2 = t5 + t6, y3 = t4 + t6, y4 = t5 + t7;
  const int32_t z5 = (y3 + y4) * 9633;                        // 1.175875602
  const int32_t a4 = t4 * 2446, a5 = t5 * 16819;              // 0.298631336, 2.053119869
  const int32_t a6 = t6 * 25172, a7 = t7 * 12299;             // 3.072711026, 1.501321110
  const int32_t m1 = -y1 * 7373, m2 = -y2 * 20995;            // 0.899976223, 2.562915447
  const int32_t m3 = -y3 * 16069 + z5, m4 = -y4 * 3196 + z5;  // 1.961570560, 0.390180644
  o[7] = descale(a4 + m1 + m3, SH);
  o[5] = descale(a5 + m2 + m4, SH);
  o[3] = descale(a6 + m2 + m3, SH);
  o[1] = descale(a7 + m1 + m4, SH);
}

// First sample of row y of 8x8 block `blk` (storage order) in the sample
// planes, and the plane's row pitch.
__device__ __forceinline__ int32_t* block_row(int32_t* s_pix, uint32_t blk, uint32_t y, uint32_t& pitch) {
  if (blk < 64) {
    pitch = 64;
    return s_pix + ((blk >> 3) * 8 + y) * 64 + (blk & 7u) * 8;
  }
  const uint32_t c = (blk - 64) & 15u;
  pitch = 32;
  return s_pix + (blk < 80 ? 4096 : 5120) + ((c >> 2) * 8 + y) * 32 + (c & 3u) * 8;
}

__global__ __launch_bounds__(JB) void k_jpeg_blocks(const uint32_t* __restrict__ rgbx, const uint32_t* __restrict__ tab,
                                                   uint8_t* __restrict__ seg, uint32_t* __restrict__ seg_len) {
  __shared__ int32_t s_pix[PIX_WORDS];  // Y 64x64 | Cb 32x32 | Cr 32x32; later the bit buffer
  __shared__ int16_t s_coef[NB8 * 64];  // quantised, zig-zag order, storage block order
  __shared__ uint32_t s_tab[TAB_WORDS];
  __shared__ uint16_t s_lpos[NB8 * 64];  // each lane's bit offset in its 8x8 block's code (scan order)
  __shared__ uint32_t s_off[NB8];
  __shared__ uint32_t s_wsum[JB / 64];
  __shared__ uint32_t s_total;
  const uint32_t tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const JpegTab& T = *reinterpret_cast<const JpegTab*>(s_tab);
  for (uint32_t i = tid; i < TAB_WORDS; i += JB) s_tab[i] = tab[i];

  // 1. colour conversion; each thread four 2x2 quads
  const uint32_t* px = rgbx + (size_t)blockIdx.x * 4096;
  for (uint32_t q = tid; q < 1024; q += JB) {
    const uint32_t qx = q & 31u, qy = q >> 5;
    int32_t cb = 0, cr = 0;
#pragma unroll
    for (uint32_t r = 0; r < 2; ++r) {
      const uint2 two = *reinterpret_cast<const uint2*>(px + (2 * qy + r) * 64 + 2 * qx);
#pragma unroll
      for (uint32_t h = 0; h < 2; ++h) {
        const uint32_t c = h ? two.y : two.x;
        const int32_t R = (int32_t)(c & 0xFFu), G = (int32_t)((c >> 8) & 0xFFu), B = (int32_t)((c >> 16) & 0xFFu);
        s_pix[(2 * qy + r) * 64 + 2 * qx + h] = ((19595 * R + 38470 * G + 7471 * B + 32768) >> 16) - 128;
        cb += (-11059 * R - 21709 * G + 32768 * B + (128 << 16) + 32767) >> 16;
        cr += (32768 * R - 27439 * G - 5329 * B + (128 << 16) + 32767) >> 16;
      }
    }
    const int32_t bias = (qx & 1u) ? 2 : 1;
    s_pix[4096 + q] = ((cb + bias) >> 2) - 128;
    s_pix[5120 + q] = ((cr + bias) >> 2) - 128;
  }
  __syncthreads();

  // 2. row pass in place
  for (uint32_t item = tid; item < NB8 * 8; item += JB) {
    uint32_t pitch;
    int32_t* row = block_row(s_pix, item >> 3, item & 7u, pitch);
    int32_t d[8], o[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) d[x] = row[x];
    fdct8<1>(d, o);
#pragma unroll
    for (int x = 0; x < 8; ++x) row[x] = o[x];
  }
  __syncthreads();

  // 3. column pass and quantisation (rounding half away from zero)
  for (uint32_t item = tid; item < NB8 * 8; item += JB) {
    const uint32_t blk = item >> 3, u = item & 7u, c = blk < 64 ? 0u : 1u;
    uint32_t pitch;
    const int32_t* col = block_row(s_pix, blk, 0, pitch) + u;
    int32_t d[8], o[8];
#pragma unroll
    for (int y = 0; y < 8; ++y) d[y] = col[y * pitch];
    fdct8<2>(d, o);
#pragma unroll
    for (uint32_t v = 0; v < 8; ++v) {
      const uint32_t nat = v * 8 + u;
      const int32_t x = o[v];
      const uint32_t n = (uint32_t)(x < 0 ? -x : x) + T.half[c][nat];
      const int32_t qv = (int32_t)__umulhi(n, T.magic[c][nat]);
      s_coef[blk * 64 + T.unzig[nat]] = (int16_t)(x < 0 ? -qv : qv);
    }
  }
  __syncthreads();

  // 4. bit counts of the 8x8 blocks in scan order; clear the bit buffer
  uint32_t* const words = reinterpret_cast<uint32_t*>(s_pix);
  for (uint32_t k = wave; k < NB8; k += JB / 64) {
    int sb, pb;
    scan_block(k, sb, pb);
    const uint32_t comp = sb < 64 ? 0u : 1u;
    const int c = s_coef[sb * 64 + lane];
    const int pred_dc = pb < 0 ? 0 : (int)s_coef[pb * 64];
    const uint64_t nz = __ballot(lane > 0 && c != 0);
    const uint32_t bits = code_lane<false>(c, pred_dc, nz, lane, T.dc[comp], T.ac[comp], nullptr, 0);
    const uint32_t incl = wave_scan_incl(bits);
    s_lpos[k * 64 + lane] = (uint16_t)(incl - bits);  // < 1660 bits per 8x8 block
    if (lane == 63) s_off[k] = incl;
  }
  for (uint32_t i = tid; i < PIX_WORDS; i += JB) words[i] = 0u;
  __syncthreads();

  // 5. block offsets (one wave scans the 96 counts), then the codes
  if (wave == 0) {
    const uint32_t a = s_off[lane], b = lane + 64 < NB8 ? s_off[lane + 64] : 0u;
    const uint32_t ia = wave_scan_incl(a);
    const uint32_t ib = wave_scan_incl(b) + __shfl(ia, 63);
    s_off[lane] = ia - a;
    if (lane + 64 < NB8) s_off[lane + 64] = ib - b;
    if (lane == 63) s_total = ib;
  }
  __syncthreads();
  const uint32_t total_bits = s_total;
  for (uint32_t k = wave; k < NB8; k += JB / 64) {
    int sb, pb;
    scan_block(k, sb, pb);
    const uint32_t comp = sb < 64 ? 0u : 1u;
    const int c = s_coef[sb * 64 + lane];
    const int pred_dc = pb < 0 ? 0 : (int)s_coef[pb * 64];
    const uint64_t nz = __ballot(lane > 0 && c != 0);
    code_lane<true>(c, pred_dc, nz, lane, T.dc[comp], T.ac[comp], words, s_off[k] + s_lpos[k * 64 + lane]);
  }
  if (tid == 0 && (total_bits & 7u)) {  // pad the last byte with 1 bits
    uint32_t pos = total_bits;
    put_bits(words, pos, 0xFFu, 8u - (total_bits & 7u));
  }
  __syncthreads();

  // 6. byte stuffing: thread t takes bytes [t * chunk, (t + 1) * chunk)
  const uint32_t nbytes = (total_bits + 7u) >> 3;
  const uint32_t chunk = (nbytes + JB - 1) / JB;
  const uint32_t b0 = min(nbytes, tid * chunk), b1 = min(nbytes, b0 + chunk);
  const auto byte_at = [&](uint32_t j) { return (words[j >> 2] >> (24u - 8u * (j & 3u))) & 0xFFu; };
  uint32_t ff = 0;
  for (uint32_t j = b0; j < b1; ++j) ff += byte_at(j) == 0xFFu ? 1u : 0u;
  uint32_t ff_total;
  const uint32_t ff_before = group_scan_excl<JB>(ff, s_wsum, ff_total);
  const uint32_t len = nbytes + ff_total + 2u;  // stuffed bytes, then EOI
  uint8_t* dst = seg + (size_t)blockIdx.x * JPEG_SEG_CAP;
  uint32_t o = b0 + ff_before;
  for (uint32_t j = b0; j < b1; ++j) {
    const uint32_t v = byte_at(j);
    dst[o++] = (uint8_t)v;
    if (v == 0xFFu) dst[o++] = 0u;
  }
  if (tid == 0) {
    dst[len - 2] = 0xFF;
    dst[len - 1] = 0xD9;
    seg_len[blockIdx.x] = len;
  }
}

// File offsets: block b's file (header + segment) starts at
// b * hdr_len + the segments before it; off[n] is the total.
__global__ __launch_bounds__(JP) void k_jpeg_offsets(const uint32_t* __restrict__ seg_len, uint32_t n,
                                                    uint32_t hdr_len, uint32_t* __restrict__ off) {
  __shared__ uint32_t s_wsum[JP / 64];
  const uint32_t t = threadIdx.x, chunk = (n + JP - 1) / JP;
  const uint32_t a = min(n, t * chunk), b = min(n, a + chunk);
  uint32_t sum = 0;
  for (uint32_t i = a; i < b; ++i) sum += seg_len[i] + hdr_len;
  uint32_t total;
  uint32_t o = group_scan_excl<JP>(sum, s_wsum, total);
  for (uint32_t i = a; i < b; ++i) {
    off[i] = o;
    o += seg_len[i] + hdr_len;
  }
  if (t == 0) off[n] = total;
}

__global__ __launch_bounds__(JP) void k_jpeg_pack(const uint8_t* __restrict__ seg, const uint32_t* __restrict__ seg_len,
                                                 const uint32_t* __restrict__ off, const uint8_t* __restrict__ hdr,
                                                 uint32_t hdr_len, uint8_t* __restrict__ files) {
  const uint32_t b = blockIdx.x, n = seg_len[b];
  uint8_t* dst = files + off[b];
  const uint8_t* src = seg + (size_t)b * JPEG_SEG_CAP;
  for (uint32_t i = threadIdx.x; i < hdr_len; i += JP) dst[i] = hdr[i];
  for (uint32_t i = threadIdx.x; i < n; i += JP) dst[hdr_len + i] = src[i];
}

}  // namespace

std::vector<uint8_t> jpeg_block_header(int quality) {
  // SOI; APP0: JFIF 1.01, no density unit, 1:1, no thumbnail
  std::vector<uint8_t> h = {0xFF, 0xD8, 0xFF, 0xE0, 0x00, 0x10, 'J',  'F',  'I',  'F',
                            0x00, 0x01, 0x01, 0x00, 0x00, 0x01, 0x00, 0x01, 0x00, 0x00};
  for (int t = 0; t < 2; ++t) {  // DQT per table, 8-bit entries in zig-zag order
    h.insert(h.end(), {0xFF, 0xDB, 0x00, 0x43, (uint8_t)t});
    for (int k = 0; k < 64; ++k) h.push_back((uint8_t)scaled_q((t ? QCHROMA : QLUMA)[ZIGZAG[k]], quality));
  }
  // SOF0: 8-bit, 64 x 64, Y 2x2 with table 0, Cb and Cr 1x1 with table 1
  h.insert(h.end(), {0xFF, 0xC0, 0x00, 0x11, 0x08, 0x00, 0x40, 0x00, 0x40, 0x03, 0x01, 0x22, 0x00, 0x02, 0x11, 0x01,
                     0x03, 0x11, 0x01});
  const auto dht = [&](uint8_t cls_id, const uint8_t* bits, const uint8_t* vals, int nvals) {
    h.insert(h.end(), {0xFF, 0xC4});
    put16(h, (uint32_t)(3 + 16 + nvals));
    h.push_back(cls_id);
    h.insert(h.end(), bits, bits + 16);
    h.insert(h.end(), vals, vals + nvals);
  };
  dht(0x00, DC_LUMA_BITS, DC_VALS, 12);
  dht(0x10, AC_LUMA_BITS, AC_LUMA_VALS, 162);
  dht(0x01, DC_CHROMA_BITS, DC_VALS, 12);
  dht(0x11, AC_CHROMA_BITS, AC_CHROMA_VALS, 162);
  // SOS: Y with tables 0/0, Cb and Cr with 1/1, coefficients 0..63
  h.insert(h.end(), {0xFF, 0xDA, 0x00, 0x0C, 0x03, 0x01, 0x00, 0x02, 0x11, 0x03, 0x11, 0x00, 0x3F, 0x00});
  return h;
}

apt_status JpegEncoder::encode(const uint32_t* d_rgbx, uint32_t n, int q, hipStream_t s) {
  APT_REQUIRE(q >= 1 && q <= 100, APT_ERR_INVALID_ARG, "jpeg: quality %d outside 1..100", q);
  APT_REQUIRE(n > 0 && n <= JPEG_MAX_BLOCKS, APT_ERR_INVALID_ARG, "jpeg: %u blocks outside 1..%u", n,
              JPEG_MAX_BLOCKS);
  if (!ev0) {
    APT_HIP(hipEventCreate(&ev0));
    APT_HIP(hipEventCreate(&ev1));
  }
  if (!d_tab) {
    APT_HIP(hipMalloc(&d_tab, sizeof(JpegTab)));
    APT_HIP(hipMalloc(&d_hdr, JPEG_HEADER_CAP));
  }
  if (q != quality) {
    header = jpeg_block_header(q);
    const JpegTab t = make_tab(q);
    APT_HIP(hipMemcpy(d_tab, &t, sizeof(t), hipMemcpyHostToDevice));
    APT_HIP(hipMemcpy(d_hdr, header.data(), header.size(), hipMemcpyHostToDevice));
    quality = q;
  }
  const uint32_t hdr_len = (uint32_t)header.size();
  if (n > block_cap) {
    for (void* p : {(void*)d_seg, (void*)d_files, (void*)d_len, (void*)d_off})
      if (p) (void)hipFree(p);
    d_seg = d_files = nullptr;
    d_len = d_off = nullptr;
    block_cap = 0;
    APT_HIP(hipMalloc(&d_seg, (size_t)n * JPEG_SEG_CAP));
    APT_HIP(hipMalloc(&d_files, (size_t)n * (JPEG_SEG_CAP + JPEG_HEADER_CAP)));
    APT_HIP(hipMalloc(&d_len, (size_t)n * sizeof(uint32_t)));
    APT_HIP(hipMalloc(&d_off, ((size_t)n + 1) * sizeof(uint32_t)));
    block_cap = n;
  }
  APT_HIP(hipEventRecord(ev0, s));
  hipLaunchKernelGGL(k_jpeg_blocks, dim3(n), dim3(JB), 0, s, d_rgbx, static_cast<const uint32_t*>(d_tab), d_seg,
                     d_len);
  hipLaunchKernelGGL(k_jpeg_offsets, dim3(1), dim3(JP), 0, s, d_len, n, hdr_len, d_off);
  hipLaunchKernelGGL(k_jpeg_pack, dim3(n), dim3(JP), 0, s, d_seg, d_len, d_off, d_hdr, hdr_len, d_files);
  APT_HIP(hipGetLastError());
  APT_HIP(hipEventRecord(ev1, s));
  offsets.resize((size_t)n + 1);
  APT_HIP(hipMemcpyAsync(offsets.data(), d_off, offsets.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  APT_HIP(hipStreamSynchronize(s));
  const size_t total = offsets[n];
  if (total > h_cap) {
    if (h_files) (void)hipHostFree(h_files);
    h_files = nullptr;
    h_cap = 0;
    APT_HIP(hipHostMalloc((void**)&h_files, total, hipHostMallocDefault));
    h_cap = total;
  }
  APT_HIP(hipMemcpyAsync(h_files, d_files, total, hipMemcpyDeviceToHost, s));
  APT_HIP(hipStreamSynchronize(s));
  APT_HIP(hipEventElapsedTime(&ms_device, ev0, ev1));
  return APT_OK;
}

void JpegEncoder::release() {
  for (void* p : {d_tab, (void*)d_hdr, (void*)d_seg, (void*)d_files, (void*)d_len, (void*)d_off})
    if (p) (void)hipFree(p);
  d_tab = nullptr;
  d_hdr = d_seg = d_files = nullptr;
  d_len = d_off = nullptr;
  block_cap = 0;
  if (h_files) (void)hipHostFree(h_files);
  h_files = nullptr;
  h_cap = 0;
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  ev0 = ev1 = nullptr;
  quality = -1;
}

}  // namespace apt

static_assert(apt::JPEG_SEG_CAP + 623 <= APT_JPEG_FILE_MAX, "a block's file fits APT_JPEG_FILE_MAX");
static_assert(apt::JPEG_MAX_BLOCKS == APT_JPEG_MAX_BLOCKS, "block limit");

extern "C" apt_status apt_jpeg_compress_blocks(const uint8_t* rgbx, uint32_t n_blocks, int quality, uint8_t* out,
                                               size_t out_cap, uint64_t* offsets, float* ms_device) {
  using namespace apt;
  clear_error();
  APT_REQUIRE(rgbx && out && offsets, APT_ERR_INVALID_ARG, "apt_jpeg_compress_blocks: null argument");
  APT_REQUIRE(n_blocks > 0 && n_blocks <= JPEG_MAX_BLOCKS, APT_ERR_INVALID_ARG,
              "apt_jpeg_compress_blocks: %u blocks outside 1..%u", n_blocks, JPEG_MAX_BLOCKS);
  APT_REQUIRE(quality >= 1 && quality <= 100, APT_ERR_INVALID_ARG,
              "apt_jpeg_compress_blocks: quality %d outside 1..100", quality);
  hipStream_t s = lib_stream();
  const size_t in_bytes = (size_t)n_blocks * 64 * 64 * 4;
  uint32_t* d_in = nullptr;
  APT_HIP(hipMalloc(&d_in, in_bytes));
  JpegEncoder enc;
  hipError_t e = hipMemcpyAsync(d_in, rgbx, in_bytes, hipMemcpyHostToDevice, s);
  apt_status st = APT_OK;
  if (e != hipSuccess) {
    set_error("apt_jpeg_compress_blocks: copy in failed: %s", hipGetErrorString(e));
    st = APT_ERR_DEVICE;
  }
  if (st == APT_OK) st = enc.encode(d_in, n_blocks, quality, s);
  (void)hipFree(d_in);
  if (st != APT_OK) return st;
  const size_t total = enc.offsets[n_blocks];
  for (uint32_t b = 0; b <= n_blocks; ++b) offsets[b] = enc.offsets[b];
  APT_REQUIRE(out_cap >= total, APT_ERR_INVALID_ARG, "apt_jpeg_compress_blocks: out holds %zu bytes, needs %zu",
              out_cap, total);
  std::memcpy(out, enc.h_files, total);
  if (ms_device) *ms_device = enc.ms_device;
  return APT_OK;
}
