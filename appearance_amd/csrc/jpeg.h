// Baseline JPEG compression of 64x64 RGBX blocks on the device (jpeg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "apt.h"

namespace apt {

// Bytes a block's entropy-coded segment can take: 96 8x8 blocks of at most
// 1660 bits (22 for the DC difference, 63 x 26 for the AC terms), doubled for
// 0xFF stuffing, plus the EOI marker.
constexpr uint32_t JPEG_SEG_CAP = 40960;
constexpr uint32_t JPEG_HEADER_CAP = 1024;  // the header is 623 bytes
// Blocks per call: a 2^26-pixel slab.
constexpr uint32_t JPEG_MAX_BLOCKS = 16384;

// Encoder state: device buffers sized for the largest batch seen, and the
// last batch's files on the host (pinned), back to back.
struct JpegEncoder {
  int quality = -1;
  std::vector<uint8_t> header;  // SOI .. SOS for `quality`
  void* d_tab = nullptr;        // Huffman codes and quantiser reciprocals
  uint8_t* d_hdr = nullptr;
  uint8_t* d_seg = nullptr;     // per block: JPEG_SEG_CAP bytes
  uint8_t* d_files = nullptr;   // the files, back to back
  uint32_t* d_len = nullptr;    // per block: segment bytes
  uint32_t* d_off = nullptr;    // per block: first byte of its file; [n]: total
  uint32_t block_cap = 0;
  uint8_t* h_files = nullptr;   // pinned
  size_t h_cap = 0;
  std::vector<uint32_t> offsets;  // n + 1 entries after encode()
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float ms_device = 0.0f;       // the last encode's kernels

  // Compresses n_blocks blocks of 64 x 64 RGBX pixels (4096 words each,
  // rows top to bottom) from device memory into complete JFIF files, copied
  // to h_files with offsets; waits for the stream.
  apt_status encode(const uint32_t* d_rgbx, uint32_t n_blocks, int quality, hipStream_t s);
  void release();
  ~JpegEncoder() { release(); }
};

// The SOI .. SOS marker bytes every block's file starts with.
std::vector<uint8_t> jpeg_block_header(int quality);

}  // namespace apt
