// sa_internal.h -- types shared by the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "sa_math.h"
#include "sa_palette.h"

namespace sa {

// Pixel formats (values match SA_PIX_* in include/stipple_amd.h).
enum PixFmt : int32_t {
  kRgbaF32 = 0,
  kRgb8 = 1,
  kRgba8 = 2,
  kRgb16 = 3,
  kRgba16 = 4,
  kIndex8 = 5,
  kIndex16 = 6,
  kYuv444 = 7,  // planar 8-bit Y, Cb, Cr planes of w*h bytes (block engine only)
  kYuv420 = 8,  // planar 8-bit Y (w*h), Cb and Cr ((w/2)*(h/2)); block engine input only
};

inline int pix_bytes(int fmt) {
  switch (fmt) {
    case kRgbaF32: return 16;
    case kRgb8: return 3;
    case kRgba8: return 4;
    case kRgb16: return 6;
    case kRgba16: return 8;
    case kIndex8: return 1;
    case kIndex16: return 2;
  }
  return 0;
}

// Bytes of one w x h image in a planar format.
inline int64_t planar_bytes(int fmt, int64_t w, int64_t h) {
  return fmt == kYuv444 ? 3 * w * h : fmt == kYuv420 ? w * h + 2 * (w / 2) * (h / 2) : 0;
}

// swscale's 8-bit YUV to 16-bit RGB in integers (SwsYuvToRgb16,
// include/rd_types.h:75-97), returned as r | g << 16 and b | 0xFFFF << 16:
// the 16-bit staging word with the opaque alpha the video reader synthesises.
__device__ __forceinline__ int sws_clip16(int x) {
  const int v = x + (1 << 15);
  return v < 0 ? 0 : (v > 65535 ? 65535 : v);
}
__device__ __forceinline__ uint2 sws_yuv_rgb16(unsigned y, unsigned u, unsigned v) {
  // Luma: (y - 128) * 512 + 2^16, less the 8192 black level, times the
  // 9539/2^13 range gain, plus the rounding term 2^13 and the -2^29 that
  // re-centres the later +2^15.
  int yy = (int)((y - 128u) * 512u);
  const int uu = (int)((u - 128u) * 512u);
  const int vv = (int)((v - 128u) * 512u);
  yy = (yy + 0x10000 - 8192) * 9539 + (1 << 13) - (1 << 29);
  const int r = sws_clip16((vv * 13075 + yy) >> 14);
  const int g = sws_clip16((vv * -6660 + uu * -3209 + yy) >> 14);
  const int b = sws_clip16((uu * 16525 + yy) >> 14);
  return make_uint2((uint32_t)r | ((uint32_t)g << 16), (uint32_t)b | 0xFFFF0000u);
}

// 4:2:0 chroma at luma (x, y): the 2x2 neighbourhood of (x/2, y/2), edges
// repeated, weighted (2 - ox, ox) x (2 - oy, oy) with ox = x & 1, oy = y & 1,
// rounded by +2 >> 2 (src/rd_blocks_cuda.cu:651-675).  Even coordinates copy.
__device__ __forceinline__ unsigned chroma420(const uint8_t* plane, int64_t w2, int64_t h2,
                                              int64_t x, int64_t y) {
  const int64_t cx = x >> 1, cy = y >> 1;
  const int ox = (int)(x & 1), oy = (int)(y & 1);
  const int64_t cx1 = cx + 1 < w2 ? cx + 1 : cx;
  const int64_t cy1 = cy + 1 < h2 ? cy + 1 : cy;
  const int a = plane[cy * w2 + cx], b = plane[cy * w2 + cx1];
  const int c = plane[cy1 * w2 + cx], d = plane[cy1 * w2 + cx1];
  const int top = a * (2 - ox) + b * ox;
  const int bot = c * (2 - ox) + d * ox;
  return (unsigned)((top * (2 - oy) + bot * oy + 2) >> 2);
}

// Pixel p of a planar YUV image as the 16-bit staging word (sws_yuv_rgb16).
__device__ __forceinline__ uint2 load_yuv(int fmt, const void* base, int64_t p, int64_t w,
                                          int64_t h) {
  const uint8_t* s = (const uint8_t*)base;
  const int64_t pp = w * h;
  if (fmt == kYuv444) return sws_yuv_rgb16(s[p], s[pp + p], s[2 * pp + p]);
  const int64_t w2 = w >> 1, h2 = h >> 1;
  const int64_t y = p / w, x = p - y * w;
  return sws_yuv_rgb16(s[p], chroma420(s + pp, w2, h2, x, y),
                       chroma420(s + pp + w2 * h2, w2, h2, x, y));
}

// Limited-range BT.601 8-bit Y, Cb, Cr of Q16 floats r, g, b (their integer
// parts' high bytes), clamped to 0..255 (src/rd_riemersma_cpu.cpp:718-730).
__device__ __forceinline__ void store_yuv444(void* dst, int64_t pix, int64_t pp, float r, float g,
                                             float b) {
  const int r8 = (int)r >> 8, g8 = (int)g >> 8, b8 = (int)b >> 8;
  const int yv = ((66 * r8 + 129 * g8 + 25 * b8 + 128) >> 8) + 16;
  const int uv = ((-38 * r8 - 74 * g8 + 112 * b8 + 128) >> 8) + 128;
  const int vv = ((112 * r8 - 94 * g8 - 18 * b8 + 128) >> 8) + 128;
  uint8_t* d = (uint8_t*)dst;
  d[pix] = (uint8_t)(yv < 0 ? 0 : (yv > 255 ? 255 : yv));
  d[pp + pix] = (uint8_t)(uv < 0 ? 0 : (uv > 255 ? 255 : uv));
  d[2 * pp + pix] = (uint8_t)(vv < 0 ? 0 : (vv > 255 ? 255 : vv));
}

// Loads one source pixel as Q16 floats: 8-bit samples widen x257 (exact),
// 16-bit samples as-is, float as-is; alpha 65535 when the format has none.
__device__ __forceinline__ float4 load_pixel(int fmt, const void* base, int64_t p) {
  float4 o;
  switch (fmt) {
    case kRgb8: {
      const uint8_t* s = (const uint8_t*)base + 3 * p;
      o = make_float4((float)s[0] * 257.0f, (float)s[1] * 257.0f, (float)s[2] * 257.0f, 65535.0f);
      break;
    }
    case kRgba8: {
      const uint32_t v = ((const uint32_t*)base)[p];
      o = make_float4((float)(v & 255u) * 257.0f, (float)((v >> 8) & 255u) * 257.0f,
                      (float)((v >> 16) & 255u) * 257.0f, (float)(v >> 24) * 257.0f);
      break;
    }
    case kRgb16: {
      const uint16_t* s = (const uint16_t*)base + 3 * p;
      o = make_float4((float)s[0], (float)s[1], (float)s[2], 65535.0f);
      break;
    }
    case kRgba16: {
      const uint2 v = ((const uint2*)base)[p];
      o = make_float4((float)(v.x & 0xFFFFu), (float)(v.x >> 16), (float)(v.y & 0xFFFFu),
                      (float)(v.y >> 16));
      break;
    }
    default:
      o = ((const float4*)base)[p];
      break;
  }
  return o;
}

// Any source pixel as Q16 floats, planar YUV (w x h) included; the block
// engine and the palette build read every format through this.
__device__ __forceinline__ float4 load_any(int fmt, const void* base, int64_t p, int64_t w,
                                           int64_t h) {
  if (fmt == kYuv444 || fmt == kYuv420) {
    const uint2 v = load_yuv(fmt, base, p, w, h);
    return make_float4((float)(v.x & 0xFFFFu), (float)(v.x >> 16), (float)(v.y & 0xFFFFu),
                       65535.0f);
  }
  return load_pixel(fmt, base, p);
}

// Where a resumable walk stopped: the next visit, and the 64 lanes' in-flight
// running sums and residuals (the rotating queue of sa_walk.hip).
struct SeqState {
  int64_t t;
  int64_t pad_;
  double acc[64];
  double e[64];
};

// Device scratch for the duration of a call's work on stream `s`, reused
// across calls (sa_scratch.cpp); scratch_free hands it back once the work
// queued on `s` so far is done with it.  Same contract as hipMallocAsync /
// hipFreeAsync, which a stream under graph capture gets.
hipError_t scratch_alloc(void** out, size_t bytes, hipStream_t s);
hipError_t scratch_free(void* p, hipStream_t s);

// The sequential walk stages each visit's source pixel in visit order, in the
// source's own width: 8-bit RGB(A) as 4 bytes (alpha 255 for RGB), 16-bit as
// 8 bytes (alpha 65535 for RGB), float RGBA as 16.  log2 of that width:
__host__ __device__ inline int seq_stage_shift(int fmt) {
  return (fmt == kRgb8 || fmt == kRgba8) ? 2 : (fmt == kRgb16 || fmt == kRgba16) ? 3 : 4;
}
// LDS the walk keeps beside the memo cache and the search image: the next
// 32 visits' associated inputs (4 doubles each).
constexpr int kSeqRingBytes = 32 * 4 * 8;

// One image of the sequential engine (device array of these).
struct SeqItem {
  const void* src;        // source pixels (raster)
  void* dst;              // output (raster), may alias src for kRgbaF32
  const int32_t* curve;   // visit order: w*h pixel ids
  void* stage;            // scratch: the source pixel of each visit, in visit order
                          // (1 << seq_stage_shift(src_fmt) bytes each, +16 B of padding)
  uint16_t* idx;          // scratch: palette index per visit
  void* memo;             // scratch: 1<<K memo entries of memo_eb bytes (K = 18, or 24
                          // with alpha)
  const PaletteDev* pal;  // device palette header
  SeqState* state;        // null: walk to the end in one launch
  unsigned char* snap;    // with state: the LDS memo cache between launches
                          // (seq_memo_bytes(memo_slots, memo_eb) bytes)
  int64_t n;              // visits == w*h
  int32_t src_fmt;
  int32_t dst_fmt;
  int32_t memo_slots;     // LDS memo line slots (a multiple of 4)
  int32_t memo_eb;        // bytes per memo entry (index + 1): 1 for palettes of at most
                          // 255 colours, 2 up to 65535, 4 for 65536
};

// Palette search images up to this size are staged in LDS for the sequential
// walk; larger ones (palettes of more than about 1400 colours) stay in HBM and
// the walk reads them there, leaving the LDS to its memo cache.
constexpr int kSeqImageLdsMax = 96 * 1024;
__host__ __device__ inline int seq_img_lds_bytes(int bytes) {
  return bytes <= kSeqImageLdsMax ? bytes : 0;
}

// LDS bytes of the walk's memo cache: per slot a 4-byte tag and 64 entries.
__host__ __device__ inline uint32_t seq_memo_bytes(int slots, int eb) {
  return (uint32_t)slots * (4u + 64u * (uint32_t)eb);
}

// One frame of the block engine.
struct BlockFrame {
  const void* src;
  void* dst;
  const PaletteDev* pal;
  int32_t src_fmt;
  int32_t dst_fmt;
};

// One frame of the screen dithers: the palette's header and its entries as
// doubles (clustered-dot's luma is taken from those).
struct ScreenFrame {
  const void* src;
  void* dst;
  const PaletteDev* pal;
  const double* entries;
  int32_t src_fmt;
  int32_t dst_fmt;
};

// Writes a plugin dither's choice at pixel `pix`: the entry's r, g, b (WriteChosen,
// src/rd_dither_common.h) with the source's alpha `a`, in the frame's output
// format (the index itself for the index formats).
__device__ __forceinline__ void store_pick(const ScreenFrame& f, int64_t pix, int64_t pp, int idx,
                                           const float4& e, float a) {
  switch (f.dst_fmt) {
    case kIndex8: ((uint8_t*)f.dst)[pix] = (uint8_t)idx; break;
    case kIndex16: ((uint16_t*)f.dst)[pix] = (uint16_t)idx; break;
    case kRgb8: {
      uint8_t* d = (uint8_t*)f.dst + 3 * pix;
      d[0] = (uint8_t)scale_char(e.x);
      d[1] = (uint8_t)scale_char(e.y);
      d[2] = (uint8_t)scale_char(e.z);
      break;
    }
    case kRgba8:
      ((uint32_t*)f.dst)[pix] = scale_char(e.x) | (scale_char(e.y) << 8) |
                                (scale_char(e.z) << 16) | (scale_char(a) << 24);
      break;
    case kRgb16: {
      uint16_t* d = (uint16_t*)f.dst + 3 * pix;
      d[0] = (uint16_t)e.x;
      d[1] = (uint16_t)e.y;
      d[2] = (uint16_t)e.z;
      break;
    }
    case kRgba16: {
      ushort4 o;
      o.x = (uint16_t)e.x;
      o.y = (uint16_t)e.y;
      o.z = (uint16_t)e.z;
      o.w = (uint16_t)a;
      ((ushort4*)f.dst)[pix] = o;
      break;
    }
    case kYuv444: store_yuv444(f.dst, pix, pp, e.x, e.y, e.z); break;
    default: ((float4*)f.dst)[pix] = make_float4(e.x, e.y, e.z, a); break;
  }
}


struct Launch {
  hipStream_t stream;
  int cus;  // compute units on the device
};

// ---- launchers (return hipError_t) ----------------------------------------
hipError_t launch_curve_count(int level, int64_t w, int64_t h, int32_t* chunk_counts,
                              int64_t chunks, hipStream_t s);
hipError_t launch_curve_scan(int32_t* chunk_counts, int64_t chunks, hipStream_t s);
hipError_t launch_curve_emit(int level, int64_t w, int64_t h, const int32_t* chunk_offsets,
                             int64_t chunks, int32_t* order, hipStream_t s);
int64_t curve_chunks(int level);

// sa_video.hip: the palette montage of `sampled` frames (device pointer array)
// as RGBA16, and whether every montage pixel is grey (*gray, device int).
hipError_t launch_video_montage(const void* const* frames, int sampled, int src_fmt, int64_t w,
                                int64_t h, int64_t cell_w, int64_t cell_h, int cols, int64_t mw,
                                int64_t mh, void* out, int* gray, hipStream_t s);

hipError_t launch_seq_gather(const SeqItem* items, int nitems, int64_t max_n, hipStream_t s);
// Sets every item's state->t to 0.
hipError_t launch_seq_state_init(const SeqItem* items, int nitems, hipStream_t s);
// Zeroes the items' HBM memo tables (2 << key_bits bytes each).
hipError_t launch_memo_clear(const SeqItem* items, int nitems, int key_bits, hipStream_t s);
// state->t of every item (n for items without a state).
hipError_t launch_seq_progress(const SeqItem* items, int nitems, int64_t* out, hipStream_t s);
// `lds_bytes`: the largest memo cache + search image + input ring of the items.
// Items with a state advance by at most `slice` visits when slice > 0 (a
// multiple of 48), and stop at the first 48-visit boundary `budget`
// s_memrealtime ticks (100 MHz) after they start when budget > 0; otherwise,
// and items without a state, walk to the end.
hipError_t launch_seq_walk(const SeqItem* items, int nitems, int32_t* queue, const double* sw,
                           int use_cache, size_t lds_bytes, int waves, int assoc_any,
                           int64_t slice, uint64_t budget, uint32_t* stats, hipStream_t s);
hipError_t launch_seq_pack(const SeqItem* items, int nitems, int64_t max_n, hipStream_t s);

// LDS of the block walk's workgroups at most (gfx950: 160 KiB per CU).
constexpr size_t kBlocksMaxLds = 160 * 1024;
// Source class of a pixel format for the block walk (the batch stages every
// frame as the widest class among them: 0 8-bit, 1 16-bit, 2 float).
int blocks_src_class(int fmt);
size_t blocks_lds_bytes(int img_bytes, int count, int src_class, int threads);
// img_bytes / count: the largest search image and palette of the frames.
hipError_t launch_blocks(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t w,
                         int64_t h, int64_t nb, int block, const double* sw, int assoc,
                         int img_bytes, int count, int src_class, int threads, hipStream_t s);
hipError_t launch_blocks_fill(const BlockFrame* frames, int nframes, int64_t gap_pixel,
                              int64_t w, int64_t h, hipStream_t s);

// The screen dithers (screen 0 Bayer, 1 void-and-cluster 64, 2 void-and-cluster
// 32, 3 clustered-dot); `table` is the screen's side x side thresholds.
// clustered-dot's per-frame candidates: darkest luma, its span, the entries.
struct ScreenRange {
  double lo, span;
  int32_t lo_i, hi_i;
};
// `ranges`: nframes entries of scratch, written and read by clustered-dot.
hipError_t launch_screen(const ScreenFrame* frames, int nframes, int64_t w, int64_t h,
                         const uint16_t* table, int side, int screen, double diffusion,
                         ScreenRange* ranges, hipStream_t s);
std::vector<uint16_t> void_cluster_ranks(int n);
// sa_diffuse.hip: the error-diffusion dithers (kind 0 floyd-steinberg, 1
// atkinson, 2 jarvis), one frame per wavefront; LDS bytes per workgroup.
size_t diffuse_lds_bytes(int kind, int64_t w, int count);
hipError_t launch_diffuse(const ScreenFrame* frames, int nframes, int64_t w, int64_t h, int kind,
                          int max_count, float diffusion, hipStream_t s);
extern const uint16_t kBayer8[64];
extern const uint16_t kDot8[64];

size_t approx_lds_bytes(int img_bytes, int taps, int fp64, int assoc, int per);
hipError_t launch_approx(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t nb,
                         int iterations, int taps, const void* g, int fp64, int use_tree,
                         int assoc, int img_bytes, void* cx, uint8_t* q0, uint8_t* q1,
                         uint8_t** result, int per, hipStream_t s);

}  // namespace sa
