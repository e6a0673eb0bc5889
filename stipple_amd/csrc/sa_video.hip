// sa_video.hip -- the palette montage of sampled video frames, on the device.
//
// VideoBuildPalette's montage path (src/rd_video.cpp:2301-2470, :2636-2665)
// lays the sampled frames out as cells of a grid, `cols` cells wide, in sample
// order, and hands the montage to the octree build:
//   * a cell is the whole frame (tile <= 0), or a tile x tile nearest-neighbour
//     sample of it, source (tx*W/tile, ty*H/tile) in 64-bit integer arithmetic
//     (place_raw, :2441-2470);
//   * cells past the last sample repeat the last sample's cell (:2636-2660);
//   * 8-bit samples widen x257 and a 3-channel source gets alpha 65535 (widen,
//     :2404-2418); 16-bit samples are used as they are.
// Every widened value is an integer in [0, 65535], so the montage is written
// as RGBA16 and the octree build reads exactly the floats the reference's
// RgbaF montage holds.  One thread per montage pixel; consecutive threads
// write consecutive pixels (8-byte stores), and within a cell row they read
// consecutive source pixels.
#include "sa_internal.h"

namespace sa {
namespace {

__global__ __launch_bounds__(256) void montage_kernel(const void* const* __restrict__ frames,
                                                      int sampled, int src_fmt, int64_t w,
                                                      int64_t h, int64_t cell_w, int64_t cell_h,
                                                      int cols, int64_t mw, int64_t mh,
                                                      ushort4* __restrict__ out,
                                                      int* __restrict__ gray) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= mw * mh) return;
  const int64_t my = i / mw, mx = i - my * mw;
  const int64_t cy = my / cell_h, cx = mx / cell_w;
  int64_t s = cy * cols + cx;
  if (s >= sampled) s = sampled - 1;  // the repeated last sample
  const int64_t ty = my - cy * cell_h, tx = mx - cx * cell_w;
  int64_t sx = tx, sy = ty;
  if (cell_w != w || cell_h != h) {  // a tile: nearest-neighbour sample
    sx = tx * w / cell_w;
    sy = ty * h / cell_h;
  }
  const int64_t p = sy * w + sx;
  const void* f = frames[s];
  ushort4 o;
  switch (src_fmt) {
    case kRgb8: {
      const uint8_t* q = (const uint8_t*)f + 3 * p;
      o = make_ushort4((uint16_t)(q[0] * 257), (uint16_t)(q[1] * 257), (uint16_t)(q[2] * 257),
                       65535);
      break;
    }
    case kRgba8: {
      const uint32_t v = ((const uint32_t*)f)[p];
      o = make_ushort4((uint16_t)((v & 0xFF) * 257), (uint16_t)(((v >> 8) & 0xFF) * 257),
                       (uint16_t)(((v >> 16) & 0xFF) * 257), (uint16_t)((v >> 24) * 257));
      break;
    }
    case kRgb16: {
      const uint16_t* q = (const uint16_t*)f + 3 * p;
      o = make_ushort4(q[0], q[1], q[2], 65535);
      break;
    }
    default:  // kRgba16
      o = ((const ushort4*)f)[p];
      break;
  }
  out[i] = o;
  if (o.x != o.y || o.y != o.z) *gray = 0;  // every writer stores the same value
}

__global__ void flag_set_kernel(int* flag) { *flag = 1; }

}  // namespace

hipError_t launch_video_montage(const void* const* frames, int sampled, int src_fmt, int64_t w,
                                int64_t h, int64_t cell_w, int64_t cell_h, int cols, int64_t mw,
                                int64_t mh, void* out, int* gray, hipStream_t s) {
  flag_set_kernel<<<1, 1, 0, s>>>(gray);
  const int64_t n = mw * mh;
  if (n <= 0) return hipGetLastError();
  montage_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(
      frames, sampled, src_fmt, w, h, cell_w, cell_h, cols, mw, mh, (ushort4*)out, gray);
  return hipGetLastError();
}

}  // namespace sa
