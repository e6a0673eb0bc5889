// sa_curve.hip -- Hilbert visit order generated on the device.
//
// The sequential walk visits HilbertPoint(n) for n in [0, 4^L), in-bounds
// only (the recursion of src/rd_riemersma_cpu.cpp:126-153 covers 4^L - 1
// cells and the trailing ForgetGravity visit lands on the missing one,
// (2^L-1, 0), which is also HilbertPoint(4^L - 1)).  The block curve of
// src/rd_riemersma_cpu.cpp:269-356 is the same list without that last cell
// plus a trailing origin, so one compaction serves both engines.
//
// Compaction: the index range is cut into chunks of 4096 consecutive n, which
// at level >= 6 is exactly one aligned 64x64 square of the grid, so a chunk
// whose square lies outside the image is rejected with one d2xy call.  Three
// passes: count per chunk, exclusive scan of the counts, ordered emit.
#include "sa_internal.h"

namespace sa {
namespace {

constexpr int kChunk = 4096;
constexpr int kThreads = 256;
constexpr int kPerThread = kChunk / kThreads;  // 16 consecutive n per thread

// Textbook d2xy in the reference's orientation (src/rd_riemersma_cpu.cpp:167-189).
__device__ __forceinline__ void d2xy(uint32_t n, int level, uint32_t* ox, uint32_t* oy) {
  uint32_t x = 0, y = 0;
  for (int b = 0; b < level; ++b) {
    const uint32_t s = 1u << b;
    const uint32_t rx = 1u & (n >> 1);
    const uint32_t ry = 1u & (n ^ rx);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      const uint32_t t = x;
      x = y;
      y = t;
    }
    x += s * rx;
    y += s * ry;
    n >>= 2;
  }
  *ox = x;
  *oy = y;
}

__device__ __forceinline__ bool chunk_outside(int64_t chunk, int level, int64_t w, int64_t h) {
  if (level < 6) return false;
  uint32_t x, y;
  d2xy((uint32_t)(chunk * kChunk), level, &x, &y);
  return (int64_t)(x & ~63u) >= w || (int64_t)(y & ~63u) >= h;
}

__global__ __launch_bounds__(kThreads) void curve_count_kernel(int level, int64_t w, int64_t h,
                                                               int64_t total,
                                                               int32_t* counts) {
  const int64_t chunk = blockIdx.x;
  __shared__ int32_t part[kThreads / 64];
  if (chunk_outside(chunk, level, w, h)) {
    if (threadIdx.x == 0) counts[chunk] = 0;
    return;
  }
  int c = 0;
  const int64_t first = chunk * kChunk + (int64_t)threadIdx.x * kPerThread;
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t n = first + k;
    if (n >= total) break;
    uint32_t x, y;
    d2xy((uint32_t)n, level, &x, &y);
    c += ((int64_t)x < w && (int64_t)y < h) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) counts[chunk] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of up to any number of chunk counts, one workgroup.
__global__ __launch_bounds__(1024) void curve_scan_kernel(int32_t* counts, int64_t chunks) {
  __shared__ int32_t buf[1024];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < chunks; base += 1024) {
    const int64_t i = base + threadIdx.x;
    const int32_t v = i < chunks ? counts[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int32_t t = threadIdx.x >= (unsigned)off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < chunks) counts[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void curve_emit_kernel(int level, int64_t w, int64_t h,
                                                              int64_t total,
                                                              const int32_t* offsets,
                                                              int32_t* order) {
  const int64_t chunk = blockIdx.x;
  if (chunk_outside(chunk, level, w, h)) return;
  __shared__ int32_t wave_tot[kThreads / 64];
  uint32_t px[kPerThread];
  int c = 0;
  const int64_t first = chunk * kChunk + (int64_t)threadIdx.x * kPerThread;
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    const int64_t n = first + k;
    px[k] = 0xFFFFFFFFu;
    if (n < total) {
      uint32_t x, y;
      d2xy((uint32_t)n, level, &x, &y);
      if ((int64_t)x < w && (int64_t)y < h) {
        px[k] = (uint32_t)((int64_t)y * w + x);
        ++c;
      }
    }
  }
  // Exclusive prefix of c across the workgroup, in thread order.
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  int incl = c;
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wave_tot[wv] = incl;
  __syncthreads();
  int before = 0;
  for (int k = 0; k < wv; ++k) before += wave_tot[k];
  int64_t pos = (int64_t)offsets[chunk] + before + (incl - c);
#pragma unroll
  for (int k = 0; k < kPerThread; ++k) {
    if (px[k] != 0xFFFFFFFFu) order[pos++] = (int32_t)px[k];
  }
}

}  // namespace

int64_t curve_chunks(int level) {
  const int64_t total = (int64_t)1 << (2 * level);
  return (total + kChunk - 1) / kChunk;
}

hipError_t launch_curve_count(int level, int64_t w, int64_t h, int32_t* counts, int64_t chunks,
                              hipStream_t s) {
  const int64_t total = (int64_t)1 << (2 * level);
  curve_count_kernel<<<dim3((unsigned)chunks), dim3(kThreads), 0, s>>>(level, w, h, total, counts);
  return hipGetLastError();
}

hipError_t launch_curve_scan(int32_t* counts, int64_t chunks, hipStream_t s) {
  curve_scan_kernel<<<dim3(1), dim3(1024), 0, s>>>(counts, chunks);
  return hipGetLastError();
}

hipError_t launch_curve_emit(int level, int64_t w, int64_t h, const int32_t* offsets,
                             int64_t chunks, int32_t* order, hipStream_t s) {
  const int64_t total = (int64_t)1 << (2 * level);
  curve_emit_kernel<<<dim3((unsigned)chunks), dim3(kThreads), 0, s>>>(level, w, h, total, offsets,
                                                                       order);
  return hipGetLastError();
}

}  // namespace sa
