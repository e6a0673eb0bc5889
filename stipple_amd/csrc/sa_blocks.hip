// sa_blocks.hip -- the block-parallel walk.
//
// Restates RiemersmaBlocksCpu (src/rd_riemersma_cpu.cpp:477-805): the block
// curve (in-bounds of [0,4^L-1) plus a trailing origin) is cut into blocks of
// B positions; every block runs the sequential visit arithmetic from a zeroed
// 16-entry queue, with no memo, and the chosen colour is written to the pixel
// whose LAST visit this is.  Associated alpha is handled once, as the CPU
// engine does (the reference's CUDA kernel associates twice; see
// src/rd_blocks_cuda.cu:204-207 against src/rd_riemersma_cpu.cpp:652-682).
//
// Mapping to CDNA4: one lane per block, so a wavefront advances 64 blocks in
// lock step and the serial chain of each block is hidden by the other lanes
// and the other wavefronts.  A workgroup owns NT consecutive blocks and walks
// them 16 positions at a time:
//   1. stage -- the workgroup reads the 16 curve entries of each of its blocks
//      (16 consecutive ints per block, so a wavefront reads four 64-byte runs)
//      and gathers those pixels into LDS, position-major, as the batch's
//      widest source class (8-bit words, 16-bit pairs or float4);
//   2. walk -- each lane runs its block's 16 positions from LDS: no global
//      load on the serial chain;
//   3. write -- the workgroup writes the 16 x NT results back, each thread the
//      entries it staged (their pixel ids stay in its registers).
// The 16-entry queue is a register ring: the 16 positions of a chunk are
// unrolled, so the oldest entry's slot is a compile-time constant and
// shifting the queue costs nothing.  The palette's search image (sa_palette.h)
// is staged in LDS once per workgroup, with each candidate's distance
// constants (its colour, premultiplied when associated) as doubles beside it;
// every workgroup serves blocks of a single frame.  The weights are
// wave-uniform and live in SGPRs.
//
// Only the origin is visited twice (position 0 and the trailing position), so
// every position except 0 owns its pixel and the write needs no index buffer.
#include "sa_internal.h"
#include "sa_math.h"
#include "sa_search.h"

namespace sa {
namespace {

constexpr int kChunk = 16;  // positions per stage/walk/write round (the queue length)

// Source classes: every format of a batch is staged as the widest one.
enum SrcClass : int { kSrc8 = 0, kSrc16 = 1, kSrcF = 2 };

template <int SRC>
struct Raw;
template <>
struct Raw<kSrc8> {
  typedef uint32_t T;  // r | g << 8 | b << 16 | a << 24
};
template <>
struct Raw<kSrc16> {
  typedef uint2 T;  // r | g << 16, b | a << 16
};
template <>
struct Raw<kSrcF> {
  typedef float4 T;
};

// One source pixel in class SRC.  Widening is exact: an 8-bit sample is
// staged as itself (8-bit class) or as sample * 257 (16-bit class), which is
// the value load_pixel gives it; formats without alpha stage it opaque.
// Planar YUV (w x h planes) widens through swscale's matrix to 16 bits and is
// never in the 8-bit class.
template <int SRC>
__device__ __forceinline__ typename Raw<SRC>::T stage_pixel(int fmt, const void* base, int64_t p,
                                                            int64_t w, int64_t h) {
  typedef typename Raw<SRC>::T T;
  if constexpr (SRC == kSrc8) {
    if (fmt == kRgba8) return ((const uint32_t*)base)[p];
    const uint8_t* s = (const uint8_t*)base + 3 * p;
    return (T)s[0] | ((T)s[1] << 8) | ((T)s[2] << 16) | 0xFF000000u;
  } else if constexpr (SRC == kSrc16) {
    switch (fmt) {
      case kRgba16: return ((const uint2*)base)[p];
      case kYuv444:
      case kYuv420: return load_yuv(fmt, base, p, w, h);
      case kRgb16: {
        const uint16_t* s = (const uint16_t*)base + 3 * p;
        return make_uint2((uint32_t)s[0] | ((uint32_t)s[1] << 16), (uint32_t)s[2] | 0xFFFF0000u);
      }
      case kRgba8: {
        const uint32_t v = ((const uint32_t*)base)[p];
        return make_uint2((v & 255u) * 257u | (((v >> 8) & 255u) * 257u) << 16,
                          ((v >> 16) & 255u) * 257u | ((v >> 24) * 257u) << 16);
      }
      default: {
        const uint8_t* s = (const uint8_t*)base + 3 * p;
        return make_uint2((uint32_t)s[0] * 257u | ((uint32_t)s[1] * 257u) << 16,
                          (uint32_t)s[2] * 257u | 0xFFFF0000u);
      }
    }
  } else {
    return load_any(fmt, base, p, w, h);
  }
}

// The staged pixel as the Q16 floats load_pixel gives.
template <int SRC>
__device__ __forceinline__ float4 unstage(const typename Raw<SRC>::T& v) {
  if constexpr (SRC == kSrc8) {
    return make_float4((float)(v & 255u) * 257.0f, (float)((v >> 8) & 255u) * 257.0f,
                       (float)((v >> 16) & 255u) * 257.0f, (float)(v >> 24) * 257.0f);
  } else if constexpr (SRC == kSrc16) {
    return make_float4((float)(v.x & 0xFFFFu), (float)(v.x >> 16), (float)(v.y & 0xFFFFu),
                       (float)(v.y >> 16));
  } else {
    return v;
  }
}

template <int SRC>
__device__ __forceinline__ float unstage_alpha(const typename Raw<SRC>::T& v) {
  if constexpr (SRC == kSrc8) return (float)(v >> 24) * 257.0f;
  else if constexpr (SRC == kSrc16) return (float)(v.y >> 16);
  else return v.w;
}

// The output pixel for palette entry `e` (src/rd_riemersma_cpu.cpp:700-760):
// alpha is the entry's when associated, else the source pixel's (`a`).
// Planar output has planes of pp bytes.
__device__ __forceinline__ void store_out(const BlockFrame& f, int64_t pix, int idx,
                                          const float4& e, float a, int64_t pp) {
  switch (f.dst_fmt) {
    case kYuv444: store_yuv444(f.dst, pix, pp, e.x, e.y, e.z); break;
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
    default:
      // rd_riemersma_cpu.cpp:704-708: r,g,b written, alpha only when associated.
      ((float4*)f.dst)[pix] = make_float4(e.x, e.y, e.z, a);
      break;
  }
}

template <bool ASSOC>
struct Queue {
  double e[ASSOC ? 4 : 3][kChunk];
};

// Position S of a chunk (ring slot S holds the oldest queue entry).  `cur` is
// its staged pixel, read one position ahead; the scheduling barrier keeps the
// compiler from hoisting all sixteen reads (and their conversions) into
// registers at once.
template <int S, bool ASSOC, int SRC>
__device__ __forceinline__ void block_step(Queue<ASSOC>& q, const double* __restrict__ sw,
                                           const LSearch& s, const typename Raw<SRC>::T* raw,
                                           uint16_t* oidx, int stride, int lane, int live,
                                           typename Raw<SRC>::T& cur) {
  if (S >= live) return;
  typename Raw<SRC>::T nxt = cur;
  if (S + 1 < kChunk && S + 1 < live) nxt = raw[(S + 1) * stride + lane];
  const float4 p = unstage<SRC>(cur);
  double x[4];
  assoc_pixel(ASSOC, p.x, p.y, p.z, p.w, x);
#pragma unroll
  for (int k = 0; k < kQueue; ++k) {
    const double wk = sw[k];
    x[0] += wk * q.e[0][(S + k) & 15];
    x[1] += wk * q.e[1][(S + k) & 15];
    x[2] += wk * q.e[2][(S + k) & 15];
    if (ASSOC) x[3] += wk * q.e[ASSOC ? 3 : 0][(S + k) & 15];
  }
  float pc[4];
  pc[0] = clamp_pixel(x[0]);
  pc[1] = clamp_pixel(x[1]);
  pc[2] = clamp_pixel(x[2]);
  pc[3] = ASSOC ? clamp_pixel(x[3]) : 0.0f;
  x[0] = (double)pc[0];
  x[1] = (double)pc[1];
  x[2] = (double)pc[2];
  if (ASSOC) x[3] = (double)pc[3];
  const int idx = lane_lookup<ASSOC>(s, pc, x);
  const double* c = s.col + 4 * idx;
  q.e[0][S] = x[0] - c[0];
  q.e[1][S] = x[1] - c[1];
  q.e[2][S] = x[2] - c[2];
  if (ASSOC) q.e[ASSOC ? 3 : 0][S] = x[3] - c[3];
  oidx[S * stride + lane] = (uint16_t)idx;
  cur = nxt;
  __builtin_amdgcn_sched_barrier(0);
}

template <bool ASSOC, int SRC, int... Ss>
__device__ __forceinline__ void walk_chunk(Queue<ASSOC>& q, const double* __restrict__ sw,
                                           const LSearch& s, const typename Raw<SRC>::T* raw,
                                           uint16_t* oidx, int stride, int lane, int live,
                                           std::integer_sequence<int, Ss...>) {
  typename Raw<SRC>::T cur = raw[lane];
  (block_step<Ss, ASSOC, SRC>(q, sw, s, raw, oidx, stride, lane, live, cur), ...);
}

}  // namespace

// LDS bytes of a block-walk workgroup: the search image (img bytes, 16-byte
// multiple), the candidates' distance constants, and one chunk's staged
// pixels and result indices (rows of NT + 1 to spread the banks).
size_t blocks_lds_bytes(int img_bytes, int count, int src_class, int threads) {
  const size_t raw = src_class == kSrcF ? 16 : src_class == kSrc16 ? 8 : 4;
  const size_t row = (size_t)threads + 1;
  const size_t oidx = (2 * kChunk * row + 15) & ~(size_t)15;
  return (size_t)img_bytes + 32 * (size_t)count + raw * kChunk * row + oidx;
}

namespace {

// Workgroup w walks blocks [x * NT, x * NT + NT) of frame w % nframes, x =
// w / nframes: the frames of one block range are dispatched together, so the
// XCDs they land on share that range of the curve in their L2.
// Occupancy: 3 wavefronts per SIMD (2 with alpha), the queue's registers
// allowing; the 16 staged reads are otherwise hoisted and kept live.
template <bool ASSOC, int SRC, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(ASSOC ? 2 : 3)))
void blocks_kernel(const BlockFrame* __restrict__ frames,
                                                    int nframes,
                                                    const int32_t* __restrict__ curve,
                                                    int64_t nb, int block, int64_t nblocks,
                                                    const double* __restrict__ sw, int img_bytes,
                                                    int count, int64_t w, int64_t h) {
  typedef typename Raw<SRC>::T T;
  extern __shared__ __align__(16) unsigned char lds[];
  const int fi = (int)(blockIdx.x % (unsigned)nframes);
  const int64_t x = blockIdx.x / (unsigned)nframes;
  const BlockFrame f = frames[fi];
  const PaletteDev pd = *f.pal;
  const int tid = threadIdx.x;
  double4* dcent = (double4*)(lds + img_bytes);
  T* raw = (T*)(lds + img_bytes + 32 * count);
  uint16_t* oidx = (uint16_t*)(raw + kChunk * (NT + 1));
  const LSearch s = stage_search<ASSOC>(pd, lds, dcent, tid, NT);
  const int64_t b0 = x * NT;
  const int64_t m = nb - 1;  // positions [0, m) follow the curve; m is the origin
  // This lane's block and its length (0 past the last block).
  const int64_t mb = b0 + tid;
  const int64_t mbegin = mb * block;
  const int mlen = mb < nblocks ? (int)(nb - mbegin < block ? nb - mbegin : block) : 0;
  Queue<ASSOC> q;
#pragma unroll
  for (int c = 0; c < (ASSOC ? 4 : 3); ++c)
#pragma unroll
    for (int k = 0; k < kQueue; ++k) q.e[c][k] = 0.0;
  // Staged entry k of this thread: position c0 + sp of block b0 + jt + k NT/16
  // (sp = tid & 15, jt = tid >> 4), i.e. curve position ibase + c0 + k istep.
  // Past the curve's end (i >= nb) an entry is empty; so is every entry of a
  // block past the last one, whose positions all lie past the end.
  const int sp = tid & 15, jt = tid >> 4;
  const int64_t ibase = (b0 + jt) * block + sp;
  const int64_t istep = (int64_t)(NT / 16) * block;
  T* const rawk = raw + sp * (NT + 1) + jt;  // entry k at rawk[k NT/16]
  for (int c0 = 0; c0 < block; c0 += kChunk) {
    // The chunk's base is opaque to the optimiser, so the sixteen entries'
    // positions are formed here each chunk instead of being hoisted out of the
    // loop (and held in registers beside the queue).
    int64_t i0 = ibase + c0;
    asm volatile("" : "+v"(i0));
    const bool in_block = c0 + sp < block;
    // Four entries at a time: the queue is live throughout, so the loads of
    // all sixteen (and their addresses) must not be in flight at once.
    int32_t pix[kChunk];
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      const int64_t i = i0 + k * istep;
      pix[k] = -1;
      if (in_block && i < nb) {
        const int64_t p = i < m ? (int64_t)curve[i] : 0;
        rawk[k * (NT / 16)] = stage_pixel<SRC>(f.src_fmt, f.src, p, w, h);
        pix[k] = i != 0 ? (int32_t)p : -1;
      }
      if ((k & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    // Walk: this lane's block, positions c0 .. c0 + 15.
    const int live = mlen - c0;
    if (live > 0)
      walk_chunk<ASSOC, SRC>(q, sw, s, raw, oidx, NT + 1, tid, live,
                             std::make_integer_sequence<int, kChunk>{});
    __syncthreads();
    // Write: the entries this thread staged.
    uint16_t* const oidxk = oidx + sp * (NT + 1) + jt;
#pragma unroll
    for (int k = 0; k < kChunk; ++k) {
      if (pix[k] >= 0) {
        const int idx = oidxk[k * (NT / 16)];
        const float4 en = s.ent[idx];
        const float a = ASSOC ? en.w : unstage_alpha<SRC>(rawk[k * (NT / 16)]);
        store_out(f, pix[k], idx, en, a, w * h);
      }
      if ((k & 3) == 3) __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// The pixel the block curve never reaches keeps its source value
// (src/rd_riemersma_cpu.cpp:776-803, src/rd_blocks_cuda.cu:389-443); planar
// output gets the source converted, as the CPU engine's fused scatter does.
__global__ void blocks_fill_kernel(const BlockFrame* __restrict__ frames, int nframes,
                                   int64_t pix, int64_t w, int64_t h) {
  const int fi = blockIdx.x * blockDim.x + threadIdx.x;
  if (fi >= nframes) return;
  const BlockFrame f = frames[fi];
  const float4 p = load_any(f.src_fmt, f.src, pix, w, h);
  switch (f.dst_fmt) {
    case kYuv444: store_yuv444(f.dst, pix, w * h, p.x, p.y, p.z); break;
    case kIndex8: ((uint8_t*)f.dst)[pix] = 0; break;
    case kIndex16: ((uint16_t*)f.dst)[pix] = 0; break;
    case kRgb8: {
      uint8_t* d = (uint8_t*)f.dst + 3 * pix;
      d[0] = (uint8_t)scale_char(p.x);
      d[1] = (uint8_t)scale_char(p.y);
      d[2] = (uint8_t)scale_char(p.z);
      break;
    }
    case kRgba8:
      ((uint32_t*)f.dst)[pix] = scale_char(p.x) | (scale_char(p.y) << 8) |
                                (scale_char(p.z) << 16) | (scale_char(p.w) << 24);
      break;
    case kRgb16: {
      uint16_t* d = (uint16_t*)f.dst + 3 * pix;
      d[0] = (uint16_t)p.x;
      d[1] = (uint16_t)p.y;
      d[2] = (uint16_t)p.z;
      break;
    }
    case kRgba16: {
      ushort4 o;
      o.x = (uint16_t)p.x;
      o.y = (uint16_t)p.y;
      o.z = (uint16_t)p.z;
      o.w = (uint16_t)p.w;
      ((ushort4*)f.dst)[pix] = o;
      break;
    }
    default:
      ((float4*)f.dst)[pix] = p;
      break;
  }
}

}  // namespace

template <int SRC, int NT>
void launch_blocks_nt(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t nb,
                      int block, int64_t nblocks, const double* sw, int assoc, int img_bytes,
                      int count, int64_t w, int64_t h, size_t lds, hipStream_t s) {
  const dim3 grid((unsigned)(((nblocks + NT - 1) / NT) * nframes));
  if (assoc)
    blocks_kernel<true, SRC, NT><<<grid, dim3(NT), lds, s>>>(frames, nframes, curve, nb, block,
                                                             nblocks, sw, img_bytes, count, w, h);
  else
    blocks_kernel<false, SRC, NT><<<grid, dim3(NT), lds, s>>>(frames, nframes, curve, nb, block,
                                                              nblocks, sw, img_bytes, count, w, h);
}

template <int SRC>
void launch_blocks_src(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t nb,
                       int block, int64_t nblocks, const double* sw, int assoc, int img_bytes,
                       int count, int threads, int64_t w, int64_t h, size_t lds, hipStream_t s) {
  switch (threads) {
    case 64: launch_blocks_nt<SRC, 64>(frames, nframes, curve, nb, block, nblocks, sw, assoc, img_bytes, count, w, h, lds, s); break;
    case 128: launch_blocks_nt<SRC, 128>(frames, nframes, curve, nb, block, nblocks, sw, assoc, img_bytes, count, w, h, lds, s); break;
    case 512: launch_blocks_nt<SRC, 512>(frames, nframes, curve, nb, block, nblocks, sw, assoc, img_bytes, count, w, h, lds, s); break;
    default: launch_blocks_nt<SRC, 256>(frames, nframes, curve, nb, block, nblocks, sw, assoc, img_bytes, count, w, h, lds, s); break;
  }
}

int blocks_src_class(int fmt) {
  if (fmt == kRgbaF32) return kSrcF;
  return (fmt == kRgb16 || fmt == kRgba16 || fmt == kYuv444 || fmt == kYuv420) ? kSrc16 : kSrc8;
}

// `threads` (64, 128, 256 or 512) blocks per workgroup: every workgroup stages
// the palette once, so larger ones amortise it and smaller ones fill more CUs.
// A shape whose LDS would not fit is halved until it does (the result does not
// depend on the shape).
hipError_t launch_blocks(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t w,
                         int64_t h, int64_t nb, int block, const double* sw, int assoc,
                         int img_bytes, int count, int src_class, int threads, hipStream_t s) {
  if (nframes <= 0 || nb <= 0) return hipSuccess;
  if (threads != 64 && threads != 128 && threads != 512) threads = 256;
  while (threads > 64 && blocks_lds_bytes(img_bytes, count, src_class, threads) > kBlocksMaxLds)
    threads /= 2;
  const size_t lds = blocks_lds_bytes(img_bytes, count, src_class, threads);
  if (lds > kBlocksMaxLds) return hipErrorInvalidValue;
  const int64_t nblocks = (nb + block - 1) / block;
  switch (src_class) {
    case kSrcF: launch_blocks_src<kSrcF>(frames, nframes, curve, nb, block, nblocks, sw, assoc, img_bytes, count, threads, w, h, lds, s); break;
    case kSrc16: launch_blocks_src<kSrc16>(frames, nframes, curve, nb, block, nblocks, sw, assoc, img_bytes, count, threads, w, h, lds, s); break;
    default: launch_blocks_src<kSrc8>(frames, nframes, curve, nb, block, nblocks, sw, assoc, img_bytes, count, threads, w, h, lds, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_blocks_fill(const BlockFrame* frames, int nframes, int64_t gap_pixel,
                              int64_t w, int64_t h, hipStream_t s) {
  if (nframes <= 0) return hipSuccess;
  blocks_fill_kernel<<<dim3((unsigned)((nframes + 63) / 64)), dim3(64), 0, s>>>(
      frames, nframes, gap_pixel, w, h);
  return hipGetLastError();
}

}  // namespace sa
