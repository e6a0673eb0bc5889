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
// and the other wavefronts.  The 16-entry queue is a register ring: the
// position loop is unrolled 16 times, so the oldest entry's slot is a
// compile-time constant and shifting the queue costs nothing.  The palette's
// search image (sa_palette.h) is staged in LDS once per workgroup; every
// workgroup serves blocks of a single frame.  The weights are wave-uniform
// and live in SGPRs.
//
// Output is fused (no index buffer, no separate scatter): the only pixel the
// block curve visits twice is the origin (position 0 and the trailing
// position), so every position except 0 owns its pixel.
#include "sa_internal.h"
#include "sa_math.h"

namespace sa {
namespace {


__device__ __forceinline__ void store_out(const BlockFrame& f, int64_t pix, int idx,
                                          const float4& e, bool assoc) {
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
    case kRgba8: {
      const float a = assoc ? e.w : load_pixel(f.src_fmt, f.src, pix).w;
      ((uint32_t*)f.dst)[pix] = scale_char(e.x) | (scale_char(e.y) << 8) |
                                (scale_char(e.z) << 16) | (scale_char(a) << 24);
      break;
    }
    case kRgb16: {
      uint16_t* d = (uint16_t*)f.dst + 3 * pix;
      d[0] = (uint16_t)e.x;
      d[1] = (uint16_t)e.y;
      d[2] = (uint16_t)e.z;
      break;
    }
    case kRgba16: {
      // rd_riemersma_cpu.cpp:747-752: alpha is the SOURCE alpha unless associated.
      const float a = assoc ? e.w : load_pixel(f.src_fmt, f.src, pix).w;
      ushort4 o;
      o.x = (uint16_t)e.x;
      o.y = (uint16_t)e.y;
      o.z = (uint16_t)e.z;
      o.w = (uint16_t)a;
      ((ushort4*)f.dst)[pix] = o;
      break;
    }
    default: {
      // rd_riemersma_cpu.cpp:704-708: r,g,b written, alpha only when associated.
      const float a = assoc ? e.w : load_pixel(f.src_fmt, f.src, pix).w;
      ((float4*)f.dst)[pix] = make_float4(e.x, e.y, e.z, a);
      break;
    }
  }
}

struct LSearch {
  const double* col;
  const float4* ent;
  const int2* sub;
  const int16_t* child;
  const uint16_t* cand;
  int kids;
};

// Per-lane descent and ordered scan (src/rd_riemersma_cpu.cpp:671-682 with
// src/rd_octree.cpp:178-217).
template <bool ASSOC>
__device__ __forceinline__ int lane_lookup(const LSearch& s, const double t[4]) {
  const unsigned c0 = scale_char(clamp_pixel(t[0]));
  const unsigned c1 = scale_char(clamp_pixel(t[1]));
  const unsigned c2 = scale_char(clamp_pixel(t[2]));
  const unsigned c3 = ASSOC ? scale_char(clamp_pixel(t[3])) : 0u;
  int row = 0, prev = 0;
  for (int lev = kMaxDepth - 1; lev > 0; --lev) {
    const int ch = s.child[row * s.kids + node_id(ASSOC, c0, c1, c2, c3, lev)];
    if (ch == kNoChild) break;
    prev = row;
    if (ch == kLeaf) break;
    row = ch;
  }
  const int2 r = s.sub[prev];
  double best = kFarDistance;
  int bidx = 0;
  for (int k = 0; k < r.y; ++k) {
    const int ci = s.cand[r.x + k];
    const float4 e4 = s.ent[ci];
    const float e[4] = {e4.x, e4.y, e4.z, e4.w};
    const double d = candidate_distance(ASSOC, e, t);
    if (d <= best) {
      best = d;
      bidx = ci;
    }
  }
  return bidx;
}

template <bool ASSOC>
struct Queue {
  double e[ASSOC ? 4 : 3][16];
};

// One position; R is the ring slot of the oldest queue entry.
template <int R, bool ASSOC>
__device__ __forceinline__ void block_step(Queue<ASSOC>& q, const double* __restrict__ sw,
                                           const BlockFrame& f, const LSearch& s,
                                           int64_t pix, bool write) {
  const float4 p = load_pixel(f.src_fmt, f.src, pix);
  double x[4];
  assoc_pixel(ASSOC, p.x, p.y, p.z, p.w, x);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const double wk = sw[k];
    x[0] += wk * q.e[0][(R + k) & 15];
    x[1] += wk * q.e[1][(R + k) & 15];
    x[2] += wk * q.e[2][(R + k) & 15];
    if (ASSOC) x[3] += wk * q.e[ASSOC ? 3 : 0][(R + k) & 15];
  }
  x[0] = (double)clamp_pixel(x[0]);
  x[1] = (double)clamp_pixel(x[1]);
  x[2] = (double)clamp_pixel(x[2]);
  if (ASSOC) x[3] = (double)clamp_pixel(x[3]);
  const int idx = lane_lookup<ASSOC>(s, x);
  const double* c = s.col + 4 * idx;
  q.e[0][R] = x[0] - c[0];
  q.e[1][R] = x[1] - c[1];
  q.e[2][R] = x[2] - c[2];
  if (ASSOC) q.e[ASSOC ? 3 : 0][R] = x[3] - c[3];
  if (write) store_out(f, pix, idx, s.ent[idx], ASSOC);
}

template <bool ASSOC, int... Rs>
__device__ __forceinline__ void block_group(Queue<ASSOC>& q, const double* __restrict__ sw,
                                            const BlockFrame& f, const LSearch& s,
                                            const int32_t* __restrict__ curve, int64_t m,
                                            int64_t i0, int64_t end,
                                            std::integer_sequence<int, Rs...>) {
  ((i0 + Rs < end
        ? (block_step<Rs, ASSOC>(q, sw, f, s, (i0 + Rs < m) ? (int64_t)curve[i0 + Rs] : 0,
                                 i0 + Rs != 0),
           0)
        : 0),
   ...);
}

template <bool ASSOC, int NT>
__global__ __launch_bounds__(NT) void blocks_kernel(const BlockFrame* __restrict__ frames,
                                                          const int32_t* __restrict__ curve,
                                                          int64_t nb, int block, int64_t nblocks,
                                                          const double* __restrict__ sw) {
  extern __shared__ __align__(16) unsigned char lds[];
  const BlockFrame f = frames[blockIdx.y];
  const PaletteDev pd = *f.pal;
  {
    const uint4* src = (const uint4*)pd.image;
    uint4* dst = (uint4*)lds;
    for (int k = threadIdx.x; k < pd.bytes / 16; k += NT) dst[k] = src[k];
    __syncthreads();
  }
  LSearch s;
  s.col = (const double*)lds;
  s.ent = (const float4*)(lds + pd.off_ent);
  s.sub = (const int2*)(lds + pd.off_sub);
  s.child = (const int16_t*)(lds + pd.off_child);
  s.cand = (const uint16_t*)(lds + pd.off_cand);
  s.kids = pd.kids;
  const int64_t b = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (b >= nblocks) return;
  const int64_t begin = b * block;
  const int64_t end = begin + block < nb ? begin + block : nb;
  const int64_t m = nb - 1;  // positions [0, m) follow the curve; m is the origin
  Queue<ASSOC> q;
#pragma unroll
  for (int c = 0; c < (ASSOC ? 4 : 3); ++c)
#pragma unroll
    for (int k = 0; k < 16; ++k) q.e[c][k] = 0.0;
  for (int64_t i0 = begin; i0 < end; i0 += 16)
    block_group<ASSOC>(q, sw, f, s, curve, m, i0, end, std::make_integer_sequence<int, 16>{});
}

// The pixel the block curve never reaches keeps its source value
// (src/rd_riemersma_cpu.cpp:776-803, src/rd_blocks_cuda.cu:389-443).
__global__ void blocks_fill_kernel(const BlockFrame* __restrict__ frames, int nframes,
                                   int64_t pix) {
  const int fi = blockIdx.x * blockDim.x + threadIdx.x;
  if (fi >= nframes) return;
  const BlockFrame f = frames[fi];
  const float4 p = load_pixel(f.src_fmt, f.src, pix);
  switch (f.dst_fmt) {
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

template <int NT>
void launch_blocks_nt(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t nb,
                      int block, int64_t nblocks, const double* sw, int assoc,
                      int lds_image_bytes, hipStream_t s) {
  const dim3 grid((unsigned)((nblocks + NT - 1) / NT), (unsigned)nframes);
  if (assoc)
    blocks_kernel<true, NT><<<grid, dim3(NT), (size_t)lds_image_bytes, s>>>(frames, curve, nb,
                                                                         block, nblocks, sw);
  else
    blocks_kernel<false, NT><<<grid, dim3(NT), (size_t)lds_image_bytes, s>>>(frames, curve, nb,
                                                                          block, nblocks, sw);
}

// `threads` (64, 128, 256 or 512) blocks per workgroup: every workgroup stages
// the palette once, so larger ones amortise it and smaller ones fill more CUs.
hipError_t launch_blocks(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t w,
                         int64_t h, int64_t nb, int block, const double* sw, int assoc,
                         int lds_image_bytes, int threads, hipStream_t s) {
  (void)w;
  (void)h;
  if (nframes <= 0 || nb <= 0) return hipSuccess;
  const int64_t nblocks = (nb + block - 1) / block;
  switch (threads) {
    case 64: launch_blocks_nt<64>(frames, nframes, curve, nb, block, nblocks, sw, assoc, lds_image_bytes, s); break;
    case 128: launch_blocks_nt<128>(frames, nframes, curve, nb, block, nblocks, sw, assoc, lds_image_bytes, s); break;
    case 512: launch_blocks_nt<512>(frames, nframes, curve, nb, block, nblocks, sw, assoc, lds_image_bytes, s); break;
    default: launch_blocks_nt<256>(frames, nframes, curve, nb, block, nblocks, sw, assoc, lds_image_bytes, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_blocks_fill(const BlockFrame* frames, int nframes, int64_t gap_pixel,
                              hipStream_t s) {
  if (nframes <= 0) return hipSuccess;
  blocks_fill_kernel<<<dim3((unsigned)((nframes + 63) / 64)), dim3(64), 0, s>>>(frames, nframes,
                                                                              gap_pixel);
  return hipGetLastError();
}

}  // namespace sa
