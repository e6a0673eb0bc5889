// sa_jacobi.hip -- the Jacobi approximation of the walk.
//
// Restates the reference's approximate engine (src/rd_approx_cuda.cu:63-384):
// along the block curve (in-bounds of [0,4^L-1) plus the origin), with
// F = (I - H)^-1 truncated to `taps`+1 coefficients g (computed on the host,
// src/rd_approx_cuda.cu:73-88),
//
//     q0[i] = Lookup(x[i])
//     t[i]  = pal[q[i]] + sum_{d=0..min(taps,i)} g[d] * (x[i-d] - pal[q[i-d]])
//     q'[i] = Lookup(clamp(t[i]))                       (`iterations` sweeps)
//
// then an owner-filtered scatter.  The per-tap arithmetic is the reference's:
// one multiply and one add in T (float, or double with fp64), d ascending, no
// FMA -- so the result is the reference's fixed point iterate, not merely near it.
//
// Mapping to CDNA4: a sweep is a 1-D convolution along the curve.  Each
// workgroup stages the residual r = x - pal[q] of its tile plus a `taps` halo
// in LDS once (the reference re-derives it per tap); each thread then produces
// kPer consecutive outputs, sliding a kPer-wide register window down the tile,
// so one LDS load feeds kPer multiply-adds and the window rotates without moves
// (the tap loop is unrolled kPer times).  The palette search image sits in LDS
// beside the tile for the per-output octree lookup.
#include "sa_internal.h"
#include "sa_math.h"
#include "sa_search.h"

namespace sa {
namespace {

constexpr int kThreads = 256;
// Outputs per thread of a sweep (its tile is kThreads * per outputs): 2, 4 or
// 8; 4 unless tuned.
constexpr int kPerDefault = 4;

// d_nearest_linear (include/rd_cuda_common.cuh:451-476): ascending scan over
// the whole palette, d = dr^2 + (dg^2 + db^2), beta from the target's alpha.
template <bool ASSOC>
__device__ __forceinline__ int linear_lookup(const LSearch& s, const double t[4]) {
  double dist = kFarDistance;
  int best = 0;
  const double beta = kScale * t[3];
  for (int i = 0; i < s.count; ++i) {
    const float4 e = s.ent[i];
    const double alpha = ASSOC ? kScale * (double)e.w : 1.0;
    const double dr = alpha * (double)e.x + -beta * t[0];
    const double dg = alpha * (double)e.y + -beta * t[1];
    const double db = alpha * (double)e.z + -beta * t[2];
    double d = dr * dr + (dg * dg + db * db);
    if (ASSOC) {
      const double da = (double)e.w + -t[3];
      d = d + da * da;
    }
    if (d <= dist) {
      dist = d;
      best = i;
    }
  }
  return best;
}

// The tree search (sa_search.h) on the clamped target, or the linear scan.
template <bool ASSOC>
__device__ __forceinline__ int lookup(const LSearch& s, const double t[4], int use_tree) {
  if (!use_tree) return linear_lookup<ASSOC>(s, t);
  const float pc[4] = {clamp_pixel(t[0]), clamp_pixel(t[1]), clamp_pixel(t[2]),
                       ASSOC ? clamp_pixel(t[3]) : 0.0f};
  return lane_lookup<ASSOC>(s, pc, t);
}

// The search image without its first array (the associated colours, which
// the engine never reads) at lds[0, bytes - off_ent), and the candidates'
// distance constants in the room that array leaves: 32 bytes per colour,
// like it, so the workgroup's LDS is the image's size.
template <bool ASSOC>
__device__ __forceinline__ LSearch stage(const PaletteDev& pd, unsigned char* lds) {
  const int skip = pd.off_ent;
  const uint4* src = (const uint4*)((const unsigned char*)pd.image + skip);
  uint4* dst = (uint4*)lds;
  const int words = (pd.bytes - skip) / 16;
  for (int k = threadIdx.x; k < words; k += blockDim.x) dst[k] = src[k];
  double4* dcent = (double4*)(lds + (pd.bytes - skip));
  const float4* cent = (const float4*)((const unsigned char*)pd.image + pd.off_cent);
  for (int k = threadIdx.x; k < pd.count; k += blockDim.x) {
    const float4 e = cent[k];
    double a = 1.0;
    if (ASSOC) a = kScale * (double)e.w;
    dcent[k] = make_double4(a * (double)e.x, a * (double)e.y, a * (double)e.z, (double)e.w);
  }
  LSearch s;
  s.col = nullptr;
  s.ent = (const float4*)(lds + (pd.off_ent - skip));
  s.sub = (const int2*)(lds + (pd.off_sub - skip));
  s.desc = (const uint32_t*)(lds + (pd.off_desc - skip));
  s.child = (const int16_t*)(lds + (pd.off_child - skip));
  s.cand = (const uint16_t*)(lds + (pd.off_cand - skip));
  s.dcent = dcent;
  s.kids = pd.kids;
  s.count = pd.count;
  return s;
}

// Gather into curve order: cx[i] = pixel(curve position i), T per channel.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gather(const BlockFrame* __restrict__ frames,
                                                     const int32_t* __restrict__ curve, int64_t nb,
                                                     T* __restrict__ cx) {
  const BlockFrame f = frames[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nb) return;
  const int64_t pix = i < nb - 1 ? (int64_t)curve[i] : 0;
  const float4 p = load_pixel(f.src_fmt, f.src, pix);
  T* o = cx + 4 * ((int64_t)blockIdx.y * nb + i);
  o[0] = (T)p.x;
  o[1] = (T)p.y;
  o[2] = (T)p.z;
  o[3] = (T)p.w;
}

// InitIndexKernel (src/rd_approx_cuda.cu:131-152): nearest colour of the raw pixel.
template <typename T, bool ASSOC>
__global__ __launch_bounds__(kThreads) void k_init(const BlockFrame* __restrict__ frames,
                                                   const T* __restrict__ cx, int64_t nb,
                                                   int use_tree, uint8_t* __restrict__ q) {
  extern __shared__ __align__(16) unsigned char lds[];
  const BlockFrame f = frames[blockIdx.y];
  const PaletteDev pd = *f.pal;
  const LSearch s = stage<ASSOC>(pd, lds);
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nb) return;
  const T* x = cx + 4 * ((int64_t)blockIdx.y * nb + i);
  double t[4];
  assoc_pixel(ASSOC, (float)x[0], (float)x[1], (float)x[2], (float)x[3], t);
  q[(int64_t)blockIdx.y * nb + i] = (uint8_t)lookup<ASSOC>(s, t, use_tree);
}

// The filter's slots in LDS (taps + 1 rounded up to 8, zero padded) and its
// offset after n values of the residual planes (16-byte aligned).
__host__ __device__ inline int g_slots(int taps) { return (taps + 1 + 7) & ~7; }
template <typename T>
__host__ __device__ inline int gl_offset(int n) {
  return (int)(((size_t)n * sizeof(T) + 15) / 16 * 16 / sizeof(T));
}

// One sweep (JacobiStepKernel, src/rd_approx_cuda.cu:158-228).
template <typename T, bool ASSOC, int kPer>
__global__ __launch_bounds__(kThreads) void k_sweep(const BlockFrame* __restrict__ frames,
                                                    const T* __restrict__ cx,
                                                    const uint8_t* __restrict__ q_in,
                                                    uint8_t* __restrict__ q_out,
                                                    const T* __restrict__ g, int taps, int64_t nb,
                                                    int use_tree, int img_bytes) {
  extern __shared__ __align__(16) unsigned char lds[];
  constexpr int C = ASSOC ? 4 : 3;
  const BlockFrame f = frames[blockIdx.y];
  const PaletteDev pd = *f.pal;
  const LSearch s = stage<ASSOC>(pd, lds);
  __syncthreads();  // the tile below reads the staged entries
  constexpr int kTile = kThreads * kPer;
  const int window = kTile + taps;
  T* r = (T*)(lds + img_bytes);  // C planes of `window` values
  // The filter beside the planes, padded with zeros to a whole unrolled step:
  // read from LDS, the taps keep the loop's waits in order (a scalar load per
  // tap shares their counter and would drain every LDS read in flight).
  T* gl = r + gl_offset<T>(C * window);
  for (int k = threadIdx.x; k < g_slots(taps); k += kThreads) gl[k] = k <= taps ? g[k] : T(0);
  const int64_t fbase = (int64_t)blockIdx.y * nb;
  const int64_t tile = (int64_t)blockIdx.x * kTile;
  const int64_t base = tile - taps;
  for (int k = threadIdx.x; k < window; k += kThreads) {
    const int64_t m = base + k;
    if (m >= 0 && m < nb) {
      const T* x = cx + 4 * (fbase + m);
      const float4 e = s.ent[q_in[fbase + m]];
      r[k] = x[0] - (T)e.x;
      r[window + k] = x[1] - (T)e.y;
      r[2 * window + k] = x[2] - (T)e.z;
      if (ASSOC) r[3 * window + k] = x[3] - (T)e.w;
    } else {
      for (int c = 0; c < C; ++c) r[c * window + k] = T(0);
    }
  }
  __syncthreads();
  const int o0 = threadIdx.x * kPer;     // first output of this thread, in the tile
  const int64_t i0 = tile + o0;
  if (i0 >= nb) return;
  T acc[kPer][C];
#pragma unroll
  for (int j = 0; j < kPer; ++j)
#pragma unroll
    for (int c = 0; c < C; ++c) acc[j][c] = T(0);
  if (i0 >= taps) {
    // Every output has the full taps+1 terms: slide the window.
    T wv[C][kPer];
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int j = 0; j < kPer; ++j) wv[c][j] = r[c * window + o0 + taps + j];
    int d = 0;
    // Unrolled by kPer: at step d output j reads slot (j - d) mod kPer.
    for (; d + kPer <= taps + 1; d += kPer) {
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const T gk = gl[d + u];
#pragma unroll
        for (int j = 0; j < kPer; ++j)
#pragma unroll
          for (int c = 0; c < C; ++c) acc[j][c] += gk * wv[c][(j - u + kPer) % kPer];
        // Slot (kPer-1-u) mod kPer now holds the value no output needs again;
        // refill it with the next step's new entry, window slot
        // o0 + taps - (d+u) - 1.  After the last step (d+u == taps) the slot is
        // never read; its load (slot o0 - 1, at worst the word before the
        // planes, inside the staged search image) is left unconditional.
#pragma unroll
        for (int c = 0; c < C; ++c)
          wv[c][(kPer - 1 - u) % kPer] = r[c * window + o0 + taps - (d + u) - 1];
      }
    }
    for (; d <= taps; ++d) {
      const T gk = gl[d];
#pragma unroll
      for (int j = 0; j < kPer; ++j)
#pragma unroll
        for (int c = 0; c < C; ++c) acc[j][c] += gk * r[c * window + o0 + taps + j - d];
    }
  } else {
    // Start of the curve: output i sums d = 0..i only.
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int64_t i = i0 + j;
      const int dmax = (int)(i < taps ? i : taps);
      for (int d = 0; d <= dmax; ++d) {
        const T gk = gl[d];
#pragma unroll
        for (int c = 0; c < C; ++c) acc[j][c] += gk * r[c * window + o0 + taps + j - d];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int64_t i = i0 + j;
    if (i >= nb) break;
    const float4 e = s.ent[q_in[fbase + i]];
    const double vr = (double)e.x + (double)acc[j][0];
    const double vg = (double)e.y + (double)acc[j][1];
    const double vb = (double)e.z + (double)acc[j][2];
    const double va = ASSOC ? (double)e.w + (double)acc[j][ASSOC ? 3 : 0] : kRange;
    // d_associate_from_double (include/rd_cuda_common.cuh:57-72).
    double t[4];
    const double ca = (double)clamp_pixel(va);
    t[3] = ca;
    if (!ASSOC || va == kRange) {
      t[0] = (double)clamp_pixel(vr);
      t[1] = (double)clamp_pixel(vg);
      t[2] = (double)clamp_pixel(vb);
    } else {
      const double al = kScale * ca;
      t[0] = al * (double)clamp_pixel(vr) + 0.0;
      t[1] = al * (double)clamp_pixel(vg) + 0.0;
      t[2] = al * (double)clamp_pixel(vb) + 0.0;
    }
    q_out[fbase + i] = (uint8_t)lookup<ASSOC>(s, t, use_tree);
  }
}

// ScatterKernel (src/rd_approx_cuda.cu:90-109): owner-filtered; position 0 is
// superseded by the trailing origin.
__global__ __launch_bounds__(kThreads) void k_scatter(const BlockFrame* __restrict__ frames,
                                                      const int32_t* __restrict__ curve,
                                                      int64_t nb, const uint8_t* __restrict__ q) {
  const BlockFrame f = frames[blockIdx.y];
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= nb || i == 0) return;
  const int64_t pix = i < nb - 1 ? (int64_t)curve[i] : 0;
  const int idx = q[(int64_t)blockIdx.y * nb + i];
  const PaletteDev* pd = f.pal;
  const float4 e = ((const float4*)((const unsigned char*)pd->image + pd->off_ent))[idx];
  const bool assoc = pd->assoc != 0;
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
                                (scale_char(e.z) << 16) |
                                (scale_char(assoc ? e.w : 65535.0f) << 24);
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
      o.w = (uint16_t)(assoc ? e.w : 65535.0f);
      ((ushort4*)f.dst)[pix] = o;
      break;
    }
    default:
      // The reference writes alpha 65535 unless associated (rd_approx_cuda.cu:107).
      ((float4*)f.dst)[pix] = make_float4(e.x, e.y, e.z, assoc ? e.w : 65535.0f);
      break;
  }
}

template <typename K>
hipError_t allow_lds(K kernel, size_t bytes) {
  return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)bytes);
}

template <typename T, bool ASSOC, int kPer>
hipError_t run(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t nb,
               int iterations, int taps, const T* g, int use_tree, int img_bytes, T* cx,
               uint8_t* q0, uint8_t* q1, uint8_t** result, hipStream_t s) {
  constexpr int kTile = kThreads * kPer;
  const dim3 grid_pos((unsigned)((nb + kThreads - 1) / kThreads), (unsigned)nframes);
  k_gather<T><<<grid_pos, kThreads, 0, s>>>(frames, curve, nb, cx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = allow_lds(k_init<T, ASSOC>, (size_t)img_bytes);
  if (e != hipSuccess) return e;
  k_init<T, ASSOC><<<grid_pos, kThreads, (size_t)img_bytes, s>>>(frames, cx, nb, use_tree, q0);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t lds = (size_t)img_bytes +
                    sizeof(T) * ((size_t)gl_offset<T>((ASSOC ? 4 : 3) * (kTile + taps)) +
                                 (size_t)g_slots(taps));
  e = allow_lds(k_sweep<T, ASSOC, kPer>, lds);
  if (e != hipSuccess) return e;
  const dim3 grid_tile((unsigned)((nb + kTile - 1) / kTile), (unsigned)nframes);
  for (int it = 0; it < iterations; ++it) {
    k_sweep<T, ASSOC, kPer><<<grid_tile, kThreads, lds, s>>>(frames, cx, q0, q1, g, taps, nb,
                                                             use_tree, img_bytes);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint8_t* t = q0;
    q0 = q1;
    q1 = t;
  }
  k_scatter<<<grid_pos, kThreads, 0, s>>>(frames, curve, nb, q0);
  *result = q0;
  return hipGetLastError();
}

}  // namespace

size_t approx_lds_bytes(int img_bytes, int taps, int fp64, int assoc, int per) {
  const int tile = kThreads * (per == 2 || per == 8 ? per : kPerDefault);
  const int n = (assoc ? 4 : 3) * (tile + taps);
  const size_t planes = fp64 ? (size_t)gl_offset<double>(n) * 8 : (size_t)gl_offset<float>(n) * 4;
  return (size_t)img_bytes + planes + (fp64 ? 8 : 4) * (size_t)g_slots(taps);
}

template <int kPer>
hipError_t run_per(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t nb,
                   int iterations, int taps, const void* g, int fp64, int use_tree, int assoc,
                   int img_bytes, void* cx, uint8_t* q0, uint8_t* q1, uint8_t** result,
                   hipStream_t s) {
  if (fp64) {
    return assoc ? run<double, true, kPer>(frames, nframes, curve, nb, iterations, taps,
                                           (const double*)g, use_tree, img_bytes, (double*)cx,
                                           q0, q1, result, s)
                 : run<double, false, kPer>(frames, nframes, curve, nb, iterations, taps,
                                            (const double*)g, use_tree, img_bytes, (double*)cx,
                                            q0, q1, result, s);
  }
  return assoc ? run<float, true, kPer>(frames, nframes, curve, nb, iterations, taps,
                                        (const float*)g, use_tree, img_bytes, (float*)cx, q0, q1,
                                        result, s)
               : run<float, false, kPer>(frames, nframes, curve, nb, iterations, taps,
                                         (const float*)g, use_tree, img_bytes, (float*)cx, q0, q1,
                                         result, s);
}

hipError_t launch_approx(const BlockFrame* frames, int nframes, const int32_t* curve, int64_t nb,
                         int iterations, int taps, const void* g, int fp64, int use_tree,
                         int assoc, int img_bytes, void* cx, uint8_t* q0, uint8_t* q1,
                         uint8_t** result, int per, hipStream_t s) {
  if (nframes <= 0 || nb <= 0) return hipSuccess;
  switch (per) {
    case 2:
      return run_per<2>(frames, nframes, curve, nb, iterations, taps, g, fp64, use_tree, assoc,
                        img_bytes, cx, q0, q1, result, s);
    case 8:
      return run_per<8>(frames, nframes, curve, nb, iterations, taps, g, fp64, use_tree, assoc,
                        img_bytes, cx, q0, q1, result, s);
    default:
      return run_per<kPerDefault>(frames, nframes, curve, nb, iterations, taps, g, fp64, use_tree,
                                  assoc, img_bytes, cx, q0, q1, result, s);
  }
}

}  // namespace sa
