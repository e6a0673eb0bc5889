// sa_diffuse.hip -- the error-diffusion dithers of the plugin seam on the
// device: floyd-steinberg, atkinson and jarvis (src/rd_dithers.cc:74-230).
//
// Each reference plugin scans a frame in raster order in float arithmetic:
// the pixel plus diffusion x its accumulated error picks the nearest entry
// (NearestEntry, src/rd_dither_common.h: float squared distance, the first of
// equal minima), the entry's r, g, b are written (alpha untouched), and the
// residual is added, weighted, into the error of later pixels
// (e[t] = e[t] + w * residual, one rounding each).  A pixel's error is thus a
// fixed sum, in source raster order, of weighted residuals of a few earlier
// pixels; both kernels reproduce each such sum in exactly that order.
//
// floyd-steinberg and atkinson reach only the row below (and two below), so a
// row can run two pixels behind the row above it: one wavefront per frame,
// lane L owning row y0+L of a 64-row strip, pixel X of lane L at step
// X + 2L.  After each step every lane hands its residual (and, for atkinson,
// the residual it holds from two rows up) to the next lane (wave shift by
// one); a lane keeps the three residuals of the row above around its column
// and its own last two, and forms its pixel's error from them in the
// reference's order.  Strips run one after another; the last two rows' residuals
// carry over to the next strip through LDS.
//
// jarvis cannot overlap rows: the reference forms tap columns as size_t
// x + dx, so at x = 1 a dx = -2 tap wraps to the previous row's last column
// and at x = w-2 a dx = +2 tap wraps to the next row's first column -- the
// first pixel of a row needs the second-last pixel of the row above.  It runs
// one frame per wavefront in raster order: the lanes share each pixel's
// nearest-entry search, and the error of the three rows a pixel can reach is
// an LDS ring; taps into pixels already dithered are dropped (the reference
// adds them but never reads them again).
#include "sa_internal.h"
#include "sa_math.h"

namespace sa {
namespace {

constexpr int kWave = 64;

struct F3 {
  float r, g, b;
};

__device__ __forceinline__ F3 f3(float r, float g, float b) { return F3{r, g, b}; }

// acc = acc + w * v, per channel, one rounding each (Accumulate).
__device__ __forceinline__ void acc_add(F3& acc, float w, const F3& v) {
  acc.r = acc.r + w * v.r;
  acc.g = acc.g + w * v.g;
  acc.b = acc.b + w * v.b;
}

// Global-address-space views: their loads and stores count only against the
// vector-memory counter, so the LDS waits of a step (shuffles, carried rows,
// error ring) never wait on a pixel load or store in flight.
typedef const uint8_t __attribute__((address_space(1))) gcu8;
typedef const uint16_t __attribute__((address_space(1))) gcu16;
typedef const uint32_t __attribute__((address_space(1))) gcu32;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef const u32x2 __attribute__((address_space(1))) gcu2;
typedef const u32x4 __attribute__((address_space(1))) gcu4;
typedef uint8_t __attribute__((address_space(1))) gu8;
typedef uint16_t __attribute__((address_space(1))) gu16;
typedef uint32_t __attribute__((address_space(1))) gu32;
typedef u16x4 __attribute__((address_space(1))) gus4;
typedef f32x4 __attribute__((address_space(1))) gf4;

// A source pixel's raw bits, loaded early and converted when used (load_pixel's
// conversions: 8-bit x257, 16-bit as-is, alpha 65535 without an alpha
// channel).  Planar YUV is converted at load (load_any).
template <int FMT>
__device__ __forceinline__ uint4 load_raw(const void* src, int64_t p, int64_t w, int64_t h) {
  switch (FMT) {
    case kRgb8: {
      gcu8* q = (gcu8*)src + 3 * p;
      return make_uint4(q[0], q[1], q[2], 0);
    }
    case kRgba8: return make_uint4(((gcu32*)src)[p], 0, 0, 0);
    case kRgb16: {
      gcu16* q = (gcu16*)src + 3 * p;
      return make_uint4(q[0], q[1], q[2], 0);
    }
    case kRgba16: {
      const u32x2 v = ((gcu2*)src)[p];
      return make_uint4(v.x, v.y, 0, 0);
    }
    case kRgbaF32: {
      const u32x4 v = ((gcu4*)src)[p];
      return make_uint4(v.x, v.y, v.z, v.w);
    }
    default: {
      const float4 v = load_any(FMT, src, p, w, h);
      return make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z),
                        __float_as_uint(v.w));
    }
  }
}

template <int FMT>
__device__ __forceinline__ float4 raw_pixel(uint4 r) {
  switch (FMT) {
    case kRgb8:
      return make_float4((float)r.x * 257.0f, (float)r.y * 257.0f, (float)r.z * 257.0f, 65535.0f);
    case kRgba8:
      return make_float4((float)(r.x & 255u) * 257.0f, (float)((r.x >> 8) & 255u) * 257.0f,
                         (float)((r.x >> 16) & 255u) * 257.0f, (float)(r.x >> 24) * 257.0f);
    case kRgb16: return make_float4((float)r.x, (float)r.y, (float)r.z, 65535.0f);
    case kRgba16:
      return make_float4((float)(r.x & 0xFFFFu), (float)(r.x >> 16), (float)(r.y & 0xFFFFu),
                         (float)(r.y >> 16));
    default:
      return make_float4(__uint_as_float(r.x), __uint_as_float(r.y), __uint_as_float(r.z),
                         __uint_as_float(r.w));
  }
}

// store_pick through global-space stores (planar 4:4:4 through store_pick).
__device__ __forceinline__ void store_choice(const ScreenFrame& f, int64_t pix, int64_t pp, int idx,
                                             const float4& e, float a) {
  switch (f.dst_fmt) {
    case kIndex8: ((gu8*)f.dst)[pix] = (uint8_t)idx; break;
    case kIndex16: ((gu16*)f.dst)[pix] = (uint16_t)idx; break;
    case kRgb8: {
      gu8* d = (gu8*)f.dst + 3 * pix;
      d[0] = (uint8_t)scale_char(e.x);
      d[1] = (uint8_t)scale_char(e.y);
      d[2] = (uint8_t)scale_char(e.z);
      break;
    }
    case kRgba8:
      ((gu32*)f.dst)[pix] = scale_char(e.x) | (scale_char(e.y) << 8) | (scale_char(e.z) << 16) |
                            (scale_char(a) << 24);
      break;
    case kRgb16: {
      gu16* d = (gu16*)f.dst + 3 * pix;
      d[0] = (uint16_t)e.x;
      d[1] = (uint16_t)e.y;
      d[2] = (uint16_t)e.z;
      break;
    }
    case kRgba16: {
      const u16x4 o = {(uint16_t)e.x, (uint16_t)e.y, (uint16_t)e.z, (uint16_t)a};
      ((gus4*)f.dst)[pix] = o;
      break;
    }
    case kRgbaF32: {
      const f32x4 o = {e.x, e.y, e.z, a};
      ((gf4*)f.dst)[pix] = o;
      break;
    }
    default: store_pick(f, pix, pp, idx, e, a); break;
  }
}

__device__ __forceinline__ F3 shift_up(const F3& v) {
  return F3{__shfl_up(v.r, 1, kWave), __shfl_up(v.g, 1, kWave), __shfl_up(v.b, 1, kWave)};
}

// NearestEntry over the frame's entries, staged in LDS (every lane reads the
// same entry: a broadcast): the first entry is taken unconditionally, then
// only a strictly smaller distance replaces it.
__device__ __forceinline__ int nearest(const float4* __restrict__ ent, int count, float r, float g,
                                       float b, float4& c) {
  float4 e = ent[0];
  float dr = r - e.x, dg = g - e.y, db = b - e.z;
  float bd = dr * dr + dg * dg + db * db;
  int bi = 0;
  c = e;
#pragma unroll 4
  for (int i = 1; i < count; ++i) {
    e = ent[i];
    dr = r - e.x;
    dg = g - e.y;
    db = b - e.z;
    const float d = dr * dr + dg * dg + db * db;
    if (d < bd) {
      bd = d;
      bi = i;
      c = e;
    }
  }
  return bi;
}

// KIND 0: floyd-steinberg (src/rd_dithers.cc:74-110); 1: atkinson (:112-148).
// FMT: the frame's source format, fixed for the loop.
template <int KIND, int FMT>
__device__ __forceinline__ void rows_frame(const ScreenFrame& f, const float4* __restrict__ ent,
                                           int count, float* carry, int64_t w, int64_t h,
                                           float diffusion) {
  const int lane = threadIdx.x;
  const int64_t pp = w * h;
  float* c1r = carry;
  float* c1g = carry + w;
  float* c1b = carry + 2 * w;
  float* c2r = carry + 3 * w;
  float* c2g = carry + 4 * w;
  float* c2b = carry + 5 * w;
  const F3 zero = f3(0.f, 0.f, 0.f);
  for (int64_t y0 = 0; y0 < h; y0 += kWave) {
    const int rows = (int)(h - y0 < kWave ? h - y0 : kWave);
    const int64_t y = y0 + lane;
    const bool live = lane < rows;
    const bool more = y0 + kWave < h;  // a strip follows: carry the last rows' residuals
    F3 um1 = zero, u0 = zero, up1 = zero;  // residuals of the row above at X-1, X, X+1
    F3 vprev = zero;                       // atkinson: residual two rows up at X
    F3 own1 = zero, own2 = zero;           // own residuals at X-1, X-2
    F3 pub = zero, pub_u = zero;           // handed to the next lane after each step
    if (lane == 0) {
      // Lane 0 has no step before its first pixel to receive column 0 of the
      // rows above in; it takes them from the carried rows.
      if (y0 >= 1) up1 = f3(c1r[0], c1g[0], c1b[0]);
      if (KIND == 1 && y0 >= 2) vprev = f3(c2r[0], c2g[0], c2b[0]);
    }
    // The row's first pixel (lane 0 has no step before its first pixel to
    // prefetch it in).
    uint4 pcur = live ? load_raw<FMT>(f.src, y * w, w, h) : make_uint4(0, 0, 0, 0);
    const int64_t steps = w + 2 * (int64_t)(rows - 1);
    for (int64_t t = 0; t < steps; ++t) {
      const int64_t X = t - 2 * (int64_t)lane;
      // What the lane above produced last step: its residual at X+1 (and, for
      // atkinson, the residual two rows up it held at X+1).  Lane 0 reads the
      // previous strip's rows.
      F3 in_e = shift_up(pub);
      F3 in_u = KIND == 1 ? shift_up(pub_u) : zero;
      if (lane == 0) {
        const bool ok = X + 1 >= 0 && X + 1 < w;
        in_e = (ok && y0 >= 1) ? f3(c1r[X + 1], c1g[X + 1], c1b[X + 1]) : zero;
        if (KIND == 1) in_u = (ok && y0 >= 2) ? f3(c2r[X + 1], c2g[X + 1], c2b[X + 1]) : zero;
      }
      um1 = u0;
      u0 = up1;
      up1 = in_e;
      const F3 v2 = vprev;  // atkinson: the residual two rows up at X
      vprev = in_u;
      // Prefetch the next step's pixel.
      uint4 pnext = make_uint4(0, 0, 0, 0);
      if (live && X >= -1 && X + 1 < w) pnext = load_raw<FMT>(f.src, y * w + X + 1, w, h);
      F3 e_out = zero;
      if (live && X >= 0 && X < w) {
        // This pixel's error, summed in the reference's source order.
        F3 acc = zero;
        if (KIND == 0) {
          if (y >= 1) {
            if (X >= 1) acc_add(acc, 1.0f / 16.0f, um1);
            if (X == 0) acc_add(acc, 3.0f / 16.0f, u0);
            acc_add(acc, 5.0f / 16.0f, u0);
            if (X + 1 < w) acc_add(acc, 3.0f / 16.0f, up1);
          }
          if (X >= 1) acc_add(acc, 7.0f / 16.0f, own1);
        } else {
          const float k = 1.0f / 8.0f;
          if (y >= 2) acc_add(acc, k, v2);
          if (y >= 1) {
            if (X >= 1) acc_add(acc, k, um1);
            acc_add(acc, k, u0);
            if (X + 1 < w) acc_add(acc, k, up1);
          }
          if (X >= 2) acc_add(acc, k, own2);
          if (X >= 1) acc_add(acc, k, own1);
        }
        const float4 p = raw_pixel<FMT>(pcur);
        const float r = p.x + acc.r * diffusion;
        const float g = p.y + acc.g * diffusion;
        const float b = p.z + acc.b * diffusion;
        float4 c;
        const int bi = nearest(ent, count, r, g, b, c);
        e_out = f3(r - c.x, g - c.y, b - c.z);
        store_choice(f, y * w + X, pp, bi, c, p.w);
        own2 = own1;
        own1 = e_out;
        if (more && lane == rows - 1) {
          c1r[X] = e_out.r;
          c1g[X] = e_out.g;
          c1b[X] = e_out.b;
        }
        if (KIND == 1 && more && lane == rows - 2) {
          c2r[X] = e_out.r;
          c2g[X] = e_out.g;
          c2b[X] = e_out.b;
        }
      }
      pub = e_out;
      pub_u = u0;
      // What the next step reads: pixel X+1 when this step loaded it, and a
      // pixel no step of this lane reads otherwise.
      pcur = pnext;
    }
    __syncthreads();  // the carried rows are complete before the next strip reads them
  }
}

template <int KIND>
__global__ __launch_bounds__(kWave) void k_diffuse_rows(const ScreenFrame* __restrict__ frames,
                                                        int64_t w, int64_t h, float diffusion) {
  // LDS: the entries (float4 each), then the carried rows' r, g, b planes:
  // row y0-1, then (atkinson) row y0-2.
  extern __shared__ float4 lds4[];
  const ScreenFrame f = frames[blockIdx.x];
  const int count = __builtin_amdgcn_readfirstlane(f.pal->count);
  const int lane = threadIdx.x;
  float4* ent = lds4;
  {
    const float4* g = (const float4*)((const unsigned char*)f.pal->image + f.pal->off_ent);
    for (int k = lane; k < count; k += kWave) ent[k] = g[k];
    __syncthreads();
  }
  float* carry = (float*)(lds4 + count);
  switch (f.src_fmt) {
    case kRgb8: rows_frame<KIND, kRgb8>(f, ent, count, carry, w, h, diffusion); break;
    case kRgba8: rows_frame<KIND, kRgba8>(f, ent, count, carry, w, h, diffusion); break;
    case kRgb16: rows_frame<KIND, kRgb16>(f, ent, count, carry, w, h, diffusion); break;
    case kRgba16: rows_frame<KIND, kRgba16>(f, ent, count, carry, w, h, diffusion); break;
    case kRgbaF32: rows_frame<KIND, kRgbaF32>(f, ent, count, carry, w, h, diffusion); break;
    case kYuv444: rows_frame<KIND, kYuv444>(f, ent, count, carry, w, h, diffusion); break;
    default: rows_frame<KIND, kYuv420>(f, ent, count, carry, w, h, diffusion); break;
  }
}

template <int CTRL>
__device__ __forceinline__ float dppf(float v, float fill) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(fill), __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// The first index at the wave's smallest d (d >= 0 or +inf, never NaN; lane l
// offers its own first minimum li): a row-wise DPP minimum, the four row
// minima, then the lowest index among the lanes holding it.
__device__ __forceinline__ int wave_argmin(float d, int li) {
  const float inf = __builtin_inff();
  float m = d;
  m = fminf(m, dppf<0x111>(m, inf));  // row_shr:1
  m = fminf(m, dppf<0x112>(m, inf));  // row_shr:2
  m = fminf(m, dppf<0x114>(m, inf));  // row_shr:4
  m = fminf(m, dppf<0x118>(m, inf));  // row_shr:8
  const float m0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), 15));
  const float m1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), 31));
  const float m2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), 47));
  const float m3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m), 63));
  const float mm = fminf(fminf(m0, m1), fminf(m2, m3));
  uint64_t tied = __ballot(d == mm);
  int best = 0x7FFFFFFF;
  while (tied) {
    const int l = __builtin_ctzll(tied);
    tied &= tied - 1;
    const int v = __builtin_amdgcn_readlane(li, l);
    best = v < best ? v : best;
  }
  return best;
}

struct JTap {
  int dy, dx;
  float w;
};
// The taps of src/rd_dithers.cc:200-204, in their order.
__constant__ JTap kJarvis[12] = {
    {2, -2, 1.0f / 48.0f}, {2, -1, 3.0f / 48.0f}, {2, 0, 5.0f / 48.0f}, {2, 1, 3.0f / 48.0f},
    {1, -2, 7.0f / 48.0f}, {1, -1, 5.0f / 48.0f}, {1, 0, 7.0f / 48.0f}, {1, 1, 5.0f / 48.0f},
    {0, -2, 3.0f / 48.0f}, {0, -1, 1.0f / 48.0f}, {0, 1, 3.0f / 48.0f}, {0, 2, 1.0f / 48.0f}};

// The target pixel of tap k from (x, y), as the reference forms it in size_t
// arithmetic (src/rd_dithers.cc:205-212), or UINT64_MAX when the reference's
// Accumulate skips it (index >= n) or it is a pixel already dithered (<= i).
__device__ __forceinline__ uint64_t jarvis_target(int k, uint64_t x, uint64_t y, uint64_t w,
                                                  uint64_t h, uint64_t i) {
  const JTap t = kJarvis[k];
  uint64_t tx = x + (uint64_t)(int64_t)t.dx;
  uint64_t ty = y + (uint64_t)(int64_t)t.dy;
  if (t.dx < 0 && x == 0) tx = 0;
  if (t.dx > 0 && x + 1 >= w) tx = w - 1;
  if (t.dy > 0 && y + 1 >= h) ty = h - 1;
  const uint64_t idx = ty * w + tx;
  return (idx >= w * h || idx <= i) ? ~(uint64_t)0 : idx;
}

// jarvis (src/rd_dithers.cc:150-220), one frame per wavefront.
template <int FMT>
__device__ __forceinline__ void jarvis_frame(const ScreenFrame& f, const float4* __restrict__ ent,
                                             int count, float* ring, int64_t w, int64_t h,
                                             float diffusion) {
  const int lane = threadIdx.x;
  const int64_t n = w * h;
  for (int64_t k = lane; k < 9 * w; k += kWave) ring[k] = 0.f;
  __syncthreads();
  // Pixels in chunks of 64 in raster order, one per lane; the next chunk is
  // loaded while this one is dithered, the choices stored after it.
  float4 cur = raw_pixel<FMT>(lane < n ? load_raw<FMT>(f.src, lane, w, h) : make_uint4(0, 0, 0, 0));
  int chosen = 0;
  float4 chosen_c = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int64_t base = 0; base < n; base += kWave) {
    const int64_t nb = base + kWave + lane;
    const uint4 next = nb < n ? load_raw<FMT>(f.src, nb, w, h) : make_uint4(0, 0, 0, 0);
    const int m = (int)(n - base < kWave ? n - base : kWave);
    for (int j = 0; j < m; ++j) {
      const int64_t i = base + j;
      const int64_t y = i / w, x = i - y * w;
      if (x == 0 && y >= 1) {
        // Row y-1 is done: its slot becomes row y+2, which nothing has reached yet.
        const int64_t s = ((y + 2) % 3) * 3 * w;
        for (int64_t q = lane; q < 3 * w; q += kWave) ring[s + q] = 0.f;
        __syncthreads();
      }
      const int64_t slot = (y % 3) * 3 * w;
      const float er0 = ring[slot + x], eg0 = ring[slot + w + x], eb0 = ring[slot + 2 * w + x];
      const float px = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.x), j));
      const float py = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.y), j));
      const float pz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cur.z), j));
      const float r = px + er0 * diffusion;
      const float g = py + eg0 * diffusion;
      const float b = pz + eb0 * diffusion;
      // NearestEntry: the lanes split the entries; the first of equal minima
      // wins.  A NaN channel makes every distance NaN: the reference keeps
      // entry 0.
      int bi = 0;
      if (!(r != r || g != g || b != b)) {
        // Each lane scans its entries in order, taking its first one
        // unconditionally and then only strictly smaller distances; lanes
        // without entries offer (+inf, INT_MAX).
        float bd = __builtin_inff();
        int li = 0x7FFFFFFF;
        for (int e = lane; e < count; e += kWave) {
          const float4 c = ent[e];
          const float dr = r - c.x, dg = g - c.y, db = b - c.z;
          const float d = dr * dr + dg * dg + db * db;
          if (e == lane || d < bd) {
            bd = d;
            li = e;
          }
        }
        bi = wave_argmin(bd, li);
      }
      const float4 c = ent[bi];
      const float resr = r - c.x, resg = g - c.y, resb = b - c.z;
      if (lane == j) {
        chosen = bi;
        chosen_c = c;
      }
      // The taps: distinct targets in the interior, one lane each; near the
      // borders and on the last row clamped taps can coincide, and those are
      // added one after another in tap order.
      const bool interior = x >= 2 && x + 3 <= w && y + 2 <= h;
      if (interior) {
        if (lane < 12) {
          const uint64_t t = jarvis_target(lane, x, y, w, h, i);
          if (t != ~(uint64_t)0) {
            const int64_t tyy = (int64_t)(t / w), txx = (int64_t)(t - (uint64_t)tyy * w);
            const int64_t s = (tyy % 3) * 3 * w + txx;
            const float wt = kJarvis[lane].w;
            ring[s] = ring[s] + wt * resr;
            ring[s + w] = ring[s + w] + wt * resg;
            ring[s + 2 * w] = ring[s + 2 * w] + wt * resb;
          }
        }
      } else if (lane == 0) {
        for (int k = 0; k < 12; ++k) {
          const uint64_t t = jarvis_target(k, x, y, w, h, i);
          if (t == ~(uint64_t)0) continue;
          const int64_t tyy = (int64_t)(t / w), txx = (int64_t)(t - (uint64_t)tyy * w);
          const int64_t s = (tyy % 3) * 3 * w + txx;
          const float wt = kJarvis[k].w;
          ring[s] = ring[s] + wt * resr;
          ring[s + w] = ring[s + w] + wt * resg;
          ring[s + 2 * w] = ring[s + 2 * w] + wt * resb;
        }
      }
      __syncthreads();
    }
    if (lane < m) store_choice(f, base + lane, n, chosen, chosen_c, cur.w);
    cur = raw_pixel<FMT>(next);
  }
}

__global__ __launch_bounds__(kWave) void k_diffuse_jarvis(const ScreenFrame* __restrict__ frames,
                                                          int64_t w, int64_t h, float diffusion) {
  // LDS: the entries (float4 each), then 3 rows x (r, g, b planes) of w floats.
  extern __shared__ float4 lds4[];
  const ScreenFrame f = frames[blockIdx.x];
  const int count = __builtin_amdgcn_readfirstlane(f.pal->count);
  const int lane = threadIdx.x;
  float4* ent = lds4;
  {
    const float4* g = (const float4*)((const unsigned char*)f.pal->image + f.pal->off_ent);
    for (int k = lane; k < count; k += kWave) ent[k] = g[k];
  }
  float* ring = (float*)(lds4 + count);
  switch (f.src_fmt) {
    case kRgb8: jarvis_frame<kRgb8>(f, ent, count, ring, w, h, diffusion); break;
    case kRgba8: jarvis_frame<kRgba8>(f, ent, count, ring, w, h, diffusion); break;
    case kRgb16: jarvis_frame<kRgb16>(f, ent, count, ring, w, h, diffusion); break;
    case kRgba16: jarvis_frame<kRgba16>(f, ent, count, ring, w, h, diffusion); break;
    case kRgbaF32: jarvis_frame<kRgbaF32>(f, ent, count, ring, w, h, diffusion); break;
    case kYuv444: jarvis_frame<kYuv444>(f, ent, count, ring, w, h, diffusion); break;
    default: jarvis_frame<kYuv420>(f, ent, count, ring, w, h, diffusion); break;
  }
}

}  // namespace

size_t diffuse_lds_bytes(int kind, int64_t w, int count) {
  const size_t ent = 16 * (size_t)count;
  if (kind == 2) return ent + sizeof(float) * 9 * (size_t)w;
  return ent + sizeof(float) * (kind == 1 ? 6 : 3) * (size_t)w;
}

hipError_t launch_diffuse(const ScreenFrame* frames, int nframes, int64_t w, int64_t h, int kind,
                          int max_count, float diffusion, hipStream_t s) {
  if (nframes <= 0 || w * h <= 0) return hipSuccess;
  const size_t lds = diffuse_lds_bytes(kind, w, max_count);
  auto go = [&](auto kernel) -> hipError_t {
    hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    kernel<<<dim3((unsigned)nframes), dim3(kWave), lds, s>>>(frames, w, h, diffusion);
    return hipGetLastError();
  };
  if (kind == 0) return go(k_diffuse_rows<0>);
  if (kind == 1) return go(k_diffuse_rows<1>);
  return go(k_diffuse_jarvis);
}

}  // namespace sa
