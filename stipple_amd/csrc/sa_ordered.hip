// sa_ordered.hip -- the threshold-screen dithers of the plugin seam, one
// thread per pixel: bayer-ordered and void-and-cluster (OrderedPick against a
// screen threshold, src/rd_dithers_ordered.cc:79-204, :447-486) and
// clustered-dot (the palette's darkest or lightest entry by luma against an
// 8x8 cell, src/rd_dithers.cc:294-358).
//
// Every pixel is independent, so the kernel is a stream: read the pixel,
// pick, write the entry.  Each thread takes 4 pixels, so each palette entry
// (the search image's `ent`, the reference's static_cast<float>(entry); a
// wave-uniform load) serves four distance updates; clustered-dot's two
// candidates come from one wavefront per frame over the entries' doubles,
// before the stream.  Arithmetic is the reference's, one
// rounding per operation in its order (no contraction, correctly rounded
// sqrt and division), so results are bit-exact.
#include "sa_internal.h"
#include "sa_math.h"

namespace sa {
namespace {

constexpr int kThreads = 256;
constexpr int kPer = 4;  // consecutive pixels per thread (RGB8: three dwords in, three out)

// OrderedPick (src/rd_dithers_ordered.cc:79-145) for kPer pixels at once: the
// nearest and second nearest entries by float distance (first wins ties),
// then the second when the threshold falls below the nearer's share of the
// two distances.  Entries are wave-uniform loads, shared by the kPer pixels.
// The reference's "no candidate yet" tests (d < 0) can only hold at entries 0
// and 1, since distances are sums of squares, so those two are peeled.
__device__ __forceinline__ void ordered_pick4(const float4* __restrict__ ent, int count,
                                             const float4 (&p)[kPer], const float (&t)[kPer],
                                             int (&idx)[kPer]) {
  float d1[kPer], d2[kPer];
  int first[kPer], second[kPer];
  const float4 e0 = ent[0];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const float dr = p[k].x - e0.x, dg = p[k].y - e0.y, db = p[k].z - e0.z;
    d1[k] = dr * dr + dg * dg + db * db;
    first[k] = 0;
    second[k] = 0;
    d2[k] = -1.0f;
  }
  if (count > 1) {
    const float4 e1 = ent[1];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const float dr = p[k].x - e1.x, dg = p[k].y - e1.y, db = p[k].z - e1.z;
      const float d = dr * dr + dg * dg + db * db;
      if (d < d1[k]) {
        d2[k] = d1[k];
        second[k] = first[k];
        d1[k] = d;
        first[k] = 1;
      } else {
        d2[k] = d;
        second[k] = 1;
      }
    }
  }
  for (int i = 2; i < count; ++i) {
    const float4 e = ent[i];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const float dr = p[k].x - e.x, dg = p[k].y - e.y, db = p[k].z - e.z;
      const float d = dr * dr + dg * dg + db * db;
      if (d < d1[k]) {
        d2[k] = d1[k];
        second[k] = first[k];
        d1[k] = d;
        first[k] = i;
      } else if (d < d2[k]) {
        d2[k] = d;
        second[k] = i;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    if (d2[k] < 0.0f) {
      idx[k] = first[k];
    } else if (d2[k] <= 0.0f) {
      idx[k] = second[k];
    } else {
      const float s1 = sqrtf(d1[k]), s2 = sqrtf(d2[k]);
      const float frac = s1 / (s1 + s2);
      idx[k] = t[k] < frac ? second[k] : first[k];
    }
  }
}

__device__ __forceinline__ double luma(double r, double g, double b) {
  return 0.299 * r + 0.587 * g + 0.114 * b;
}

// clustered-dot's two candidates per frame, the palette's darkest and
// lightest entries by luma, first wins (src/rd_dithers.cc:317-328): one
// wavefront per frame, lexicographic on (luma, index) through the shuffles.
__global__ __launch_bounds__(64) void k_luma_range(const ScreenFrame* __restrict__ frames,
                                                   ScreenRange* __restrict__ out) {
  const ScreenFrame f = frames[blockIdx.x];
  const int count = f.pal->count;
  double lv = 0.0, hv = 0.0;
  int li = -1, hi = -1;
  for (int i = threadIdx.x; i < count; i += 64) {
    const double* c = f.entries + 4 * (int64_t)i;
    const double l = luma(c[0], c[1], c[2]);
    if (li < 0 || l < lv) { lv = l; li = i; }
    if (hi < 0 || l > hv) { hv = l; hi = i; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double olv = __shfl_xor(lv, o), ohv = __shfl_xor(hv, o);
    const int oli = __shfl_xor(li, o), ohi = __shfl_xor(hi, o);
    if (oli >= 0 && (li < 0 || olv < lv || (olv == lv && oli < li))) { lv = olv; li = oli; }
    if (ohi >= 0 && (hi < 0 || ohv > hv || (ohv == hv && ohi < hi))) { hv = ohv; hi = ohi; }
  }
  if (threadIdx.x == 0) out[blockIdx.x] = ScreenRange{lv, hv > lv ? hv - lv : 1.0, li, hi};
}

// Workgroup w serves kThreads * P pixels of frame w % nframes (chunk
// w / nframes).  screen: 0 Bayer 8x8, 1/2 void-and-cluster 64/32 (ranks),
// 3 clustered-dot cell (CLUSTER).
template <bool CLUSTER, int P = CLUSTER ? 16 : kPer>
__global__ __launch_bounds__(kThreads) void k_screen(const ScreenFrame* __restrict__ frames,
                                                     int nframes, int64_t w, int64_t h,
                                                     const uint16_t* __restrict__ table,
                                                     int side, int screen, double diffusion,
                                                     const ScreenRange* __restrict__ ranges) {
  const int fi = (int)(blockIdx.x % (unsigned)nframes);
  const int64_t chunk = blockIdx.x / (unsigned)nframes;
  const ScreenFrame f = frames[fi];
  const PaletteDev pd = *f.pal;
  const int count = pd.count;
  const float4* ent = (const float4*)((const unsigned char*)pd.image + pd.off_ent);
  const int64_t pp = w * h;
  const int64_t base = (chunk * kThreads + threadIdx.x) * P;
  const bool whole = base + P <= pp;
  int64_t pix[P];
  float4 p[P];
  float t[P];
  int idx[P];
  if (whole && f.src_fmt == kRgb8 && ((uintptr_t)f.src & 3) == 0) {
    // RGB8 pixels four at a time from 12 aligned bytes (base is a multiple of 4).
    const uint32_t* q = (const uint32_t*)((const uint8_t*)f.src + 3 * base);
#pragma unroll
    for (int g = 0; g < P / 4; ++g) {
      const uint32_t a = q[3 * g], b = q[3 * g + 1], c = q[3 * g + 2];
      const uint32_t by[12] = {a & 255u, (a >> 8) & 255u, (a >> 16) & 255u, a >> 24,
                               b & 255u, (b >> 8) & 255u, (b >> 16) & 255u, b >> 24,
                               c & 255u, (c >> 8) & 255u, (c >> 16) & 255u, c >> 24};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        p[4 * g + k] = make_float4((float)by[3 * k] * 257.0f, (float)by[3 * k + 1] * 257.0f,
                                   (float)by[3 * k + 2] * 257.0f, 65535.0f);
    }
  } else {
#pragma unroll
    for (int k = 0; k < P; ++k)
      p[k] = base + k < pp ? load_any(f.src_fmt, f.src, base + k, w, h)
                           : make_float4(0.f, 0.f, 0.f, 0.f);
  }
  int lo_i = 0, hi_i = 0;
  double lo = 0.0, span = 1.0;
  if (CLUSTER) {
    const ScreenRange r = ranges[fi];
    lo = r.lo;
    span = r.span;
    lo_i = r.lo_i;
    hi_i = r.hi_i;
  }
  // Coordinates in 32 bits when the frame allows (screen sides are powers of two).
  const bool narrow = pp < ((int64_t)1 << 31);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    pix[k] = base + k;
    int64_t y, x;
    if (narrow) {
      const uint32_t yy = (uint32_t)pix[k] / (uint32_t)w;
      y = yy;
      x = (int64_t)((uint32_t)pix[k] - yy * (uint32_t)w);
    } else {
      y = pix[k] / w;
      x = pix[k] - y * w;
    }
    if (CLUSTER) {
      const double tt = (luma(p[k].x, p[k].y, p[k].z) - lo) / span;
      const double cell = (double)table[(y & 7) * 8 + (x & 7)];
      const double level = 1.0 - tt;
      double dark = cell / 32.0;
      if (cell == 0) dark = 0.0;
      dark = 0.5 + (dark - 0.5) * diffusion;
      idx[k] = level >= dark ? lo_i : hi_i;
    } else {
      float tk;
      if (screen == 0) {
        tk = ((float)table[(y & 7) * 8 + (x & 7)] + 0.5f) / 64.0f;
      } else {
        const uint16_t rk = table[(y & (side - 1)) * side + (x & (side - 1))];
        tk = (float)rk / (float)((double)side * (double)side);
      }
      t[k] = 0.5f + (tk - 0.5f) * (float)diffusion;
    }
  }
  if constexpr (!CLUSTER) ordered_pick4(ent, count, p, t, idx);
  if (whole && f.dst_fmt == kRgb8 && ((uintptr_t)f.dst & 3) == 0) {
    uint32_t* d = (uint32_t*)((uint8_t*)f.dst + 3 * base);
#pragma unroll
    for (int g = 0; g < P / 4; ++g) {
      uint32_t by[12];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float4 e = ent[idx[4 * g + k]];
        by[3 * k] = scale_char(e.x);
        by[3 * k + 1] = scale_char(e.y);
        by[3 * k + 2] = scale_char(e.z);
      }
      d[3 * g] = by[0] | (by[1] << 8) | (by[2] << 16) | (by[3] << 24);
      d[3 * g + 1] = by[4] | (by[5] << 8) | (by[6] << 16) | (by[7] << 24);
      d[3 * g + 2] = by[8] | (by[9] << 8) | (by[10] << 16) | (by[11] << 24);
    }
  } else {
#pragma unroll
    for (int k = 0; k < P; ++k)
      if (pix[k] < pp) store_pick(f, pix[k], pp, idx[k], ent[idx[k]], p[k].w);
  }
}

}  // namespace

hipError_t launch_screen(const ScreenFrame* frames, int nframes, int64_t w, int64_t h,
                         const uint16_t* table, int side, int screen, double diffusion,
                         ScreenRange* ranges, hipStream_t s) {
  if (nframes <= 0 || w * h <= 0) return hipSuccess;
  const int64_t per = screen == 3 ? 16 : kPer;  // pixels per thread of the kernel launched
  const int64_t chunks = (w * h + kThreads * per - 1) / (kThreads * per);
  const dim3 grid((unsigned)(chunks * nframes));
  if (screen == 3) {
    k_luma_range<<<dim3((unsigned)nframes), dim3(64), 0, s>>>(frames, ranges);
    k_screen<true><<<grid, dim3(kThreads), 0, s>>>(frames, nframes, w, h, table, side, screen,
                                                   diffusion, ranges);
  } else {
    k_screen<false><<<grid, dim3(kThreads), 0, s>>>(frames, nframes, w, h, table, side, screen,
                                                    diffusion, ranges);
  }
  return hipGetLastError();
}

}  // namespace sa
