// sa_math.h -- the Q16-HDRI numeric contract, usable from host and device.
//
// Every function here reproduces one of the reference's scalar helpers bit for
// bit (include/rd_types.h).  All translation units that include it are built
// with -ffp-contract=off, so `a*b+c` below is two roundings, never an FMA.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SA_HD __host__ __device__ __forceinline__
#else
// Plain host build (the CPU tests compile this header with g++).
#define SA_HD inline
#endif

namespace sa {

constexpr double kRange = 65535.0;        // QuantumRange
constexpr double kScale = 1.0 / 65535.0;  // QuantumScale
constexpr double kEps = 1.0e-12;          // MagickEpsilon
constexpr int kQueue = 16;                // ErrorQueueLength
constexpr int kMaxDepth = 8;              // MaxTreeDepth
// 4*(QuantumRange+1)^2 + 1, the incumbent distance a search starts from
// (src/rd_riemersma_cpu.cpp:90).
constexpr double kFarDistance = 4.0 * ((kRange + 1.0) * (kRange + 1.0)) + 1.0;
// fl64(1/257).  q * kInv257 rounded to float equals the correctly rounded
// float quotient q/257 for every float q in [0, 65535]: the exact quotient of
// a float by 257 lies at least 2^-32 (relative) from any float rounding
// midpoint, and the double product is within 2^-52 of it.  Checked
// exhaustively on the host (tests/test_numeric.py) and on the device
// (tests/test_gpu_numeric.py).
constexpr double kInv257 = 1.0 / 257.0;

// include/rd_types.h:39-43 ClampPixel (narrows through float).
SA_HD float clamp_pixel(double v) {
  // Selects rather than branches: the same three cases, no divergent flow.
  float f = (float)v;
  f = (v >= kRange) ? 65535.0f : f;
  f = (v < 0.0) ? 0.0f : f;
  return f;
}

// include/rd_types.h:133-138 ScaleQuantumToChar, with the float division.
SA_HD unsigned scale_char_div(float q) {
  if (!(q > 0.0f)) return 0;  // NaN or <= 0
  const float s = q / 257.0f;
  if (s >= 255.0f) return 255;
  return (unsigned)(s + 0.5f);
}

// Same result as scale_char_div for q = clamp_pixel(v) of any finite or NaN
// v (so q is NaN or in [0, 65535]): the reciprocal product replaces the
// division (see kInv257), q <= 0 yields 0 through the +0.5 truncation, and
// q/257 < 255 for every float q < 65535.
SA_HD unsigned scale_char(float q) {
  if (!(q > 0.0f)) return 0;
  const float s = (float)((double)q * kInv257);
  return (unsigned)(s + 0.5f);
}

// include/rd_types.h:30-33.
SA_HD double safe_recip(double x) {
  const double sign = x < 0.0 ? -1.0 : 1.0;
  return (sign * x) >= kEps ? 1.0 / x : sign * (1.0 / kEps);
}

// include/rd_types.h:179-192 AssociateAlphaPixel: premultiply unless opaque.
SA_HD void assoc_pixel(bool assoc, float r, float g, float b, float a, double out[4]) {
  out[3] = (double)a;
  if (!assoc || a == 65535.0f) {
    out[0] = (double)r;
    out[1] = (double)g;
    out[2] = (double)b;
    return;
  }
  const double al = kScale * (double)a;
  out[0] = al * (double)r;
  out[1] = al * (double)g;
  out[2] = al * (double)b;
}

// include/rd_types.h:195-208 AssociateAlphaPixelInfo (palette entry).
SA_HD void assoc_entry(bool assoc, const double e[4], double out[4]) {
  out[3] = e[3];
  if (!assoc || e[3] == kRange) {
    out[0] = e[0];
    out[1] = e[1];
    out[2] = e[2];
    return;
  }
  const double al = kScale * e[3];
  out[0] = al * e[0];
  out[1] = al * e[1];
  out[2] = al * e[2];
}

// include/rd_types.h:141-153 CacheOffset from already scaled channel bytes.
SA_HD int memo_key(bool assoc, unsigned c0, unsigned c1, unsigned c2, unsigned c3) {
  int k = (int)(c0 >> 2) | ((int)(c1 >> 2) << 6) | ((int)(c2 >> 2) << 12);
  if (assoc) k |= (int)(c3 >> 2) << 18;
  return k;
}

// Where the walk stores memo entry `memo_key` (a bijection of it, so the
// memo's first-touch semantics are untouched): entries are grouped in lines of
// 64 that cover a 4x4x4 cube of the 6-bit r,g,b codes at one alpha code --
// offset = the low two bits of r, g, b; line = the high four bits of r, g, b
// and the alpha code.  Targets along the walk stay in a few neighbouring
// cubes, where a line of one channel's 64 codes would be touched once or twice.
SA_HD uint32_t memo_store_key(bool assoc, unsigned c0, unsigned c1, unsigned c2, unsigned c3) {
  const uint32_t r = c0 >> 2, g = c1 >> 2, b = c2 >> 2;
  uint32_t line = (r >> 2) | ((g >> 2) << 4) | ((b >> 2) << 8);
  if (assoc) line |= (uint32_t)(c3 >> 2) << 12;
  return (line << 6) | (r & 3u) | ((g & 3u) << 2) | ((b & 3u) << 4);
}

// scale_char for q = clamp_pixel(v), where q is NaN or in [0, 65535]: the
// guard becomes a max (s + 0.5 >= 0.5 for every q >= 0, and a NaN sum gives
// 0, as the guard does), so the result is the same with no compare and select.
SA_HD unsigned scale_char_clamped(float q) {
  const float s = (float)((double)q * kInv257);
  return (unsigned)fmaxf(s + 0.5f, 0.0f);
}

// memo_store_key split by channel: memo_store_key(assoc, c0, c1, c2, c3) ==
// memo_key_part(0, c0) | memo_key_part(1, c1) | memo_key_part(2, c2) |
// (assoc ? memo_key_part(3, c3) : 0).
SA_HD uint32_t memo_key_part(int ch, unsigned c8) {
  if (ch == 3) return (uint32_t)(c8 >> 2) << 18;
  return ((uint32_t)(c8 >> 4) << (6 + 4 * ch)) | (((c8 >> 2) & 3u) << (2 * ch));
}

// src/rd_octree.cpp:59-67 ColorToQNodeId on packed channel bytes.
SA_HD int node_id(bool assoc, unsigned c0, unsigned c1, unsigned c2, unsigned c3, int index) {
  int id = (int)((c0 >> index) & 1u);
  id |= (int)((c1 >> index) & 1u) << 1;
  id |= (int)((c2 >> index) & 1u) << 2;
  if (assoc) id |= (int)((c3 >> index) & 1u) << 3;
  return id;
}

// One candidate's distance with quantize.c's channel order.  Equivalent to the
// early-exit chain of src/rd_octree.cpp:190-215 in its effect on the running
// minimum: partial sums of non-negative terms are monotone under round to
// nearest, so a candidate passes every partial test exactly when its full
// distance passes the last one.
SA_HD double candidate_distance(bool assoc, const float e[4], const double t[4]) {
  double alpha = 1.0, beta = 1.0;
  if (assoc) {
    alpha = kScale * (double)e[3];
    beta = kScale * t[3];
  }
  double v = alpha * (double)e[0] - beta * t[0];
  double d = v * v;
  v = alpha * (double)e[1] - beta * t[1];
  d += v * v;
  v = alpha * (double)e[2] - beta * t[2];
  d += v * v;
  if (assoc) {
    v = (double)e[3] - t[3];
    d += v * v;
  }
  return d;
}

}  // namespace sa
