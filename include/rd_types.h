// rd_types.h -- the data types of the reference engine API, restated.
//
// The C++ drop-in layer keeps the reference's type names and layouts so a
// caller written against include/rd_types.h of the reference compiles
// unchanged: RgbaF is ImageMagick's Quantum == float pixel (4 x float, 16 B),
// PaletteEntry four doubles (include/rd_types.h:51-58, :156-158 of the
// reference).  Only what the GPU engine path uses is declared here; the
// arithmetic helpers live in the device code (stipple_amd/csrc/sa_math.h).
#pragma once

#include <cstddef>
#include <cstdint>

namespace rd {

inline constexpr int kQuantumDepth = 16;
inline constexpr double kQuantumRange = 65535.0;        // QuantumRange
inline constexpr double kQuantumScale = 1.0 / 65535.0;  // QuantumScale
inline constexpr double kOpaqueAlpha = 65535.0;         // OpaqueAlpha
inline constexpr double kMagickEpsilon = 1.0e-12;       // MagickEpsilon

inline constexpr int kErrorQueueLength = 16;             // ErrorQueueLength
inline constexpr double kErrorRelativeWeight = 1.0 / 16.0;
inline constexpr int kMaxTreeDepth = 8;                  // MaxTreeDepth
inline constexpr int kMaxColormapSize = 65536;           // MaxColormapSize
inline constexpr int kCacheShift = 2;                    // quantize.c CacheShift
inline constexpr int kCacheEntries = 1 << (4 * (8 - kCacheShift));

struct RgbaD {
  double r = 0.0, g = 0.0, b = 0.0, a = 0.0;
};

// One pixel as ImageMagick stores it: Quantum == float per channel.
struct RgbaF {
  float r = 0.0f, g = 0.0f, b = 0.0f, a = 0.0f;
};

struct PaletteEntry {
  double r = 0.0, g = 0.0, b = 0.0, a = kOpaqueAlpha;
};

static_assert(sizeof(RgbaF) == 16, "RgbaF is 4 floats");
static_assert(sizeof(PaletteEntry) == 32, "PaletteEntry is 4 doubles");

}  // namespace rd
