// rd_riemersma.h -- the reference's GPU engine API, served by the MI355X path.
//
// Signatures, argument meaning and error behaviour follow the reference's
// include/rd_riemersma.h (line numbers below are that file's):
//   * engines return std::string, empty on success and never empty on
//     failure (:214-217; src/rd_cuda_stub.cpp:60-74);
//   * pixels are caller-owned host RgbaF Q16, dithered in place: one frame
//     for RiemersmaWalkCuda and RiemersmaApproxCuda, `frames` consecutive
//     frames of one geometry for RiemersmaBlocksCuda (:204-217, :224-268);
//   * calls are synchronous on return (src/rd_riemersma_cuda.cu:742-746);
//   * refusals: block < 16, frames < 1, palette > 256 for blocks and approx,
//     taps outside 1..4096 (:296-307, :450-464).
// The kernels behind them are this repository's HIP code, reached through
// the C-ABI of include/stipple_amd.h; nothing here computes a result on the
// host.  The reference's CPU engines (RiemersmaWalkCpu, RiemersmaBlocksCpu)
// are its own code and are not restated here.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "rd_octree.h"
#include "rd_types.h"

namespace rd {

enum class Gravity : int { kWest = 0, kEast = 1, kNorth = 2, kSouth = 3 };

// :63-67.  The output palette (ImageMagick's colormap).
struct Palette {
  PaletteEntry entries[kMaxColormapSize] = {};
  int count = 0;
  bool associate_alpha = false;
};

// :69-78.
struct DitherParams {
  int colors = 16;
  double diffusion = 1.0;  // dither:diffusion-amount
  bool use_cache = true;   // quantize.c's memo; true reproduces ImageMagick
};

// :125-130.
struct CurveStepEntry {
  int x = 0;
  int y = 0;
  int dir = 0;
  int dir_next = -1;
};

// :88-94, the curve level ImageMagick derives from the longer side.
inline int ComputeCurveLevel(std::size_t width, std::size_t height) {
  const std::size_t extent = width > height ? width : height;
  if (extent == 0) return 0;
  int level = static_cast<int>(std::log2(static_cast<double>(extent)));
  if ((static_cast<std::size_t>(1) << level) < extent) ++level;
  return level;
}

// Curve point n of the level-`level` Hilbert curve in ImageMagick's
// orientation (src/rd_riemersma_cpu.cpp:167-189 of the reference).
void HilbertPoint(std::int64_t n, int level, int* out_x, int* out_y);

// :178-179.  The in-bounds pixel ids of the block engine's curve (generated
// on the device): n in [0, 4^level - 1) plus a trailing origin, and the
// last-writer owner of every pixel (-1 for the one unvisited pixel of a
// power-of-two-wide image).  Returns the curve length; a 1x1 image (level 0)
// has an empty block curve, as in the reference.
std::size_t BuildCurveIndex(int level, std::size_t width, std::size_t height,
                            std::vector<int>* out, std::vector<int>* owner);

// :278-279.  Cached per thread, as the reference's thread_local cache.
const std::vector<int>* BlockCurveOrder(std::size_t width, std::size_t height);
const std::vector<int>* BlockCurveOwner(std::size_t width, std::size_t height);

enum class Engine { kCpu, kCuda, kApprox, kBlocks, kOpenCL };

// :214-217.  The exact sequential walk, bit-exact with RiemersmaWalkCpu.
// `pixels` holds one width*height frame, dithered in place.  As in the
// reference (src/rd_riemersma_cuda.cu:688, :747), `frames` copies of that one
// frame would be walked and one frame is returned; the copies are identical,
// so one walk is done.  Distinct images go through RiemersmaWalkBatch.
std::string RiemersmaWalkCuda(const Palette& palette, const DitherParams& params,
                              const ColorTree& tree, std::size_t width, std::size_t height,
                              RgbaF* pixels, int frames, std::string* device_name);

// New: the batched form of the same walk over uneven (image, tree) items,
// walked concurrently from one device queue, largest first.  Each item is
// bit-exact with RiemersmaWalkCpu on its own.
struct WalkItem {
  RgbaF* pixels = nullptr;  // width*height, dithered in place
  std::size_t width = 0;
  std::size_t height = 0;
  const ColorTree* tree = nullptr;  // its colormap is the item's palette
};
std::string RiemersmaWalkBatch(const std::vector<WalkItem>& items, const DitherParams& params,
                               std::string* device_name);

// :284.
bool CudaAvailable();
// :287.  Frames of this geometry whose device state fits `vram_budget` bytes
// (0: the free device memory).
int CudaMaxFrames(std::size_t width, std::size_t height, std::size_t vram_budget);

// :296-301.
struct ApproxOptions {
  int iterations = 20;
  int taps = 96;
  bool fp64 = false;
  bool use_tree = true;
};

// :303-307.  Jacobi sweeps along the block curve, then the owner-filtered
// scatter, in place.  `pixels` holds one image (the reference's signature).
std::string RiemersmaApproxCuda(const Palette& palette, const DitherParams& params,
                                const ColorTree& tree, std::size_t width, std::size_t height,
                                RgbaF* pixels, const ApproxOptions& options,
                                std::string* device_name);

// :309.  g[0..taps] of (I - H)^-1 (src/rd_approx_cuda.cu:73-88).
std::vector<double> ApproxInverseFilter(int taps, double diffusion);

// :325-443.
struct BlockOptions {
  int block = 512;
  int frames = 1;
  bool emit_u16 = false;     // write packed rgba64le to out_u16 instead of batch
  bool emit_yuv444 = false;  // with emit_u16: planar 8-bit 4:4:4 to out_u16 (3 bytes/px)
  bool upload_u16 = false;   // read interleaved uint16 from in_u16 instead of batch
  int in_channels = 4;       // 3 (rgb48le) or 4 (rgba64le) when upload_u16
  enum class InMode {
    Interleaved16 = 0,
    PlanarYuv444 = 1,
    PlanarYuv444Prepass = 2,
    PlanarYuv420 = 3,
  };
  // Planar modes read 8-bit planes through in_u16: 4:4:4 (both 4:4:4 modes
  // give the same values) or 4:2:0 with even width and height.
  InMode in_mode = InMode::Interleaved16;
  bool in_u16_pinned = false;
  bool out_u16_pinned = false;
  bool batch_pinned = false;
};

// :450-464.  The block-parallel walk of options.frames frames of one
// geometry, bit-exact with RiemersmaBlocksCpu at the same block length.
// `state_slot` selects one of kGpuStateSlots persistent device states; two
// threads using the same slot serialise.
std::string RiemersmaBlocksCuda(const Palette& palette, const DitherParams& params,
                                const ColorTree& tree, std::size_t width, std::size_t height,
                                RgbaF* batch, const BlockOptions& options,
                                std::string* device_name, std::uint16_t* out_u16 = nullptr,
                                int state_slot = 0, const std::uint16_t* in_u16 = nullptr);

inline constexpr int kGpuStateSlots = 2;

// :480-481.  Page-locked host memory (hipHostMalloc), for batch / in_u16 /
// out_u16 buffers whose copies should be true DMA.
void* CudaAllocPinned(std::size_t bytes);
void CudaFreePinned(void* p);

}  // namespace rd
