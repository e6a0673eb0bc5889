// rd_plugin.h -- the reference's dither plugin seam, restated.
//
// Same types and registry functions as include/rd_plugin.h:86-148 of the
// reference, so a host dither written against it registers here unchanged.
// As in the reference, the GPU engines are not plugins (its rd_plugin.h:
// 181-189): they are called through include/rd_riemersma.h.
#ifndef RD_PLUGIN_H_
#define RD_PLUGIN_H_

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "rd_riemersma.h"

namespace rd {

using PluginPalette = Palette;

struct DitherJob {
  const Palette* palette = nullptr;
  const CurveStepEntry* curve = nullptr;
  std::size_t curve_count = 0;
  std::size_t width = 0;
  std::size_t height = 0;
  int frames = 1;  // consecutive frames, each with an independent error queue
  double diffusion = 1.0;
  void* user = nullptr;  // reserved; always null
};

struct DitherResult {
  std::string error;  // empty on success
};

struct DitherAlgorithm {
  const char* name;         // what --dither takes
  const char* description;  // one line, shown by --list-dithers
  DitherResult (*run)(DitherJob& job, RgbaF* pixels);
};

// False when the name is empty, already registered, or run is null.
bool RegisterDither(const DitherAlgorithm& algo);
const DitherAlgorithm* FindDither(const char* name);
const std::vector<DitherAlgorithm>& RegisteredDithers();
// One line per algorithm: "  name  description".
std::string DitherListText();

}  // namespace rd

#endif  // RD_PLUGIN_H_
