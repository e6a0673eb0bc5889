// rd_im.h -- the palette functions of the reference's include/rd_im.h that sit
// on the hot path's edges, served by the MI355X path.
//
// Same names, signatures and error convention (bool + *error) as the
// reference (include/rd_im.h:120-186):
//   * ImWritePalette / ImReadPalette: the .txt / .text palette file
//     (src/rd_im.cpp:994-1024, :1175-1330); the image form (an indexed PNG
//     through ImageMagick) is refused with an error;
//   * ImBuildPaletteFromPixels: the palette of a pixel buffer, as the video
//     path's montage uses it (src/rd_im.cpp:535-656) -- ColorTree::Build on
//     the device, then HarvestColormap's removal of exact duplicate entries.
//     (The reference's palette comes from ImageMagick's QuantizeImage, whose
//     colormap the reference asserts equal to its own ColorTree's,
//     src/rd_cli.cpp:2506-2530; this one is that ColorTree's.)
#pragma once

#include <cstddef>
#include <string>

#include "rd_octree.h"
#include "rd_riemersma.h"

namespace rd {

bool ImWritePalette(const Palette& palette, const std::string& path, std::string* error);

bool ImReadPalette(const std::string& path, int expect_colors, Palette* palette,
                   ColorTree* tree, std::string* error);

bool ImBuildPaletteFromPixels(const RgbaF* pixels, std::size_t width, std::size_t height,
                              int colors, Palette* palette, ColorTree* tree, std::string* error);

}  // namespace rd
