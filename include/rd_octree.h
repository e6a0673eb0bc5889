// rd_octree.h -- ColorTree with the reference's interface, built on the GPU.
//
// Same names and meaning as include/rd_octree.h:35-123 of the reference:
// Build() is quantize.c's classify -> reduce -> colormap, and afterwards
// color_count(), colormap() and nodes() hold exactly what the reference's
// ColorTree::Build leaves (same node numbering, the same doubles in every
// QNode).  The difference is where it runs: Build() uploads the pixels and
// runs the device octree (sa_palette_build); the node arrays are read back so
// the accessors work on the host, and the device search structure stays
// resident for the engines (device_palette()).
//
// ClosestColor / PackChannels / NodeId, which only the reference's CPU walk
// calls, are not part of the GPU path and are not declared.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "rd_types.h"

typedef struct sa_palette_s* sa_palette;

namespace rd {

inline constexpr int kMaxQNodes = 266817;  // quantize.c: MaxQNodes
inline constexpr int kRgbChildren = 8;
inline constexpr int kRgbaChildren = 16;

struct QNode {
  std::int32_t child[kRgbaChildren] = {};  // -1 == null
  std::int32_t parent = -1;
  double total_color[4] = {0.0, 0.0, 0.0, 0.0};
  double quantize_error = 0.0;
  std::uint64_t number_unique = 0;
  std::uint32_t color_number = 0;
  std::uint32_t id = 0;
  std::uint32_t level = 0;
};

class ColorTree {
 public:
  ColorTree();
  ~ColorTree();
  ColorTree(const ColorTree&) = delete;
  ColorTree& operator=(const ColorTree&) = delete;
  ColorTree(ColorTree&&) noexcept;
  ColorTree& operator=(ColorTree&&) noexcept;

  // rd_octree.cpp:219-409, on the device.  `pixels` is host memory (w*h RgbaF).
  // The reference has no failure path; this one can fail (no device, device
  // memory): color_count() is then 0 and last_error() says why.
  void Build(const RgbaF* pixels, std::size_t width, std::size_t height, int max_colors,
             bool associate_alpha, bool grayscale);

  int color_count() const { return color_count_; }
  const PaletteEntry* colormap() const { return colormap_.data(); }
  const std::vector<QNode>& nodes() const { return nodes_; }
  bool associate_alpha() const { return associate_alpha_; }
  std::size_t node_count() const { return nodes_.size(); }
  std::int32_t root() const { return nodes_.empty() ? -1 : 0; }
  std::int32_t Child(std::int32_t node, std::size_t id) const { return nodes_[node].child[id]; }
  std::int32_t Parent(std::int32_t node) const { return nodes_[node].parent; }
  int depth() const { return depth_; }

  // The device search structure of this tree (owned by the tree).
  sa_palette device_palette() const { return device_; }
  const std::string& last_error() const { return error_; }

 private:
  void Reset();
  std::vector<QNode> nodes_;
  std::vector<PaletteEntry> colormap_;
  int color_count_ = 0;
  int depth_ = 3;
  bool associate_alpha_ = false;
  sa_palette device_ = nullptr;
  std::string error_;
};

}  // namespace rd
