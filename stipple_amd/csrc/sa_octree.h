// sa_octree.h -- device octree build (ColorTree::Build) interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sa_palette.h"

namespace sa {

// Node arrays of a built tree (index 0 is the root), device memory.
struct TreeDev {
  uint64_t* key;          // (level << 40) | channel prefixes
  int32_t* child;         // 16 per node, -1 = none
  int32_t* parent;        // root's parent is the root
  int32_t* id;            // index in the parent's child array
  int32_t* level;
  double* qe;             // quantize_error
  double* share;          // this node's part of the root's chain (approximate)
  double* total;          // total_color, 4 per node
  uint64_t* nu;           // number_unique
  uint32_t* color_number;
  int32_t* row;           // search-image row of internal nodes, else -1
};

struct PaletteBuild {
  unsigned char* image;   // search image (capacity image_cap)
  int32_t image_cap;
  PaletteDev* hdr;        // header, written by the build
  double* entries;        // colormap, 4 doubles per entry
};

struct OctreeOut {
  int keep_tree = 0;      // keep the node arrays (for the ColorTree drop-in)
  TreeDev tree{};         // valid when keep_tree; points into tree_block
  int64_t tree_nodes = 0;
  void* tree_block = nullptr;      // one allocation holding the node arrays
  void* block = nullptr;           // one allocation holding image, hdr and entries;
  unsigned char* image = nullptr;  // owned by the caller after a successful build
  PaletteDev* hdr = nullptr;
  double* entries = nullptr;
};

// One image of a build request.
struct OctreeReq {
  const void* src;        // device pixels in `fmt`
  int fmt;
  int64_t w, h;
  int maxc;               // palette size (<= 0: 65536)
  int assoc;
  int gray;
};

struct OctreeResult {
  int status = 0;         // SA_* code
  char msg[256];
  int count = 0;          // colours
  int rows = 0;
  int image_bytes = 0;
  int64_t nodes = 0;
  int depth = 0;
  int64_t live = 0;
  double root_qe = 0.0;
  int exact_root = 0;     // 1 when the root's chain was summed serially
};

int octree_depth(int maxc, bool assoc, bool gray);
OctreeResult build_octree(const void* src, int fmt, int64_t w, int64_t h, int maxc, int assoc,
                          int gray, hipStream_t s, OctreeOut* out);

// Builds n palettes together, one launch per kernel for all of them (images
// are grouped into sub-batches of at most 2^26 pixels).  Each image gets its
// own result and outputs, exactly as build_octree would give it; the stream is
// synchronised four times per sub-batch.
void build_octree_batch(const OctreeReq* req, int n, hipStream_t s, OctreeOut* out,
                        OctreeResult* res);

// 1 (the default): loop 2's pairs of an image whose cube depth leaves few cells
// are placed by a counting sort over cells; 0: every pair goes through the
// node hash and the interval sort.  Both give the same trees.
void octree_set_cells(int on);
int octree_cells();
// 1: the root's quantize_error is always summed serially (exact); 0 (the
// default): only when the rounding bound of the parallel sum cannot tell it
// from the other nodes' errors, which is when its exact value could matter.
void octree_set_exact_root(int on);
int octree_exact_root();

}  // namespace sa
