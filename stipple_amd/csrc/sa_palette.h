// sa_palette.h -- the device-resident palette + octree search structure.
//
// One contiguous "search image" per palette, laid out so an engine can copy it
// into LDS verbatim:
//
//   double4  col[count]    associated colour of each entry, r,g,b,a (the value
//                          the residual subtracts: AssociateAlphaPixelInfo)
//   float4   ent[count]    the entry itself (Quantum == float, exact)
//   int2     sub[rows]     candidate range [off, off+len) of each internal node
//   int16    child[rows*kids]  -1 none, kLeaf a child that has no children,
//                          else the child's row
//   uint16   cand[count]   palette indices of the colour-bearing nodes in
//                          post-order
//
// Why this is the reference's search, exactly: quantize.c (restated in
// src/rd_octree.cpp:178-217) scans the subtree of parent(deepest node) in
// post-order and keeps the LAST candidate whose distance is <= the running
// minimum.  A subtree's post-order is a contiguous range of the root's
// post-order, so every candidate list is a range of `cand`.  The descent
// (src/rd_riemersma_cpu.cpp:82-89) only needs to know, per child, whether it
// exists and whether it has children of its own; the node the descent stops
// below ("prev") then names the range.  Only internal nodes get a row.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SA_PAL_HD __host__ __device__
#else
#define SA_PAL_HD
#endif

namespace sa {

constexpr int16_t kNoChild = -1;
constexpr int16_t kLeaf = -2;

struct PaletteDev {   // header, in device memory
  int32_t count;      // palette entries (<= 65536)
  int32_t assoc;      // alpha associated (16-ary tree, 4-channel distance)
  int32_t kids;       // 8 or 16
  int32_t rows;       // internal nodes (row 0 is the root)
  int32_t bytes;      // size of the search image
  int32_t off_ent;    // byte offsets inside the search image
  int32_t off_cent;
  int32_t off_sub;
  int32_t off_desc;
  int32_t off_child;
  int32_t off_cand;
  int32_t pad_;
  const void* image;  // device pointer to the search image
};

// Byte layout of a search image with the given counts (16-byte aligned parts):
//   double4 col[count]   associated colour (the residual's subtrahend)
//   float4  ent[count]   the entry (Quantum == float)
//   float4  cent[count]  ent of cand[k]: the candidates in post-order
//   int2    sub[rows]    candidate range of each internal node
//   uint32  desc[8^3 or 16^3]  the descent's first three levels (see below)
//   int16   child[rows*kids]
//   uint16  cand[count]
struct PaletteLayout {
  int32_t off_col, off_ent, off_cent, off_sub, off_desc, off_child, off_cand, bytes;
};

SA_PAL_HD inline int desc_entries(int kids) { return kids == 16 ? 4096 : 512; }

SA_PAL_HD inline PaletteLayout palette_layout(int count, int rows, int kids) {
  PaletteLayout L;
  L.off_col = 0;
  L.off_ent = (L.off_col + 32 * count + 15) & ~15;
  L.off_cent = (L.off_ent + 16 * count + 15) & ~15;
  L.off_sub = (L.off_cent + 16 * count + 15) & ~15;
  L.off_desc = (L.off_sub + 8 * rows + 15) & ~15;
  L.off_child = (L.off_desc + 4 * desc_entries(kids) + 15) & ~15;
  L.off_cand = (L.off_child + 2 * rows * kids + 15) & ~15;
  L.bytes = (L.off_cand + 2 * count + 15) & ~15;
  return L;
}

// The descent of src/rd_riemersma_cpu.cpp:82-89 over its first three levels
// (channel bits 7, 6, 5), tabulated per combination of those bits: entry
// index = top3(c0) | top3(c1) << 3 | top3(c2) << 6 [| top3(c3) << 9].
// Value: bit 31 set when the descent stopped within the three levels, bits
// 16..30 the row whose candidate range applies so far ("prev"), bits 0..14 the
// row reached (where the descent continues with bit 4 when not stopped).
constexpr uint32_t kDescStop = 0x80000000u;

SA_PAL_HD inline uint32_t descent3(const int16_t* child, int kids, uint32_t e) {
  int row = 0, prev = 0;
  bool stop = false;
  for (int b = 2; b >= 0; --b) {
    int id = (int)((e >> b) & 1u) | (int)(((e >> (3 + b)) & 1u) << 1) |
             (int)(((e >> (6 + b)) & 1u) << 2);
    if (kids == 16) id |= (int)(((e >> (9 + b)) & 1u) << 3);
    const int ch = child[row * kids + id];
    if (ch == kNoChild) {
      stop = true;
      break;
    }
    prev = row;
    if (ch == kLeaf) {
      stop = true;
      break;
    }
    row = ch;
  }
  return (stop ? kDescStop : 0u) | ((uint32_t)prev << 16) | (uint32_t)row;
}

}  // namespace sa
