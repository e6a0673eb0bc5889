// sa_search.h -- the per-lane palette search of the block and Jacobi engines.
//
// d_select_index without a memo (include/rd_cuda_common.cuh:408-446): the
// descent of src/rd_riemersma_cpu.cpp:82-89 and the ordered post-order scan of
// src/rd_octree.cpp:178-217, with candidate_distance's arithmetic.  Each lane
// searches on its own (the sequential walk scores one search across the whole
// wavefront instead, sa_walk.hip).
//
// The search image (sa_palette.h) is staged in LDS by the workgroup, and each
// candidate's distance constants beside it as doubles:
//   dcent[k] = (alpha e0, alpha e1, alpha e2, e3) of the k-th candidate in
//   post-order, alpha = QuantumScale e3 when associated, 1 otherwise --
// the products candidate_distance forms from the entry, formed the same way,
// so the scan reads 32 bytes per candidate and converts nothing.  The block
// walk stages the whole image and the constants after it; the Jacobi sweep,
// which never reads the associated colours, puts the constants in their room.
#pragma once

#include "sa_math.h"
#include "sa_palette.h"

namespace sa {

struct LSearch {
  const double* col;      // associated colour per entry (the residual's subtrahend)
  const float4* ent;      // the entries
  const int2* sub;        // candidate range per internal node
  const uint32_t* desc;   // the descent's first three levels
  const int16_t* child;
  const uint16_t* cand;   // palette index per post-order candidate
  const double4* dcent;   // distance constants per post-order candidate
  int kids;
  int count;
};

// Copies the search image to lds[0, pd.bytes) and builds dcent[0, pd.count)
// (16-byte aligned, in LDS); every thread of the workgroup calls it, and the
// caller synchronises before the first search.
template <bool ASSOC>
__device__ __forceinline__ LSearch stage_search(const PaletteDev& pd, unsigned char* lds,
                                                double4* dcent, int tid, int nthreads) {
  const uint4* src = (const uint4*)pd.image;
  uint4* dst = (uint4*)lds;
  for (int k = tid; k < pd.bytes / 16; k += nthreads) dst[k] = src[k];
  const float4* cent = (const float4*)((const unsigned char*)pd.image + pd.off_cent);
  for (int k = tid; k < pd.count; k += nthreads) {
    const float4 e = cent[k];
    double a = 1.0;
    if (ASSOC) a = kScale * (double)e.w;
    dcent[k] = make_double4(a * (double)e.x, a * (double)e.y, a * (double)e.z, (double)e.w);
  }
  LSearch s;
  s.col = (const double*)lds;
  s.ent = (const float4*)(lds + pd.off_ent);
  s.sub = (const int2*)(lds + pd.off_sub);
  s.desc = (const uint32_t*)(lds + pd.off_desc);
  s.child = (const int16_t*)(lds + pd.off_child);
  s.cand = (const uint16_t*)(lds + pd.off_cand);
  s.dcent = dcent;
  s.kids = pd.kids;
  s.count = pd.count;
  return s;
}

// The palette index for target t (doubles) whose channels clamp to pc
// (pc[c] = clamp_pixel(t[c])): descent on the channel bytes scale_char(pc),
// then the last candidate at the running minimum of the distance.
template <bool ASSOC>
__device__ __forceinline__ int lane_lookup(const LSearch& s, const float pc[4], const double t[4]) {
  const unsigned c0 = scale_char(pc[0]);
  const unsigned c1 = scale_char(pc[1]);
  const unsigned c2 = scale_char(pc[2]);
  const unsigned c3 = ASSOC ? scale_char(pc[3]) : 0u;
  uint32_t de = (c0 >> 5) | ((c1 >> 5) << 3) | ((c2 >> 5) << 6);
  if (ASSOC) de |= (c3 >> 5) << 9;
  const uint32_t d3 = s.desc[de];
  int prev = (int)((d3 >> 16) & 0x7FFFu);
  if (!(d3 & kDescStop)) {
    int row = (int)(d3 & 0x7FFFu);
    for (int lev = 4; lev > 0; --lev) {
      const int ch = s.child[row * s.kids + node_id(ASSOC, c0, c1, c2, c3, lev)];
      if (ch == kNoChild) break;
      prev = row;
      if (ch == kLeaf) break;
      row = ch;
    }
  }
  const int2 r = s.sub[prev];
  // candidate_distance's per-target factors: beta = QuantumScale t3 when
  // associated (beta t_c formed once), 1 otherwise (1 * t_c == t_c).
  double bt0 = t[0], bt1 = t[1], bt2 = t[2];
  if (ASSOC) {
    const double beta = kScale * t[3];
    bt0 = beta * t[0];
    bt1 = beta * t[1];
    bt2 = beta * t[2];
  }
  double best = kFarDistance;
  int bk = 0;
  const double4* dc = s.dcent + r.x;
  for (int k = 0; k < r.y; ++k) {
    const double4 e = dc[k];
    double v = e.x - bt0;
    double d = v * v;
    v = e.y - bt1;
    d += v * v;
    v = e.z - bt2;
    d += v * v;
    if (ASSOC) {
      v = e.w - t[3];
      d += v * v;
    }
    if (d <= best) {
      best = d;
      bk = k;
    }
  }
  return (int)s.cand[r.x + bk];
}

}  // namespace sa
