// sa_octree.hip -- ColorTree::Build on the device, bit for bit.
//
// Restates src/rd_octree.cpp:219-409 (quantize.c's Classify / PruneToCubeDepth /
// Reduce / DefineImageColormap).  The O(pixels) classification is data-parallel
// integer work; the O(nodes) tree walks run as level-synchronous sweeps of one
// workgroup.  Every double the reference accumulates is accumulated here in the
// reference's order.
//
//   1. Runs.  quantize.c collapses runs of identical adjacent pixels per row and
//      weights them by their length (rd_octree.cpp:305-313).  Run starts are
//      counted per tile of pixels, a scan over tiles numbers them, and a second
//      pass writes one record per run.
//   2. Loop-1 extent.  Loop 1 classifies rows to depth 8 until the number of
//      distinct depth-8 leaves exceeds the palette size; the row where that
//      happens is found by one workgroup inserting runs in order into a hash set
//      and ranking first occurrences with a prefix sum (rd_octree.cpp:364-375).
//   3. Nodes.  A node is (level, top `level` bits of each channel byte).  Every
//      (run, level) pair of loop 1 (levels 1..8) and loop 2 (levels 1..depth)
//      is inserted into a hash table that keeps the pair's first index; sorting
//      nodes by first pair reproduces the reference's creation order, hence its
//      node numbering.
//   4. Ordered sums.  quantize_error, total_color and number_unique of a node
//      are sums over its pairs in run order.  Pairs are stored level-major
//      (each level's pairs are one run sequence), so the consecutive runs that
//      share a node at a level form an interval of consecutive pairs; a stable
//      radix sort of the intervals by node, expanded, puts each node's pairs in
//      run order (at low levels an interval spans many runs, so there are far
//      fewer intervals than pairs) -- or, when the cube depth leaves few cells,
//      loop 2's pairs are placed by a counting sort over cells (section 3b);
//      every pair's term is computed in parallel,
//      then each node's chain is folded in order over contiguous memory.  The
//      totals of loop 2 are folded after
//      PruneToCubeDepth has merged loop 1's deep leaves, exactly when the
//      reference adds them.
//   5. Tree walks (section 5 explains the restatement): the gated
//      PruneToCubeDepth, the reduction threshold, the Reduce passes,
//      DefineImageColormap and the search image (sa_palette.h).
//
// The root's quantize_error.  The reference adds every node's running total
// into the root after each level update -- one serial chain over all pairs.  Its
// exact value can only matter if it comes within 1 of another node's error (it
// is then the flattening's maximum by more than MagickEpsilon, never the
// threshold, never the minimum a Reduce pass can reach, never pruned).  The
// chain is summed in parallel, its rounding error bounded, and the exact serial
// chain is run only when the bound cannot separate it (tiny images).
//
// Trees whose live node count could pass MaxQNodes (266817) during
// classification, where the reference starts pruning levels mid-image
// (rd_octree.cpp:366-369, 380-383), are classified again row by row (section
// 6, k_rows); the tree logic and the colormap are then the same kernels'.
#include <hipcub/hipcub.hpp>
#include <rocprim/block/block_radix_sort.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <type_traits>
#include <vector>

#include "sa_internal.h"
#include "sa_math.h"
#include "sa_octree.h"

namespace sa {
namespace {

constexpr uint32_t kEmpty32 = 0xFFFFFFFFu;
constexpr uint64_t kEmpty64 = 0ull;
constexpr int kMaxQNodes = 266817;

__device__ __forceinline__ bool same_px(const float4& a, const float4& b) {
  return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;
}

__device__ __forceinline__ uint32_t hash32(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return (uint32_t)k;
}

// Node key: level in bits 40..43, the top `level` bits of channel c in bits
// 8c..8c+7 (alpha only when associated).  Never 0 for level >= 1.
__device__ __forceinline__ uint64_t node_key(uint32_t pack, int level, bool assoc) {
  const int sh = 8 - level;
  uint64_t k = (uint64_t)((pack & 0xFFu) >> sh) | ((uint64_t)(((pack >> 8) & 0xFFu) >> sh) << 8) |
               ((uint64_t)(((pack >> 16) & 0xFFu) >> sh) << 16);
  if (assoc) k |= (uint64_t)((pack >> 24) >> sh) << 24;
  return k | ((uint64_t)level << 40);
}

__device__ __forceinline__ uint64_t parent_key(uint64_t k) {
  const int level = (int)(k >> 40);
  const uint64_t p = k & 0xFFFFFFFFull;
  const uint64_t q = ((p >> 1) & 0x7F7F7F7Full);
  return q | ((uint64_t)(level - 1) << 40);
}

__device__ __forceinline__ int child_index(uint64_t k) {
  return (int)((k & 1) | (((k >> 8) & 1) << 1) | (((k >> 16) & 1) << 2) | (((k >> 24) & 1) << 3));
}

// Midpoint of a node's cube (rd_octree.cpp:318-329); every step is exact.
__device__ __forceinline__ void node_mid(uint64_t k, double mid[4]) {
  const int level = (int)(k >> 40);
  for (int c = 0; c < 4; ++c) {
    const uint32_t bits = (uint32_t)((k >> (8 * c)) & 0xFF);
    double m = kRange / 2.0;
    double bisect = (kRange + 1.0) / 2.0;
    for (int l = 1; l <= level; ++l) {
      bisect *= 0.5;
      const uint32_t b = (bits >> (level - l)) & 1u;
      m += b ? bisect : -bisect;
    }
    mid[c] = m;
  }
}

struct Terms {
  double* term;   // count * |pixel - mid| per sorted pair
  float4* lval;   // clamped leaf channels (leaf pairs only)
  int32_t* lcnt;  // run length per sorted pair
};

// Per-node scratch of the tree sweeps (M + 1 entries each unless noted).
struct TreeWork {
  int32_t* lvl;      // node ids grouped by level
  int32_t* lvl_off;  // 10 level offsets into lvl
  int32_t* att;      // attached to the root at the start of the sweep
  int32_t* anc;      // lowest S ancestor-or-self, -1 when none
  int32_t* size;     // attached subtree size
  int32_t* gsz;      // size of the node's part of its group
  int32_t* start;    // post-order index of the first node of the subtree
  int32_t* gstart;   // the same, counted inside the group
  int32_t* pre;      // pre-order index
  int32_t* merged;   // merged by this pass (later: internal-node flag)
  int32_t* arr;      // scan buffers, indexed by post- or pre-order
  int32_t* arr2;
};

// ---- batches ----------------------------------------------------------------------
// A batch of images is built phase by phase, each phase one launch per kernel
// for every image: element-parallel kernels flatten (image, element) onto one
// grid through a table of each image's first block, the tree-logic kernels run
// one workgroup per image, and the batch's pair and node arrays are the
// images' arrays laid end to end (an image's part starts at its pair_base /
// node_base).  The host reads back between phases only the few scalars that
// size the next phase's arrays.
struct Frame {
  const void* src;
  int32_t fmt, assoc, depth, kids;
  int32_t maxc;
  int32_t ok;              // 0 once the image has failed: later phases skip it
  int32_t exact_root;      // 1: the root's chain is summed serially whatever the bound
  int64_t w, h, N;
  int32_t* info;           // 16 ints, layout in sa_octree.h
  double* dinfo;           // [0] parallel root chain, [1] max non-root qe, [2] root qe used
  // 1. runs
  int32_t px_pack;         // 1: no pixel copy, run_px rebuilds it from the pack
  int64_t R;
  int32_t* start;
  int32_t* count;
  int32_t* row;
  float4* px;              // the run's pixel as loaded (associated where it is read)
  uint32_t* pack;
  // 2. loop-1 extent
  uint64_t* bkeys;
  uint32_t* bfirst;
  uint32_t bcap;
  // 3. nodes
  uint32_t ncap;
  int64_t R1, P;
  int64_t pair_base;       // offset of this image's pairs in the batch's pair arrays
  uint64_t* nkeys;
  uint32_t* nfirst;
  uint32_t* pair_slot;
  // 4. ordered sums and tree logic
  int64_t M;
  int64_t node_base;       // offset of this image's nodes in the batch's node arrays
  int64_t entry_base;      // offset of this image's nodes in the sorted node entries
  int32_t* slot_node;
  TreeDev t;
  TreeWork W;
  int32_t* seg_b;
  int32_t* seg_e;
  const uint32_t* spair;   // pair indices of this image in sorted pair order
  Terms T;
  double* record;
  PaletteBuild pb;
  // 3b. loop 2 placed by cell (images whose cube depth has at most
  // kDenseCellsMax cells in all)
  int32_t dense;
  int32_t dcells;          // cells of levels 1..depth
  int32_t dtshift;         // log2 of the loop-2 runs per tile
  int32_t dtiles;
  int64_t P_ins;           // pairs inserted through the hash: 8 R1 when dense, else P
  int64_t dbase;           // this image's first cell in the batch's cell arrays
  int32_t* dtab;           // [tile][cell] pair counts, then their exclusive prefix over tiles
  uint32_t* dfirst;        // per cell: its first run
  int32_t* dtot;           // per cell: its loop-2 pairs
  uint32_t* dslot;         // per cell: the hash slot of its node
  const uint32_t* dpos;    // per cell: first sorted position (batch-wide) of its loop-2 pairs
  int32_t* l2s;            // per node: first sorted position of its loop-2 pairs (INT32_MAX: none)
  int32_t exp1;            // its k_compact intervals are expanded one per thread (k_expand_short)
};

// Batch-wide arrays and counters.
struct Batch {
  int32_t* ctl;            // [0] node entries; medium [1] [2], long [3] [4], huge [5] [6]
                           // segments of the loop-1 / loop-2 folds; [8 + FoldMode] queue heads
                           // of the long folds, [13] [14] of the medium ones
  uint32_t* ekey;          // node entries: pair_base + first pair
  uint32_t* eslot;         // node entries: hash slot
  const uint32_t* skey;    // the same, sorted
  const uint32_t* sslot;
  uint64_t* mid_list;      // (image << 32) | node, segments folded by one wavefront
  uint64_t* long_list;     // the same, segments folded by one workgroup
  uint64_t* huge_list;     // the same, the longest: taken first
};

struct Map {
  const int64_t* first;    // first block of image f; first[F] = grid size
  int32_t F;
};

struct At {
  int32_t f;
  int64_t i;
};

// The image of this block (the last one whose first block is <= blockIdx.x,
// which skips images with no blocks) and the thread's element in it.
__device__ __forceinline__ At map_at(const Map& m, int per) {
  int32_t lo = 0, hi = m.F - 1;
  const int64_t b = blockIdx.x;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (m.first[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return {lo, (b - m.first[lo]) * per + (int64_t)threadIdx.x};
}

struct TileAt {
  int32_t f;
  int64_t t;
};

// Tile of block b (one block per tile).
__device__ __forceinline__ TileAt map_tile(const Map& m, int64_t b) {
  int32_t lo = 0, hi = m.F - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (m.first[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return {lo, b - m.first[lo]};
}

// ---- 1. runs ------------------------------------------------------------------
// info zeroed, info[5] = h, the double results zeroed (one workgroup per image).
__global__ void k_frame_init(const Frame* __restrict__ fr) {
  const Frame& d = fr[blockIdx.x];
  if (threadIdx.x < 16) d.info[threadIdx.x] = threadIdx.x == 5 ? (int32_t)d.h : 0;
  if (threadIdx.x < 4) d.dinfo[threadIdx.x] = 0.0;
}

// Runs are found a tile of kRunTile pixels at a time (one 256-thread workgroup
// per tile, each wavefront a quarter of it; a tile never spans two images):
// k_run_tiles counts the run starts of every quarter, an exclusive scan over
// the batch's quarters gives each quarter's first run, and k_run_emit writes
// every run's record, its length too unless it is its quarter's last run,
// whose length k_run_tail takes from the next start.  The wavefronts need no
// workgroup synchronisation.
constexpr int kRunTile = 4096;

// The run's pixel as loaded.  An 8-bit source read without association keeps
// no copy: its channels are the packed bytes times 257 (load_pixel), and its
// alpha is never read.
__device__ __forceinline__ float4 run_px(const float4* px, const uint32_t* pack, int px_pack,
                                         int64_t r) {
  if (!px_pack) return px[r];
  const uint32_t pk = pack[r];
  return make_float4((float)(pk & 0xFFu) * 257.0f, (float)((pk >> 8) & 0xFFu) * 257.0f,
                     (float)((pk >> 16) & 0xFFu) * 257.0f, 65535.0f);
}

// Pixel p (consecutive in the wavefront's lanes) starts a run: the first of
// its row, or unlike the pixel before it (rd_octree.cpp:305-313), which the
// lane below holds (lane 0 loads it).  Every lane of the wavefront calls it.
__device__ __forceinline__ bool run_start(const Frame& d, int64_t p, bool valid, float4* v) {
  *v = valid ? load_any(d.fmt, d.src, p, d.w, d.h) : make_float4(0.f, 0.f, 0.f, 0.f);
  float4 u;
  u.x = __shfl_up(v->x, 1, 64);
  u.y = __shfl_up(v->y, 1, 64);
  u.z = __shfl_up(v->z, 1, 64);
  u.w = __shfl_up(v->w, 1, 64);
  const bool row0 = !valid || p % d.w == 0;
  if ((threadIdx.x & 63) == 0 && !row0) u = load_any(d.fmt, d.src, p - 1, d.w, d.h);
  return valid && (row0 || !same_px(*v, u));
}

constexpr int kRunQuarter = kRunTile / 4;  // pixels of one wavefront

__global__ __launch_bounds__(256) void k_run_tiles(const Frame* __restrict__ fr, Map m,
                                                   int32_t* __restrict__ tcnt) {
  const TileAt tl = map_tile(m, blockIdx.x);
  const Frame& d = fr[tl.f];
  const int wave = threadIdx.x >> 6;
  const int64_t p0 = tl.t * kRunTile + wave * kRunQuarter;
  int c = 0;
  for (int k = 0; k < kRunQuarter / 64; ++k) {
    const int64_t p = p0 + k * 64 + (threadIdx.x & 63);
    float4 v;
    c += __popcll(__ballot(run_start(d, p, p < d.N, &v)));
  }
  if ((threadIdx.x & 63) == 0) tcnt[4 * (int64_t)blockIdx.x + wave] = c;
}

// info[12] <- R, from the scan over the image's quarters.
__global__ void k_run_count(const Frame* __restrict__ fr, Map m, const int32_t* __restrict__ tcnt,
                            const int32_t* __restrict__ toff) {
  const Frame& d = fr[blockIdx.x];
  if (threadIdx.x != 0) return;
  const int64_t q0 = 4 * m.first[blockIdx.x], q1 = 4 * m.first[blockIdx.x + 1];
  d.info[12] = q1 > q0 ? toff[q1 - 1] + tcnt[q1 - 1] - toff[q0] : 0;
}

// Every run start of the wavefront's quarter, in pixel order: its start, row,
// pack, (when kept) pixel, and its length when the next start is in the
// quarter.  64 pixels at a time; a start whose successor is not among them is
// resolved by the next round with a start.  For an 8-bit source read without
// association the pack is the source bytes (scale_char of a byte times 257 is
// the byte), alpha 255 for RGB.
__global__ __launch_bounds__(256) void k_run_emit(const Frame* __restrict__ fr, Map m,
                                                  const int32_t* __restrict__ toff) {
  const TileAt tl = map_tile(m, blockIdx.x);
  const Frame& d = fr[tl.f];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t q = 4 * (int64_t)blockIdx.x + wave;
  const int64_t p0 = tl.t * kRunTile + wave * kRunQuarter;
  int64_t base = toff[q] - toff[4 * m.first[tl.f]];
  int64_t pend_r = -1, pend_p = 0;  // the start waiting for its successor (uniform)
  for (int k = 0; k < kRunQuarter / 64; ++k) {
    const int64_t pc = p0 + 64 * k;
    const int64_t p = pc + lane;
    float4 v;
    const bool st = run_start(d, p, p < d.N, &v);
    const uint64_t bal = __ballot(st);
    if (!bal) continue;
    if (st) {
      const int64_t r = base + __popcll(bal & ((1ull << lane) - 1));
      uint32_t pk;
      if (d.px_pack) {
        pk = (uint32_t)v.x / 257u | ((uint32_t)v.y / 257u) << 8 | ((uint32_t)v.z / 257u) << 16 |
             ((uint32_t)v.w / 257u) << 24;
      } else {
        double a[4];
        assoc_pixel(d.assoc != 0, v.x, v.y, v.z, v.w, a);
        d.px[r] = v;
        const uint32_t c0 = scale_char(clamp_pixel(a[0])), c1 = scale_char(clamp_pixel(a[1]));
        const uint32_t c2 = scale_char(clamp_pixel(a[2])), c3 = scale_char(clamp_pixel(a[3]));
        pk = c0 | (c1 << 8) | (c2 << 16) | (c3 << 24);
      }
      d.pack[r] = pk;
      d.start[r] = (int32_t)p;
      d.row[r] = (int32_t)(p / d.w);
      const uint64_t above = bal & ~((2ull << lane) - 1);
      if (above) d.count[r] = __ffsll((unsigned long long)above) - 1 - lane;
    }
    const int first = __ffsll((unsigned long long)bal) - 1, last = 63 - __clzll((long long)bal);
    if (lane == 0 && pend_r >= 0) d.count[pend_r] = (int32_t)(pc + first - pend_p);
    pend_r = base + __popcll(bal) - 1;
    pend_p = pc + last;
    base += __popcll(bal);
  }
}

// The length of each quarter's last run: up to the next run's start, or to
// the end of the image.
__global__ void k_run_tail(const Frame* __restrict__ fr, Map m, const int32_t* __restrict__ tcnt,
                           const int32_t* __restrict__ toff, int64_t nq) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nq || tcnt[b] == 0) return;
  const TileAt tl = map_tile(m, b / 4);
  const Frame& d = fr[tl.f];
  const int64_t r = toff[b] + tcnt[b] - 1 - toff[4 * m.first[tl.f]];
  const int64_t e = r + 1 < d.R ? d.start[r + 1] : d.N;
  d.count[r] = (int32_t)(e - d.start[r]);
}

// ---- 2. loop-1 extent -----------------------------------------------------------
// info[0] <- y1 (the row after which the distinct leaf count exceeds maxc, or h)
// info[1] <- R1 (first run of loop 2, R when loop 1 covers the image)
__global__ __launch_bounds__(1024) void k_break(const Frame* __restrict__ fr) {
  const Frame& d = fr[blockIdx.x];
  if (!d.ok) return;
  const uint32_t* pack = d.pack;
  const int32_t* row = d.row;
  const int64_t R = d.R, h = d.h;
  const int assoc = d.assoc, maxc = d.maxc;
  uint64_t* hkeys = d.bkeys;
  uint32_t* hfirst = d.bfirst;
  const uint32_t cap = d.bcap;
  int32_t* info = d.info;
  __shared__ int32_t s_total;
  __shared__ int32_t s_wave[16];
  __shared__ int64_t s_hit;
  if (threadIdx.x == 0) {
    s_total = 0;
    s_hit = -1;
  }
  __syncthreads();
  const uint32_t mask_key = assoc ? 0xFFFFFFFFu : 0x00FFFFFFu;
  for (int64_t base = 0; base < R; base += 1024) {
    const int64_t r = base + threadIdx.x;
    uint32_t slot = 0;
    bool have = r < R;
    const uint64_t key = have ? (uint64_t)(pack[r] & mask_key) + 1ull : 0ull;
    if (have) {
      slot = hash32(key) & (cap - 1);
      uint32_t probe = 0;
      for (;; ++probe) {
        if (probe >= cap) {  // table full: report, never spin
          info[6] = 4;
          have = false;
          break;
        }
        const unsigned long long prev = atomicCAS((unsigned long long*)&hkeys[slot],
                                                  (unsigned long long)kEmpty64,
                                                  (unsigned long long)key);
        if (prev == kEmpty64 || prev == key) break;
        slot = (slot + 1) & (cap - 1);
      }
      if (have) atomicMin(&hfirst[slot], (uint32_t)r);
    }
    __threadfence_block();
    __syncthreads();
    const int isnew = (have && hfirst[slot] == (uint32_t)r) ? 1 : 0;
    // Inclusive prefix of isnew over the chunk, in run order.
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int incl = isnew;
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    if (lane == 63) s_wave[wv] = incl;
    __syncthreads();
    int before = s_total;
    for (int k = 0; k < wv; ++k) before += s_wave[k];
    const int cum = before + incl;
    if (have && isnew && cum == maxc + 1) s_hit = r;  // exactly one run crosses
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 0; k < 16; ++k) s_total += s_wave[k];
    }
    __syncthreads();
    if (s_hit >= 0) break;
  }
  if (threadIdx.x == 0) {
    if (s_hit < 0) {
      info[0] = (int32_t)h;
      info[1] = (int32_t)R;
    } else {
      const int32_t y1 = row[s_hit];
      info[0] = y1;
      // First run of row y1+1: row[] does not decrease, so a binary search.
      int64_t lo = s_hit, hi = R;
      while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if (row[mid] <= y1)
          lo = mid + 1;
        else
          hi = mid;
      }
      info[1] = (int32_t)lo;
    }
  }
}

// ---- 3. nodes ---------------------------------------------------------------------
// Pair numbering.  Loop 1 classifies runs [0, R1) at levels 1..8, loop 2 runs
// [R1, R) at levels 1..depth.  Pairs are stored level-major within each loop:
//   loop 1: p = (l - 1) R1 + r,                     p in [0, 8 R1)
//   loop 2: p = 8 R1 + (l - 1) (R - R1) + (r - R1), p in [8 R1, P)
// so a level's pairs are consecutive runs.  The reference's order of the pairs
// (run-major: its creation order, and the order of the root chain) is
// pair_rmajor; both numberings put loop 1 first and order one node's pairs
// (all at one level) by run.
__device__ __forceinline__ void pair_decode(uint32_t p, int64_t R1, int64_t R, int depth,
                                            int64_t* r, int* l) {
  const int64_t l1 = 8 * R1;
  int64_t q, n;
  int64_t base;
  int levels;
  if ((int64_t)p < l1) {
    q = p;
    n = R1;
    base = 0;
    levels = 8;
  } else {
    q = (int64_t)p - l1;
    n = R - R1;
    base = R1;
    levels = depth;
  }
  int k = 0;  // level - 1: the number of level blocks wholly before q
  int64_t edge = n;
  for (int j = 1; j < 8; ++j) {
    if (j < levels && q >= edge) {
      ++k;
      edge += n;
    }
  }
  *l = k + 1;
  *r = base + q - (int64_t)k * n;
}

__device__ __forceinline__ uint32_t pair_rmajor(int64_t r, int l, int64_t R1, int depth) {
  return (uint32_t)(r < R1 ? 8 * r + (l - 1) : 8 * R1 + (r - R1) * depth + (l - 1));
}

// A pair that reaches a leaf: level 8 in loop 1, level depth in loop 2.
__device__ __forceinline__ bool leaf_pair(uint32_t p, int64_t R1, int depth, int l) {
  return p < 8ull * (uint64_t)R1 ? l == 8 : l == depth;
}

// The first pair of an interval (see the top of the file) inserts its node;
// the other pairs of an interval share the node of the pair before them, so
// they insert nothing and never need their slot: k_compact reads the node of
// each interval's first pair.  This also keeps the smooth regions of an image,
// where thousands of consecutive runs share their low-level nodes, from
// serialising on a few hash slots.  A node's first pair (its creation order)
// is the least run-major index of its pairs, which is always an interval's
// first.  wgc[workgroup] counts the intervals that begin in the workgroup's
// 256 pairs (their sum, the batch's intervals, is reduced into ctl[12]).
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

__global__ __launch_bounds__(256) void k_node_insert(const Frame* __restrict__ fr, Map m,
                                                     int32_t* __restrict__ wgc) {
  __shared__ int s_starts;
  if (threadIdx.x == 0) s_starts = 0;
  __syncthreads();
  const At at = map_at(m, 256);
  const Frame& d = fr[at.f];
  const int64_t p = at.i;
  const bool valid = p < d.P_ins;
  const int64_t R1 = d.R1;
  const int depth = d.depth, assoc = d.assoc;
  int64_t r = 0;
  int l = 1;
  uint64_t key = 0;
  bool start = false;
  if (valid) {
    pair_decode((uint32_t)p, R1, d.R, depth, &r, &l);
    key = node_key(d.pack[r], l, assoc != 0);
    const bool first_run = r == 0 || r == R1;  // the first run of its loop
    start = first_run || node_key(d.pack[r - 1], l, assoc != 0) != key;
  }
  const unsigned long long starts = __ballot(start);
  if ((threadIdx.x & 63) == 0 && starts) atomicAdd(&s_starts, __popcll(starts));
  __syncthreads();
  if (threadIdx.x == 0) wgc[blockIdx.x] = s_starts;
  if (!valid) return;
  uint32_t* pair_slot = d.pair_slot;
  if (!start) {
    pair_slot[p] = kNoSlot;
    return;
  }
  uint64_t* hkeys = d.nkeys;
  uint32_t* hfirst = d.nfirst;
  const uint32_t cap = d.ncap;
  int32_t* info = d.info;
  const uint32_t c = pair_rmajor(r, l, R1, depth);
  // A plain read first: a key never changes once set and `first` only
  // decreases, so a stale read costs at most an atomic that finds the same.
  // Most intervals find their node already inserted, by an earlier one.
  uint32_t slot = hash32(key) & (cap - 1);
  for (uint32_t probe = 0;; ++probe) {
    if (probe >= cap) {  // table full: report, never spin
      info[6] = 4;
      pair_slot[p] = 0;
      return;
    }
    const uint64_t cur = hkeys[slot];
    if (cur == key) break;
    if (cur == kEmpty64) {
      const unsigned long long prev =
          atomicCAS((unsigned long long*)&hkeys[slot], (unsigned long long)kEmpty64,
                    (unsigned long long)key);
      if (prev == kEmpty64 || prev == key) break;
    }
    slot = (slot + 1) & (cap - 1);
  }
  if (hfirst[slot] > c) atomicMin(&hfirst[slot], c);
  pair_slot[p] = slot;
}

// Node entries (first pair, slot) of every image into one list; sorting it by
// pair_base + first pair orders the batch by image, then by creation.
// A workgroup's threads belong to one image, so the workgroup takes its
// entries' places with one atomic.
__global__ __launch_bounds__(256) void k_node_collect(const Frame* __restrict__ fr, Map m, Batch B) {
  __shared__ int s_w[4];
  __shared__ int s_base;
  const At at = map_at(m, 256);
  const Frame& d = fr[at.f];
  const int64_t s = at.i;
  const bool used = s < d.ncap && d.nkeys[s] != kEmpty64;
  const uint64_t bal = __ballot(used);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_w[wave] = __popcll(bal);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int all = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    s_base = all ? atomicAdd(&B.ctl[0], all) : 0;
    if (all) atomicAdd(&d.info[13], all);
  }
  __syncthreads();
  if (!used) return;
  int k = s_base + __popcll(bal & ((1ull << lane) - 1));
  for (int w = 0; w < wave; ++w) k += s_w[w];
  B.ekey[k] = (uint32_t)d.pair_base + d.nfirst[s];
  B.eslot[k] = (uint32_t)s;
}

// Node ids in creation order (root is 0); creation counters for the colour count.
__global__ void k_node_assign(const Frame* __restrict__ fr, Map m, Batch B) {
  const At at = map_at(m, 256);
  const Frame& d = fr[at.f];
  const int64_t k = at.i;
  const bool live = k < d.M && d.ok;
  bool leaf1 = false, new2 = false, loop1 = false;
  if (live) {
    const int32_t node = (int32_t)(k + 1);
    const uint32_t slot = B.sslot[d.entry_base + k];
    const uint32_t first = B.skey[d.entry_base + k] - (uint32_t)d.pair_base;
    d.slot_node[slot] = node;
    const uint64_t key = d.nkeys[slot];
    d.t.key[node] = key;
    const int level = (int)(key >> 40);
    loop1 = first < 8ull * (uint64_t)d.R1;
    leaf1 = loop1 && level == 8;            // leaves counted in loop 1
    new2 = !loop1 && level == d.depth;      // level-depth nodes new in loop 2
  }
  // One workgroup, one image: a count per wavefront.
  const uint64_t b2 = __ballot(leaf1), b3 = __ballot(new2), b4 = __ballot(loop1);
  if ((threadIdx.x & 63) == 0) {
    int32_t* info = d.info;
    if (b2) atomicAdd(&info[2], __popcll(b2));
    if (b3) atomicAdd(&info[3], __popcll(b3));
    if (b4) atomicAdd(&info[4], __popcll(b4));  // nodes created by loop 1
  }
}

// The intervals in pair order: the workgroups of k_node_insert's grid again,
// each writing its intervals from its offset wgo (the exclusive scan of wgc)
// on.  Interval k: its first pair istart[k] (batch-wide, ascending) and its
// sort key, the batch node id of that pair (an image that failed sorts onto its
// own single node id, so its pairs stay inside its range).  An interval runs
// up to the next one's first pair (every level block, and every image, begins
// one), the end of the batch's pairs after the last.
__global__ __launch_bounds__(256) void k_compact(const Frame* __restrict__ fr, Map m,
                                                 const int32_t* __restrict__ wgo,
                                                 const uint32_t* __restrict__ pair_slot,
                                                 uint32_t* __restrict__ istart,
                                                 int32_t* __restrict__ ikey,
                                                 uint32_t* __restrict__ ival) {
  __shared__ int s_wave[4];
  const At at = map_at(m, 256);
  const Frame& d = fr[at.f];
  const int64_t p = at.i;
  const int64_t g = d.pair_base + p;
  const bool start = p < d.P_ins && pair_slot[g] != kNoSlot;
  const unsigned long long bits = __ballot(start);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_wave[wave] = __popcll(bits);
  __syncthreads();
  if (!start) return;
  int k = wgo[blockIdx.x] + __popcll(bits & ((1ull << lane) - 1));
  for (int w = 0; w < wave; ++w) k += s_wave[w];
  istart[k] = (uint32_t)g;
  ikey[k] = (int32_t)d.node_base + (d.ok ? d.slot_node[pair_slot[g]] : 0);
  ival[k] = (uint32_t)k;
}

// The image whose pair range holds batch position i: the last with
// pair_base <= i (images without pairs have empty ranges at the running
// offset; an image's sorted positions are its pairs' range).
__device__ __forceinline__ int image_of_pair(const Frame* __restrict__ fr, int F, int64_t i) {
  int lo = 0, hi = F - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (fr[mid].pair_base <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Intervals [0, ni) come from k_compact; [ni, nt) stand for the loop-2 pairs
// of one cell of an image placed by cell (k_dense_virtual), their length in
// vlen.  An interval of k_compact runs up to the next one's first pair, and at
// most to the end of its image's inserted pairs.
constexpr uint32_t kVirtual = 0xFFFFFFFFu;

// Lengths in sorted order (then an exclusive scan gives each sorted
// interval's first position).
__global__ void k_interval_len(const Frame* __restrict__ fr, int F,
                               const uint32_t* __restrict__ sival,
                               const uint32_t* __restrict__ istart, int64_t ni, int64_t nt,
                               int64_t sum_p, const int32_t* __restrict__ vlen,
                               int32_t* __restrict__ slen) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nt) return;
  const int64_t k = sival[j];
  if (k >= ni) {
    slen[j] = vlen[k - ni];
    return;
  }
  const int64_t b = istart[k];
  int64_t e = k + 1 < ni ? (int64_t)istart[k + 1] : sum_p;
  const Frame& d = fr[image_of_pair(fr, F, b)];
  e = min(e, d.pair_base + d.P_ins);
  slen[j] = (int32_t)(e - b);
}

// Sorted positions are expanded a tile of kExpandTile at a time.  Sorted
// interval j covers positions [spos[j], spos[j] + slen[j]); it is the first
// interval of every tile whose first position it covers.  A node's segment
// (its pairs' sorted positions, as positions of its image) begins with its
// first interval and ends with its last.  A cell's interval (sival >= ni)
// records its first position in dpos; the empty ones carry the key sum_nodes,
// sort after every node and are skipped.
constexpr int kExpandTile = 1024;

__global__ void k_interval_ends(const Frame* __restrict__ fr, int F,
                                const int32_t* __restrict__ sikey,
                                const uint32_t* __restrict__ sival, const int32_t* __restrict__ spos,
                                const int32_t* __restrict__ slen, int64_t ni, int64_t nt,
                                int32_t sum_nodes, uint32_t* __restrict__ dpos,
                                int32_t* __restrict__ tile_first) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nt) return;
  const int32_t n = sikey[j];
  if (n >= sum_nodes) return;
  const int64_t b = spos[j], e = b + slen[j];
  for (int64_t t = (b + kExpandTile - 1) / kExpandTile; t * kExpandTile < e; ++t)
    tile_first[t] = (int32_t)j;
  const bool cell = sival[j] >= (uint64_t)ni;
  if (cell) dpos[sival[j] - ni] = (uint32_t)b;
  const bool first = j == 0 || sikey[j - 1] != n, last = j + 1 == nt || sikey[j + 1] != n;
  if (!first && !last && !cell) return;
  const Frame& d = fr[image_of_pair(fr, F, b)];
  if (!d.ok) return;
  if (cell) d.l2s[n - d.node_base] = (int32_t)(b - d.pair_base);
  if (first) d.seg_b[n - d.node_base] = (int32_t)(b - d.pair_base);
  if (last) d.seg_e[n - d.node_base] = (int32_t)(e - d.pair_base);
}

// Sorted position i of image d (li = i - pair_base): pair p of interval
// `node` (a batch node id); its term (distance x count), run length and leaf
// value, formed as quantize.c's classification accumulates them.
__device__ __forceinline__ void expand_one(const Frame& d, int64_t li, uint32_t p, int32_t node,
                                           uint32_t* __restrict__ spair) {
  spair[d.pair_base + li] = p;
  if (!d.ok) return;
  const int64_t R1 = d.R1;
  const int depth = d.depth, assoc = d.assoc;
  int64_t r;
  int l;
  pair_decode(p, R1, d.R, depth, &r, &l);
  double mid[4];
  node_mid(d.t.key[node - d.node_base], mid);
  const float4 f = run_px(d.px, d.pack, d.px_pack, r);
  double a[4];
  assoc_pixel(assoc != 0, f.x, f.y, f.z, f.w, a);
  const double4 v = make_double4(a[0], a[1], a[2], a[3]);
  const int32_t cnt = d.count[r];
  const double e0 = kScale * (v.x - mid[0]);
  const double e1 = kScale * (v.y - mid[1]);
  const double e2 = kScale * (v.z - mid[2]);
  const double e3 = assoc ? kScale * (v.w - mid[3]) : 0.0;
  double dd = e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3;
  if (dd != dd) dd = 0.0;
  const Terms& T = d.T;
  T.term[li] = (double)cnt * sqrt(dd);
  T.lcnt[li] = cnt;
  if (leaf_pair(p, R1, depth, l))
    T.lval[li] = make_float4(clamp_pixel(v.x), clamp_pixel(v.y), clamp_pixel(v.z),
                             clamp_pixel(assoc ? v.w : kRange));
}

// Tile t: its intervals (at most kExpandTile + 1 touch it) are staged in LDS,
// and a bitmap marks the tile offsets where one of them begins (after the
// first, which covers the tile's first position).  Sorted position i takes the
// interval whose start is the last at or before it -- the count of marks up
// to its offset -- that interval's node, and the pair its first pair plus the
// offset into it.
__global__ __launch_bounds__(256) void k_expand(const Frame* __restrict__ fr, int F, int64_t sum_p,
                                                const int32_t* __restrict__ tiles,
                                                int64_t ni, const int32_t* __restrict__ tile_first,
                                                const int32_t* __restrict__ sikey,
                                                const uint32_t* __restrict__ sival,
                                                const uint32_t* __restrict__ istart,
                                                const int32_t* __restrict__ spos,
                                                uint32_t* __restrict__ spair) {
  constexpr int kWords = kExpandTile / 32;
  __shared__ int32_t l_pos[kExpandTile + 1];
  __shared__ int32_t l_node[kExpandTile + 1];
  __shared__ uint32_t l_first[kExpandTile + 1];
  __shared__ uint32_t s_mark[kWords];
  __shared__ int32_t s_before[kWords];  // marks in the words before
  __shared__ int s_f;
  __shared__ int64_t s_next;
  const int64_t tile = tiles ? tiles[blockIdx.x] : blockIdx.x;  // null: every tile
  const int64_t i0 = tile * kExpandTile;
  const int64_t i1 = i0 + kExpandTile < sum_p ? i0 + kExpandTile : sum_p;
  const int64_t j0 = tile_first[tile];
  // A tile inside one cell's interval has nothing to expand.
  if (istart[sival[j0]] == kVirtual && (j0 + 1 >= ni || spos[j0 + 1] >= i1)) return;
  if (threadIdx.x < kWords) s_mark[threadIdx.x] = 0;
  __syncthreads();
  for (int k0 = 0; k0 <= kExpandTile; k0 += blockDim.x) {
    const int k = k0 + threadIdx.x;
    const int64_t j = j0 + k;
    bool past = k > kExpandTile || j >= ni;
    if (!past) {
      const int32_t pos = spos[j];
      past = pos >= i1;
      if (!past) {
        l_pos[k] = pos;
        l_node[k] = sikey[j];
        l_first[k] = istart[sival[j]];
        if (k > 0) {
          const int o = (int)(pos - i0);  // in (0, kExpandTile)
          atomicOr(&s_mark[o >> 5], 1u << (o & 31));
        }
      }
    }
    if (__syncthreads_or(past)) break;  // intervals start in ascending order
  }
  if (threadIdx.x == 0) {
    const int f = image_of_pair(fr, F, i0);
    s_f = f;
    s_next = f + 1 < F ? fr[f + 1].pair_base : INT64_MAX;
  }
  __syncthreads();
  if (threadIdx.x < 64) {  // one wavefront: exclusive prefix of the word counts
    const int w = threadIdx.x;
    int c = w < kWords ? __popc(s_mark[w]) : 0;
    int incl = c;
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(incl, d, 64);
      if (w >= d) incl += v;
    }
    if (w < kWords) s_before[w] = incl - c;
  }
  __syncthreads();
  const int f0 = __builtin_amdgcn_readfirstlane(s_f);
  const bool one_image = i1 <= s_next;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int o = (int)(i - i0);
    const uint32_t below = (o & 31) == 31 ? 0xFFFFFFFFu : ((2u << (o & 31)) - 1u);
    const int k = s_before[o >> 5] + __popc(s_mark[o >> 5] & below);
    if (l_first[k] == kVirtual) continue;  // placed by k_dense_scatter
    const int64_t g = (int64_t)l_first[k] + (i - l_pos[k]);
    const Frame& d = fr[one_image ? f0 : image_of_pair(fr, F, i)];
    if (d.exp1) continue;  // expanded by k_expand_short
    expand_one(d, i - d.pair_base, (uint32_t)(g - d.pair_base), l_node[k], spair);
  }
}

// The k_compact intervals of an image whose loop 1 is short (placed by cell,
// a few rows at most): one thread per sorted interval writes its pairs, so
// the sorted range of such images needs no tile at all.
__global__ void k_expand_short(const Frame* __restrict__ fr, int F, int64_t ni, int64_t nt,
                               const int32_t* __restrict__ sikey,
                               const uint32_t* __restrict__ sival,
                               const uint32_t* __restrict__ istart,
                               const int32_t* __restrict__ spos,
                               const int32_t* __restrict__ slen, uint32_t* __restrict__ spair) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nt) return;
  const int64_t k = sival[j];
  if (k >= ni) return;  // a cell's interval (k_dense_scatter)
  const int64_t b = spos[j];
  const Frame& d = fr[image_of_pair(fr, F, b)];
  if (!d.exp1) return;
  const int64_t g = istart[k];
  const int32_t node = sikey[j];
  for (int32_t q = 0; q < slen[j]; ++q)
    expand_one(d, b - d.pair_base + q, (uint32_t)(g - d.pair_base + q), node, spair);
}

// ---- 3b. loop 2 placed by cell -------------------------------------------------------
// When the cube depth leaves few cells (kids + kids^2 + ... + kids^depth at most
// kDenseCellsMax: depth <= 4 without alpha, <= 3 with), loop 2's pairs skip the
// node hash and the interval sort.  A cell (level l, the top l bits of each
// channel byte) is a node, and its pairs in run order are a counting sort of
// loop 2's runs by cell:
//   k_dense_count    one workgroup per tile of runs counts every cell's pairs
//                    and notes each cell's first run;
//   k_dense_scan     an exclusive scan over tiles: each tile's offset in a cell;
//   k_dense_insert   each cell enters the node hash with its first pair, so
//                    node numbering is unchanged;
//   k_dense_virtual  one interval per cell stands for its pairs in the interval
//                    sort, after the cell's loop-1 intervals;
//   k_dense_scatter  a second pass over the tiles writes every pair's term and
//                    leaf value at its sorted position (its pair index only
//                    when the exact root chain needs it: a late pass).
// Inside a tile a wavefront takes 64 consecutive runs at a time and ranks each
// lane among the lanes with its cell (a lane mask refined by one ballot per key
// bit and level); the lowest such lane then advances the cell's count, so a
// cell's pairs keep run order.  The folds find a node's first loop-2 pair from
// l2s rather than from pair indices.
constexpr int kDenseCellsMax = 4680;  // 8 + 64 + 512 + 4096

// Workgroups are dealt to the 8 XCDs in turn; this numbering gives each XCD a
// contiguous range of tiles, so the segments that consecutive tiles extend are
// filled in one L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t q = nb / 8, rem = nb % 8, x = b % 8;
  return x * q + (x < rem ? x : rem) + b / 8;
}

// Four wavefronts per tile, each a quarter of its runs (held in registers);
// their counts are summed at the end, the first runs kept by an LDS minimum.
// LDS: dense_count_lds.
__host__ __device__ constexpr int64_t dense_count_lds(int64_t dcells) {
  return 2 * 4 * dcells + 4 * dcells;
}

__global__ __launch_bounds__(256) void k_dense_count(const Frame* __restrict__ fr, Map m) {
  extern __shared__ int32_t s_dyn[];
  const TileAt tl = map_tile(m, xcd_block(blockIdx.x, gridDim.x));
  const Frame& d = fr[tl.f];
  const int dcells = d.dcells;
  uint32_t* s_first = (uint32_t*)s_dyn;             // dcells
  uint16_t* s_cnt = (uint16_t*)(s_dyn + dcells);    // 4 x dcells
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int c = threadIdx.x; c < 4 * dcells; c += 256) s_cnt[c] = 0;
  for (int c = threadIdx.x; c < dcells; c += 256) s_first[c] = kEmpty32;
  __syncthreads();
  const int64_t tile = (int64_t)1 << d.dtshift;
  const int64_t r0 = d.R1 + tl.t * tile, r1 = min(d.R, r0 + tile);
  const int64_t wr0 = r0 + wave * (tile / 4), wr1 = min(r1, wr0 + tile / 4);
  const int depth = d.depth, kids = d.kids, nch = d.assoc ? 4 : 3;
  const uint64_t below = (1ull << lane) - 1;
  uint16_t* cw = s_cnt + wave * dcells;
  constexpr int kMaxRounds = 32;  // a quarter of the largest tile (8192) over 64 lanes
  uint32_t pk[kMaxRounds];
#pragma unroll
  for (int k = 0; k < kMaxRounds; ++k) {
    const int64_t r = wr0 + 64 * k + lane;
    pk[k] = r < wr1 ? d.pack[r] : 0u;
  }
#pragma unroll
  for (int k = 0; k < kMaxRounds; ++k) {
    const int64_t rb = wr0 + 64 * k;
    if (rb >= wr1) break;
    const int64_t r = rb + lane;
    const bool valid = r < wr1;
    uint64_t msk = __ballot(valid);
    int off = 0, pw = 1;
    for (int l = 1; l <= depth; ++l) {
      uint32_t cell = 0;
      for (int c = 0; c < nch; ++c) {
        const uint32_t top = ((pk[k] >> (8 * c)) & 0xFFu) >> (8 - l);
        const uint64_t v = __ballot((top & 1u) != 0);
        msk &= (top & 1u) ? v : ~v;
        cell |= top << (l * c);
      }
      if (valid && (msk & below) == 0) {  // the lowest lane of its cell
        const int q = off + (int)cell;
        const int n = cw[q];
        if (n == 0) atomicMin(&s_first[q], (uint32_t)r);
        cw[q] = (uint16_t)(n + __popcll(msk));
      }
      pw *= kids;
      off += pw;
    }
  }
  __syncthreads();
  int32_t* tab = d.dtab + tl.t * (int64_t)dcells;
  for (int c = threadIdx.x; c < dcells; c += 256) {
    const int32_t n = s_cnt[c] + s_cnt[dcells + c] + s_cnt[2 * dcells + c] + s_cnt[3 * dcells + c];
    tab[c] = n;
    if (n) atomicMin(&d.dfirst[c], s_first[c]);
  }
}

// One thread per cell: the tiles' counts become their exclusive prefix, and
// dtot the cell's loop-2 pairs.
__global__ void k_dense_scan(const Frame* __restrict__ fr, Map m) {
  const At at = map_at(m, 256);
  const Frame& d = fr[at.f];
  const int64_t c = at.i;
  if (c >= d.dcells) return;
  int32_t* p = d.dtab + c;
  const int64_t st = d.dcells, nt = d.dtiles;
  int32_t acc = 0;
  int64_t t = 0;
  for (; t + 8 <= nt; t += 8) {
    int32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = p[(t + k) * st];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      p[(t + k) * st] = acc;
      acc += v[k];
    }
  }
  for (; t < nt; ++t) {
    const int32_t v = p[t * st];
    p[t * st] = acc;
    acc += v;
  }
  d.dtot[c] = acc;
}

// Level of cell c (cells are numbered level by level).
__device__ __forceinline__ int dense_level(int64_t c, int kids, int* off) {
  int l = 1;
  int64_t o = 0, pw = kids;
  while (c >= o + pw) {
    o += pw;
    pw *= kids;
    ++l;
  }
  *off = (int)o;
  return l;
}

// Each cell with pairs enters the node hash (k_node_insert's table and
// protocol) with the run-major index of its first pair; a cell that loop 1
// already created keeps loop 1's first pair.
__global__ void k_dense_insert(const Frame* __restrict__ fr, Map m) {
  const At at = map_at(m, 256);
  const Frame& d = fr[at.f];
  const int64_t c = at.i;
  if (c >= d.dcells || !d.ok) return;
  const uint32_t r = d.dfirst[c];
  if (r == kEmpty32) return;
  int off;
  const int l = dense_level(c, d.kids, &off);
  const uint64_t key = node_key(d.pack[r], l, d.assoc != 0);
  const uint32_t first = pair_rmajor(r, l, d.R1, d.depth);
  uint64_t* hkeys = d.nkeys;
  uint32_t* hfirst = d.nfirst;
  const uint32_t cap = d.ncap;
  uint32_t slot = hash32(key) & (cap - 1);
  for (uint32_t probe = 0;; ++probe) {
    if (probe >= cap) {  // table full: report, never spin
      d.info[6] = 4;
      d.dslot[c] = 0;
      return;
    }
    const uint64_t cur = hkeys[slot];
    if (cur == key) break;
    if (cur == kEmpty64) {
      const unsigned long long prev =
          atomicCAS((unsigned long long*)&hkeys[slot], (unsigned long long)kEmpty64,
                    (unsigned long long)key);
      if (prev == kEmpty64 || prev == key) break;
    }
    slot = (slot + 1) & (cap - 1);
  }
  atomicMin(&hfirst[slot], first);
  d.dslot[c] = slot;
}

// Interval ni + dbase + c stands for cell c's loop-2 pairs: no first pair of
// its own (kVirtual), its node as the sort key -- the empty ones sort after
// every node -- and, since the sort is stable and these follow k_compact's
// intervals, it lands after the node's loop-1 intervals.
__global__ void k_dense_virtual(const Frame* __restrict__ fr, Map m, int64_t ni,
                                int32_t sum_nodes, uint32_t* __restrict__ istart,
                                int32_t* __restrict__ ikey, uint32_t* __restrict__ ival,
                                int32_t* __restrict__ vlen) {
  const At at = map_at(m, 256);
  const Frame& d = fr[at.f];
  const int64_t c = at.i;
  if (c >= d.dcells) return;
  const int64_t v = d.dbase + c, k = ni + v;
  const int32_t n = d.dtot[c];
  istart[k] = kVirtual;
  ival[k] = (uint32_t)k;
  vlen[v] = n;
  ikey[k] = n == 0 ? sum_nodes
                   : (int32_t)d.node_base + (d.ok ? d.slot_node[d.dslot[c]] : 0);
}

// Exclusive prefix of v over a 256-thread workgroup (s_part: 4 ints).
__device__ __forceinline__ int block_excl_256(int v, int* s_part) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) s_part[wave] = incl;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; ++w) before += s_part[w];
  __syncthreads();
  return before + incl - v;
}

// The cell of a run at level l and the lanes of the wavefront whose run has
// the same cell (msk: the candidate lanes on entry).
__device__ __forceinline__ uint32_t dense_cell(uint32_t pk, int l, int nch, uint64_t* msk) {
  uint64_t mk = *msk;
  uint32_t cell = 0;
  for (int k = 1; k <= l; ++k) {
    for (int c = 0; c < nch; ++c) {
      const uint32_t top = ((pk >> (8 * c)) & 0xFFu) >> (8 - k);
      const uint64_t v = __ballot((top & 1u) != 0);
      mk &= (top & 1u) ? v : ~v;
      if (k == l) cell |= top << (l * c);
    }
  }
  *msk = mk;
  return cell;
}

// Every loop-2 pair of a tile at its sorted position.  The tile's runs are
// first put in each level's order inside the tile (cells ascending, runs
// ascending within a cell): each wavefront counts the cells of its quarter of
// the tile, a scan over cells gives every cell's first slot and each
// wavefront's first slot in it, and a second walk places the runs into a
// slot -> run map.  The first walk keeps each run's rank among the lanes of its
// round with its cell, and whether it is the last of them, so the second walk
// needs no ballots.  The pairs are then written slot by slot, so consecutive
// threads write consecutive positions of one cell's segment (the cell's first
// global position for this tile minus its first slot, plus the slot).  The
// values are expand_one's (the cube midpoints in closed form, equal to the
// classification's level-by-level sums); a run's count is written for leaf
// pairs only (the only ones whose count is read).
// Levels are taken one at a time, so the LDS holds the counters of one
// level's cells, one map and one rank byte per run: dense_scatter_lds.
constexpr int kScatterThreads = 256;

__host__ __device__ constexpr int64_t dense_scatter_lds(int64_t tile, int64_t cmax) {
  return 4 * cmax + 2 * 4 * cmax + 2 * tile + tile;
}

// SP: the pair indices only, for an image whose root chain is summed serially
// (k_fold_record reads them); the first pass leaves them out.
template <bool SP>
__global__ __launch_bounds__(kScatterThreads) void k_dense_scatter(const Frame* __restrict__ fr,
                                                                   Map m,
                                                                   uint32_t* __restrict__ spair_all) {
  extern __shared__ int32_t s_dyn[];
  __shared__ int s_part[4];
  const TileAt tl = map_tile(m, xcd_block(blockIdx.x, gridDim.x));
  const Frame& d = fr[tl.f];
  if (SP && (d.info[6] != 0 || !d.info[14])) return;
  const int depth = d.depth, kids = d.kids, assoc = d.assoc, nch = assoc ? 4 : 3;
  int cmax = 1;
  for (int l = 1; l <= depth; ++l) cmax *= kids;
  const int dcells = d.dcells;
  const int64_t tile = (int64_t)1 << d.dtshift;
  const int64_t R1 = d.R1, n2 = d.R - d.R1;
  const int64_t r0 = R1 + tl.t * tile, r1 = min(d.R, r0 + tile);
  const int n = (int)(r1 - r0);
  const int ncell = cmax;                             // counters of one level
  int32_t* s_delta = s_dyn;                           // ncell
  uint16_t* s_cnt = (uint16_t*)(s_dyn + ncell);       // 4 x ncell
  uint16_t* s_run = s_cnt + 4 * ncell;                // tile
  uint8_t* s_rk = (uint8_t*)(s_run + tile);           // tile: rank in round | last << 6
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t quarter = tile / 4;
  const int64_t wr0 = r0 + wave * quarter, wr1 = min(r1, wr0 + quarter);
  const uint64_t below = (1ull << lane) - 1;
  const uint32_t* pack = d.pack;
  const int32_t* tab = d.dtab + tl.t * (int64_t)dcells;
  const uint32_t* dpos = d.dpos + d.dbase;
  const uint32_t l1 = (uint32_t)(8 * R1);
  const Terms T = d.T;
  uint32_t* spair = spair_all + d.pair_base;
  uint16_t* cw = s_cnt + wave * ncell;
  // This wavefront's runs, loaded once: run wr0 + 64 k + lane in pk[k].
  constexpr int kMaxRounds = 32;  // a quarter of the largest tile (8192) over 64 lanes
  uint32_t pk[kMaxRounds];
#pragma unroll
  for (int k = 0; k < kMaxRounds; ++k) {
    const int64_t r = wr0 + 64 * k + lane;
    pk[k] = r < wr1 ? pack[r] : 0u;
  }

  // Counters of cells [0, nc) (flat offset foff): per wavefront, then first
  // slots.
  auto starts = [&](int nc, int foff) {
    const int per = (nc + kScatterThreads - 1) / kScatterThreads;
    const int cb = min(nc, tid * per), ce = min(nc, cb + per);
    int sum = 0;
    for (int c = cb; c < ce; ++c)
      sum += s_cnt[c] + s_cnt[ncell + c] + s_cnt[2 * ncell + c] + s_cnt[3 * ncell + c];
    int run = block_excl_256(sum, s_part);
    for (int c = cb; c < ce; ++c) {
      int at = run;
      for (int w = 0; w < 4; ++w) {
        const int v = s_cnt[w * ncell + c];
        s_cnt[w * ncell + c] = (uint16_t)at;
        at += v;
      }
      if (at > run) {
        s_delta[c] = (int32_t)((int64_t)dpos[foff + c] - d.pair_base) + tab[foff + c] - run;
      }
      run = at;
    }
  };
  // The pairs of level l in slot order from the map.
  auto emit = [&](int l, const uint16_t* map) {
    for (int i = tid; i < n; i += kScatterThreads) {
      const int64_t r = r0 + map[i];
      if (SP) {
        uint32_t cell = 0;
        const uint32_t pk = pack[r];
        for (int c = 0; c < nch; ++c) cell |= (((pk >> (8 * c)) & 0xFFu) >> (8 - l)) << (l * c);
        spair[s_delta[cell] + i] = l1 + (uint32_t)((int64_t)(l - 1) * n2 + (r - R1));
        continue;
      }
      const uint32_t pk = pack[r];
      const float4 f = d.px_pack ? make_float4((float)(pk & 0xFFu) * 257.0f,
                                               (float)((pk >> 8) & 0xFFu) * 257.0f,
                                               (float)((pk >> 16) & 0xFFu) * 257.0f, 65535.0f)
                                 : d.px[r];
      double a[4];
      assoc_pixel(assoc != 0, f.x, f.y, f.z, f.w, a);
      // The cube midpoint, 32767.5 +- 2^14 +- ... +- 2^(15-l) by the channel's
      // top l bits B, is (B + 1/2) 2^(16-l) - 1/2: every partial sum of the
      // reference's loop is a half-integer below 2^17, so both are exact.
      const double step = (double)(1 << (16 - l));
      double mid[4];
      uint32_t cell = 0;
      for (int c = 0; c < 4; ++c) {
        const uint32_t top = c < nch ? ((pk >> (8 * c)) & 0xFFu) >> (8 - l) : 0u;
        cell |= c < nch ? top << (l * c) : 0u;
        mid[c] = ((double)top + 0.5) * step - 0.5;
      }
      const int32_t cnt = d.count[r];
      const int64_t li = s_delta[cell] + i;
      const double e0 = kScale * (a[0] - mid[0]);
      const double e1 = kScale * (a[1] - mid[1]);
      const double e2 = kScale * (a[2] - mid[2]);
      const double e3 = assoc ? kScale * (a[3] - mid[3]) : 0.0;
      double dd = e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3;
      if (dd != dd) dd = 0.0;
      T.term[li] = (double)cnt * sqrt(dd);
      if (l == depth) {
        T.lcnt[li] = cnt;
        T.lval[li] = make_float4(clamp_pixel(a[0]), clamp_pixel(a[1]), clamp_pixel(a[2]),
                                 clamp_pixel(assoc ? a[3] : kRange));
      }
    }
  };

  int off = 0, C = 1;
  for (int l = 1; l <= depth; ++l) {  // one level at a time
    C *= kids;
    for (int c = tid; c < 4 * ncell; c += kScatterThreads) s_cnt[c] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kMaxRounds; ++k) {  // count this quarter's cells
      const int64_t rb = wr0 + 64 * k;
      if (rb >= wr1) break;
      const int64_t r = rb + lane;
      const bool valid = r < wr1;
      uint64_t msk = __ballot(valid);
      const uint32_t cell = dense_cell(pk[k], l, nch, &msk);
      if (valid) {
        const uint64_t lo = msk & below;
        const bool last = (msk >> lane) == 1ull;
        s_rk[r - r0] = (uint8_t)(__popcll(lo) | (last ? 64 : 0));
        if (lo == 0) cw[cell] += (uint16_t)__popcll(msk);
      }
    }
    __syncthreads();
    starts(C, off);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kMaxRounds; ++k) {  // the slot -> run map
      const int64_t rb = wr0 + 64 * k;
      if (rb >= wr1) break;
      const int64_t r = rb + lane;
      if (r < wr1) {
        uint32_t cell = 0;
        for (int c = 0; c < nch; ++c)
          cell |= (((pk[k] >> (8 * c)) & 0xFFu) >> (8 - l)) << (l * c);
        const int rk = s_rk[r - r0];
        const int slot = cw[cell] + (rk & 63);
        s_run[slot] = (uint16_t)(r - r0);
        if (rk & 64) cw[cell] = (uint16_t)(slot + 1);  // the round's last of its cell
      }
    }
    __syncthreads();
    emit(l, s_run);
    __syncthreads();
    off += C;
  }
}

// The root of every image: key 0, its own parent, level 0, zero sums.
__global__ void k_root_init(const Frame* __restrict__ fr) {
  const Frame& d = fr[blockIdx.x];
  if (!d.ok || threadIdx.x != 0) return;
  const TreeDev& t = d.t;
  t.key[0] = 0;
  t.parent[0] = 0;
  t.id[0] = 0;
  t.level[0] = 0;
  t.qe[0] = 0.0;
  t.share[0] = 0.0;
  for (int c = 0; c < 4; ++c) t.total[c] = 0.0;
  t.nu[0] = 0;
}

__global__ void k_links(const Frame* __restrict__ fr, Map m) {
  const At at = map_at(m, 256);
  const Frame& d = fr[at.f];
  const int64_t k = at.i + 1;
  if (k > d.M || !d.ok) return;
  const uint64_t* hkeys = d.nkeys;
  const int32_t* slot_node = d.slot_node;
  const uint32_t cap = d.ncap;
  const TreeDev& t = d.t;
  int32_t* info = d.info;
  const uint64_t key = t.key[k];
  const int level = (int)(key >> 40);
  int32_t par = 0;
  if (level > 1) {
    const uint64_t pk = parent_key(key);
    uint32_t slot = hash32(pk) & (cap - 1);
    uint32_t probe = 0;
    while (hkeys[slot] != pk && probe < cap) {
      slot = (slot + 1) & (cap - 1);
      ++probe;
    }
    if (probe >= cap) {  // a parent that was never inserted: report
      info[6] = 4;
      return;
    }
    par = slot_node[slot];
  }
  const int id = child_index(key);
  t.parent[k] = par;
  t.id[k] = id;
  t.level[k] = level;
  t.child[16 * par + id] = (int32_t)k;
}

// ---- 4. ordered sums --------------------------------------------------------------
// Each node's quantize_error is a serial chain over its pairs in run order, and
// the leaves' total_color four more chains.  The chains cannot be reassociated,
// so the parallel work is everything around them: one thread per sorted pair
// computes its term (distance x count) and the leaf values, written contiguously
// in node order; then each chain is a plain in-order fold over contiguous
// memory -- one thread per node for short segments, one wavefront per node for
// long ones, staging chunks through LDS while the next chunks load.

// First sorted position in [b, e) whose pair belongs to loop 2 (pairs of a
// node are in ascending pair order, loop 1's first).
__device__ __forceinline__ int32_t loop2_start(const uint32_t* spair, int32_t b, int32_t e,
                                               uint32_t l1) {
  int32_t lo = b, hi = e;
  while (lo < hi) {
    const int32_t m = lo + ((hi - lo) >> 1);
    if (spair[m] < l1)
      lo = m + 1;
    else
      hi = m;
  }
  return lo;
}

// The same for node n of an image: one placed by cell knows it per node (its
// loop-2 pairs carry no pair index unless the exact root chain needs one).
__device__ __forceinline__ int32_t seg_loop2(const Frame& d, int64_t n, int32_t b, int32_t e) {
  return d.dense ? min(d.l2s[n], e) : loop2_start(d.spair, b, e, (uint32_t)(8 * d.R1));
}

// One leaf term of total_color, as the reference forms it: count * scale * value.
__device__ __forceinline__ double tot_term(int32_t cnt, float v) {
  return (double)cnt * kScale * (double)v;
}

constexpr int kShortSeg = 256;   // longer: one wavefront (k_fold_mid)
constexpr int kMidSeg = 8192;    // longer: one workgroup, in parallel (k_fold_long)
// Loop 2's four total_color chains: one wavefront folds them together, which
// keeps up with a workgroup's parallel fold far longer.
constexpr int kMidSegTot = 65536;
constexpr int kHugeSeg = 65536;  // longer: queued ahead of the other long segments

// L2 = false: quantize_error (+ the node's share of the root chain) over all
// pairs, and loop 1's leaf totals.  L2 = true: loop 2's leaf totals, folded onto
// what PruneToCubeDepth left.  Segments longer than kShortSeg are queued for
// k_fold_long.
template <bool L2>
__global__ void k_fold_short(const Frame* __restrict__ fr, Map m, Batch B) {
  const At at = map_at(m, 64);
  const Frame& d = fr[at.f];
  const int64_t n = at.i + 1;
  if (n > d.M || !d.ok) return;
  const TreeDev& t = d.t;
  const Terms& T = d.T;
  const int depth = d.depth;
  const int32_t b = d.seg_b[n], e = d.seg_e[n];
  const int level = t.level[n];
  if (L2 && level != depth) return;
  if (e - b > kShortSeg) {
    const uint64_t ent = ((uint64_t)at.f << 32) | (uint64_t)n;
    if (e - b > kHugeSeg)
      B.huge_list[atomicAdd(&B.ctl[L2 ? 6 : 5], 1)] = ent;
    else if (e - b > (L2 ? kMidSegTot : kMidSeg))
      B.long_list[atomicAdd(&B.ctl[L2 ? 4 : 3], 1)] = ent;
    else
      B.mid_list[atomicAdd(&B.ctl[L2 ? 2 : 1], 1)] = ent;
    return;
  }
  const int32_t s = seg_loop2(d, n, b, e);
  if (!L2) {
    double qe = 0.0, share = 0.0;
    for (int32_t i = b; i < e; ++i) {
      qe += T.term[i];
      share += qe;
    }
    t.qe[n] = qe;
    t.share[n] = share;
    double tot[4] = {0.0, 0.0, 0.0, 0.0};
    uint64_t nu = 0;
    if (level == 8) {
      for (int32_t i = b; i < s; ++i) {
        const float4 v = T.lval[i];
        const int32_t c = T.lcnt[i];
        tot[0] += tot_term(c, v.x);
        tot[1] += tot_term(c, v.y);
        tot[2] += tot_term(c, v.z);
        tot[3] += tot_term(c, v.w);
        nu += (uint64_t)c;
      }
    }
    for (int c = 0; c < 4; ++c) t.total[4 * n + c] = tot[c];
    t.nu[n] = nu;
  } else {
    if (s == e) return;
    double tot[4];
    for (int c = 0; c < 4; ++c) tot[c] = t.total[4 * n + c];
    uint64_t nu = t.nu[n];
    for (int32_t i = s; i < e; ++i) {
      const float4 v = T.lval[i];
      const int32_t c = T.lcnt[i];
      tot[0] += tot_term(c, v.x);
      tot[1] += tot_term(c, v.y);
      tot[2] += tot_term(c, v.z);
      tot[3] += tot_term(c, v.w);
      nu += (uint64_t)c;
    }
    for (int c = 0; c < 4; ++c) t.total[4 * n + c] = tot[c];
    t.nu[n] = nu;
  }
}

constexpr int kChunk = 256;           // terms staged per step (a smaller stage keeps more
                                      // wavefronts resident: 512 was 0.3 ms slower on C3)
constexpr int kPerLane = kChunk / 64;

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// quantize_error over [b, e): lane 0 folds each staged chunk in order while the
// wave loads the chunk after next.  The share (sum of the running values) is
// only needed approximately: sum_i term_i * (e - i), reduced in any order.
__device__ void fold_qe_wave(const double* term, int32_t b, int32_t e, double* s_q, double* qe_out,
                             double* share_out) {
  const int lane = threadIdx.x;
  const int32_t len = e - b;
  const int32_t nch = (len + kChunk - 1) / kChunk;
  double ra[kPerLane], rb[kPerLane];
  auto load = [&](double* r, int32_t c) {
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
      const int32_t i = b + c * kChunk + k * 64 + lane;
      r[k] = (c < nch && i < e) ? term[i] : 0.0;
    }
  };
  load(ra, 0);
  load(rb, 1);
  double qe = 0.0, sh = 0.0;
  auto step = [&](double* r, int32_t c) {
    double* buf = s_q + (c & 1) * kChunk;
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
      const int32_t i = b + c * kChunk + k * 64 + lane;
      buf[k * 64 + lane] = r[k];
      sh += r[k] * (double)(e - i);
    }
    load(r, c + 2);
    __syncthreads();
    if (lane == 0) {
      const int32_t cnt = min(kChunk, len - c * kChunk);
      if (cnt == kChunk) {
#pragma unroll 32
        for (int j = 0; j < kChunk; ++j) qe += buf[j];
      } else {
        for (int j = 0; j < cnt; ++j) qe += buf[j];
      }
    }
  };
  for (int32_t c = 0; c < nch; c += 2) {
    step(ra, c);
    if (c + 1 < nch) step(rb, c + 1);
  }
  sh = wave_sum(sh);
  qe = __shfl(qe, 0, 64);
  *qe_out = qe;
  *share_out = sh;
}

// total_color over [b, e) onto tot[] (lanes 0..3 fold one channel each);
// number_unique is an integer sum, reduced across the wave.
__device__ void fold_tot_wave(const Terms& T, int32_t b, int32_t e, double* s_p, double tot[4],
                              uint64_t* nu) {
  const int lane = threadIdx.x;
  const int32_t len = e - b;
  const int32_t nch = (len + kChunk - 1) / kChunk;
  float4 va[kPerLane], vb[kPerLane];
  int32_t ca[kPerLane], cb[kPerLane];
  auto load = [&](float4* v, int32_t* cn, int32_t c) {
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
      const int32_t i = b + c * kChunk + k * 64 + lane;
      const bool ok = c < nch && i < e;
      v[k] = ok ? T.lval[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      cn[k] = ok ? T.lcnt[i] : 0;
    }
  };
  load(va, ca, 0);
  load(vb, cb, 1);
  const int ch = lane & 3;
  double acc = tot[ch];
  uint64_t nsum = 0;
  auto step = [&](float4* v, int32_t* cn, int32_t c) {
    double* buf = s_p + (c & 1) * 4 * kChunk;
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
      const int j = k * 64 + lane;
      buf[0 * kChunk + j] = tot_term(cn[k], v[k].x);
      buf[1 * kChunk + j] = tot_term(cn[k], v[k].y);
      buf[2 * kChunk + j] = tot_term(cn[k], v[k].z);
      buf[3 * kChunk + j] = tot_term(cn[k], v[k].w);
      nsum += (uint64_t)cn[k];
    }
    load(v, cn, c + 2);
    __syncthreads();
    if (lane < 4) {
      const double* src = buf + ch * kChunk;
      const int32_t cnt = min(kChunk, len - c * kChunk);
      if (cnt == kChunk) {
#pragma unroll 32
        for (int j = 0; j < kChunk; ++j) acc += src[j];
      } else {
        for (int j = 0; j < cnt; ++j) acc += src[j];
      }
    }
  };
  for (int32_t c = 0; c < nch; c += 2) {
    step(va, ca, c);
    if (c + 1 < nch) step(vb, cb, c + 1);
  }
  nsum = wave_sum_u64(nsum);
  for (int c = 0; c < 4; ++c) tot[c] = __shfl(acc, c, 64);
  *nu += nsum;
}

// Medium segments: one wavefront each, the chain folded serially; taken from a
// queue, so the long ones do not gather on a few workgroups.
template <bool L2>
__global__ __launch_bounds__(64) void k_fold_mid(const Frame* __restrict__ fr, Batch B) {
  __shared__ double s_buf[2 * 4 * kChunk];
  __shared__ int32_t s_item;
  const int32_t items = B.ctl[L2 ? 2 : 1];
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(&B.ctl[L2 ? 14 : 13], 1);
    __syncthreads();
    const int32_t it = s_item;
    __syncthreads();
    if (it >= items) break;
    const uint64_t ent = B.mid_list[it];
    const Frame& d = fr[(int32_t)(ent >> 32)];
    const int32_t n = (int32_t)(ent & 0xFFFFFFFFu);
    const TreeDev& t = d.t;
    const Terms& T = d.T;
    const int32_t b = d.seg_b[n], e = d.seg_e[n];
    const int level = t.level[n];
    const int32_t s = seg_loop2(d, n, b, e);
    if (!L2) {
      double qe, share;
      fold_qe_wave(T.term, b, e, s_buf, &qe, &share);
      double tot[4] = {0.0, 0.0, 0.0, 0.0};
      uint64_t nu = 0;
      __syncthreads();
      if (level == 8 && s > b) fold_tot_wave(T, b, s, s_buf, tot, &nu);
      if (threadIdx.x == 0) {
        t.qe[n] = qe;
        t.share[n] = share;
        for (int c = 0; c < 4; ++c) t.total[4 * n + c] = tot[c];
        t.nu[n] = nu;
      }
    } else if (s < e) {
      double tot[4];
      for (int c = 0; c < 4; ++c) tot[c] = t.total[4 * n + c];
      uint64_t nu = t.nu[n];
      fold_tot_wave(T, s, e, s_buf, tot, &nu);
      if (threadIdx.x == 0) {
        for (int c = 0; c < 4; ++c) t.total[4 * n + c] = tot[c];
        t.nu[n] = nu;
      }
    }
    __syncthreads();
  }
}

// ---- exact serial sums, in parallel --------------------------------------------------
// Every chain folded here adds finite non-negative terms, so its running value
// only grows.  While it stays in one binade [2^x, 2^(x+1)) every double there is
// a multiple of u = 2^(x-52), so s + t rounds to s + rn(t / u) u: an integer
// addition of the term rounded on its own.  Two kinds of step are not of that
// form: a tie (t / u exactly halfway, where round-half-even looks at s) and the
// step whose exact sum leaves the binade (rounded on the next binade's grid).
// A workgroup takes the chain a window at a time: each thread rounds its terms
// on the current grid, a block scan gives every step's running value, the
// first special step is found, everything before it is committed exactly, and
// that step is done as the plain IEEE addition it is.  The result equals the
// serial loop bit for bit; the serial work left is one addition per binade
// crossing or tie, and a short serial prologue where crossings are dense.
// Workgroups of 256 threads with 4 terms each per window (1024-term windows):
// smaller and more numerous than 512 x 14, which ran the 32 x 1080p folds
// 0.8 ms slower -- half the windows hold a binade crossing, and a small
// window's resolution rounds are cheap.
constexpr int kFoldThreads = 256;
constexpr int kFoldWaves = kFoldThreads / 64;
constexpr int kFoldPrologue = 512;                 // terms added serially first
constexpr int kFoldRounds = 8;                     // special steps resolved per window in
                                                   // parallel; the rest of it serially
constexpr double kFoldSat = 0x1p54;                // saturation of the integer scan

template <int NW>  // wavefronts of the workgroup
struct FoldSharedT {
  double part[2][NW][4];                           // scan partials (alternating passes)
  int32_t first[2][NW][4];                         // min partials
  double dpart[NW];
  uint64_t upart[NW];
  double Sk[4];
  double tk[4];
  double ser[4];
};
using FoldShared = FoldSharedT<kFoldWaves>;

// Wavefront sums through DPP: quad permutes and row rotates give every lane
// its 16-lane row's sum in four steps; two shuffles join the rows.
template <int Ctrl>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, Ctrl, 0xF, 0xF, false);
}
template <int Ctrl>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i32<Ctrl>((int)(b & 0xFFFFFFFF));
  const int hi = dpp_i32<Ctrl>((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
constexpr int kDppXor1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppRor4 = 0x124;  // row_ror:4
constexpr int kDppRor8 = 0x128;  // row_ror:8

__device__ __forceinline__ double wave_total(double v) {  // every lane: the wavefront's sum
  v += dpp_f64<kDppXor1>(v);
  v += dpp_f64<kDppXor2>(v);
  v += dpp_f64<kDppRor4>(v);
  v += dpp_f64<kDppRor8>(v);
  v += __shfl_xor(v, 16, 64);
  return v + __shfl_xor(v, 32, 64);
}

__device__ __forceinline__ int wave_or(int v) {
  v |= dpp_i32<kDppXor1>(v);
  v |= dpp_i32<kDppXor2>(v);
  v |= dpp_i32<kDppRor4>(v);
  v |= dpp_i32<kDppRor8>(v);
  v |= __shfl_xor(v, 16, 64);
  return v | __shfl_xor(v, 32, 64);
}

template <int NW>
__device__ double fold_sum_d(double v, FoldSharedT<NW>& sh) {  // any order (approximate sums)
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh.dpart[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = 0.0;
  for (int k = 0; k < NW; ++k) m += sh.dpart[k];
  __syncthreads();
  return m;
}

template <int NW>
__device__ uint64_t fold_sum_u64(uint64_t v, FoldSharedT<NW>& sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) sh.upart[threadIdx.x >> 6] = v;
  __syncthreads();
  uint64_t m = 0;
  for (int k = 0; k < NW; ++k) m += sh.upart[k];
  __syncthreads();
  return m;
}

// Terms of the chains, read through LDS: `load` fetches the raw record of an
// index from global memory (coalesced, by the whole workgroup), `term` forms
// chain c's term from it.
struct QeSrc {  // side sum: the share of the root's chain, sum_i term_i * (e - i)
  using Raw = double;
  const double* term_;
  __device__ Raw load(int32_t i) const { return term_[i]; }
  __device__ double term(const Raw& r, int) const { return r; }
  __device__ double side(const Raw& r, double e_minus_i) const { return r * e_minus_i; }
};

struct TotRaw {
  float4 v;
  int32_t n;
};

struct TotSrc {  // side sum: number_unique (an integer, exact in a double here)
  using Raw = TotRaw;
  const Terms* T;
  __device__ Raw load(int32_t i) const { return TotRaw{T->lval[i], T->lcnt[i]}; }
  __device__ double term(const Raw& r, int c) const {
    const float f = c == 0 ? r.v.x : c == 1 ? r.v.y : c == 2 ? r.v.z : r.v.w;
    return tot_term(r.n, f);
  }
  __device__ double side(const Raw& r, double) const { return (double)r.n; }
};

// Records [from, to) of one window into LDS (to - from <= the stage size),
// thread t's K consecutive records contiguous, one record of padding after
// every K so that threads reading their own run do not collide on banks.
template <int K, class Src>
__device__ __forceinline__ int32_t stage_slot(int32_t p) {
  return p + p / K;
}

template <int K, class Src, int NT = kFoldThreads>
__device__ void stage_in(const Src& src, int32_t from, int32_t to, typename Src::Raw* stage) {
  typename Src::Raw r[K];
  const int32_t n = to - from;  // <= K * NT
#pragma unroll
  for (int j = 0; j < K; ++j) {  // every load in flight before the first store
    const int32_t p = (int32_t)threadIdx.x + j * NT;
    if (p < n) r[j] = src.load(from + p);
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const int32_t p = (int32_t)threadIdx.x + j * NT;
    if (p < n) stage[stage_slot<K, Src>(p)] = r[j];
  }
}

// v + term(p0, c) + ... + term(p1 - 1, c) from the stage, left to right; the
// records are read a batch ahead of the dependent additions.
template <int K, class Src>
__device__ double serial_lds(double v, const Src& src, const typename Src::Raw* stage,
                             int32_t p0, int32_t p1, int c) {
  constexpr int kB = sizeof(typename Src::Raw) > 8 ? 8 : 16;
  int32_t p = p0;
  for (; p + kB <= p1; p += kB) {
    typename Src::Raw r[kB];
#pragma unroll
    for (int k = 0; k < kB; ++k) r[k] = stage[stage_slot<K, Src>(p + k)];
#pragma unroll
    for (int k = 0; k < kB; ++k) v += src.term(r[k], c);
  }
  for (; p < p1; ++p) v += src.term(stage[stage_slot<K, Src>(p)], c);
  return v;
}

// NC chains over the same index range [b, e), each summed left to right in
// double from s[c]: s[c] + term(b, c) + ... + term(e - 1, c).  Every thread of
// the workgroup calls it and gets the results.  The chains share windows
// (each skips indices it has already passed), loads and reductions.  Integers
// on a binade's grid are held in doubles: exact below 2^53, and a sum that
// reaches 2^53 always lies past a special step.  A window without a special
// step (the common case: its total stays inside the binade and no term ties)
// needs one block reduction; only a window holding one is scanned to find it.
// The next window's records are loaded while the current one is summed.
//
// It also returns, summed in any order, the side sum src.side over [b, e)
// (each index counted once: `seen` marks how far it has been taken).
template <int NC, int K, class Src, int NT = kFoldThreads>
__device__ double exact_sums(double* s, int32_t b, int32_t e, const Src& src, FoldSharedT<NT / 64>& sh,
                             typename Src::Raw* stage) {
  using Raw = typename Src::Raw;
  constexpr int32_t kWin = NT * K;
  static_assert(kFoldPrologue <= kWin, "the prologue fits the stage");
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int32_t me = (int32_t)threadIdx.x;
  double side = 0.0;
  const int32_t pro = min(e, b + kFoldPrologue);
  int32_t seen = pro;
  if (pro > b) {  // the first terms serially: binade crossings come densely there
    __syncthreads();
    stage_in<K, Src, NT>(src, b, pro, stage);
    __syncthreads();
    if (me < NC) {
      double v = 0.0;
#pragma unroll
      for (int k = 0; k < NC; ++k)
        if (k == me) v = s[k];
      sh.ser[me] = serial_lds<K>(v, src, stage, 0, pro - b, me);
    }
    for (int32_t p = me; p < pro - b; p += NT)
      side += src.side(stage[stage_slot<K, Src>(p)], (double)(e - (b + p)));
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NC; ++k) s[k] = sh.ser[k];
  }
  int32_t pos[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) pos[c] = pro;
  int par = 0;
  // Window records, two buffers in turn: one summed while the next loads.
  // Thread t holds records W + t + j * NT (coalesced loads; the
  // index is clamped so that every load is in bounds and none waits on a branch).
  Raw bufA[K], bufB[K];
  int32_t wA = -1, wB = -1;
  auto fetch = [&](Raw (&buf)[K], int32_t& wb, int32_t w) {
#pragma unroll
    for (int j = 0; j < K; ++j) buf[j] = src.load(min(w + me + j * NT, e - 1));
    wb = w;
  };
  auto pass = [&](Raw (&cur)[K], int32_t& wc, Raw (&nxt)[K], int32_t& wn) -> bool {
    int32_t W = e;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (!(s[c] <= 0x1.fffffffffffffp1023)) pos[c] = e;  // inf (or NaN) absorbs the rest
      W = min(W, pos[c]);
    }
    if (W >= e) return false;
    const int32_t wend = min(e, W + kWin);
    if (W > seen) {  // serial stretches passed [seen, W): their side sums, in parallel
      for (int32_t i = seen + me; i < W; i += NT)
        side += src.side(src.load(i), (double)(e - i));
      seen = W;
    }
    if (wc != W) fetch(cur, wc, W);
    if (wend < e && wn != wend) fetch(nxt, wn, wend);
    // Each chain's grid: the binade of s (below 2^-1022, the subnormal grid).
    int x[NC];
    double S[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      x[c] = s[c] < 0x1p-1022 ? -1022 : ilogb(s[c]);
      S[c] = ldexp(s[c], 52 - x[c]);
    }
    // The window's total on each grid, and whether any term ties (any order:
    // sums of these integers are exact below 2^53 and stay >= 2^53 above it).
    double total[NC];
    int tie = 0;
    {
      // t / u as two multiplications by powers of two (each representable,
      // the product exact wherever it matters: an intermediate underflow
      // only meets terms that round to 0).
      double sc1[NC], sc2[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int k = 52 - x[c], k1 = k / 2;
        sc1[c] = ldexp(1.0, k1);
        sc2[c] = ldexp(1.0, k - k1);
      }
      double acc[NC][2];
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c][0] = acc[c][1] = 0.0;
      const double ei0 = (double)(e - (W + me));  // e - i, exact integers in a double
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int32_t i = W + me + j * NT;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const bool act = i < e && i >= pos[c];
          const double q = act ? (src.term(cur[j], c) * sc1[c]) * sc2[c] : 0.0;
          const double r = rint(q);
          tie |= fabs(q - r) == 0.5 ? 1 << c : 0;  // inf or NaN: the total fails below
          acc[c][j & 1] += r;
        }
        if (i < wend && i >= seen) side += src.side(cur[j], ei0 - (double)(j * NT));
      }
      seen = max(seen, wend);
      // Integers of at most 2^54 each (or inf / NaN, which fail the test
      // below): their sums are exact below 2^53 and stay >= 2^53 above it.
      double red[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) red[c] = wave_total(acc[c][0] + acc[c][1]);
      const int t = wave_or(tie);
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < NC; ++c) sh.part[par][wv][c] = red[c];
        sh.first[par][wv][0] = t;
      }
    }
    __syncthreads();
    bool clean = true;
    int ties = 0;
    for (int k = 0; k < (NT / 64); ++k) ties |= sh.first[par][k][0];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double tot = 0.0;
      for (int k = 0; k < (NT / 64); ++k) tot += sh.part[par][k][c];
      total[c] = fmin(tot, kFoldSat);
      if ((ties >> c) & 1 || !(S[c] + tot < 0x1p53 - 1.0)) clean = false;
    }
    par ^= 1;
    if (clean) {  // every step of the window stays on its binade's grid
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (pos[c] < wend) {
          s[c] = ldexp(S[c] + total[c], x[c] - 52);
          pos[c] = wend;
        }
      return true;
    }
    // A special step in this window: resolve the window from LDS, one
    // special step per round (scan for every step's running value, find each
    // chain's first special step, commit everything before it, take the step
    // as the IEEE addition it is, go on from there on the new grid).  The next
    // window is then the one already loading.
#pragma unroll
    for (int j = 0; j < K; ++j) stage[stage_slot<K, Src>(me + j * NT)] = cur[j];
    __syncthreads();
    const int32_t base = W + me * K;
    for (int round = 0;; ++round) {
      bool live = false;
#pragma unroll
      for (int c = 0; c < NC; ++c) live |= pos[c] < wend;
      if (!live) break;
      if (round == kFoldRounds) {  // specials come densely (ties): the rest serially
        if (me < NC) {
          double v = 0.0;
          int32_t from = wend;
#pragma unroll
          for (int k = 0; k < NC; ++k)
            if (k == me) {
              v = s[k];
              from = pos[k];
            }
          sh.ser[me] = serial_lds<K>(v, src, stage, from - W, wend - W, me);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (pos[c] < wend) {
            s[c] = sh.ser[c];
            pos[c] = wend;
          }
        __syncthreads();
        break;
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        x[c] = s[c] < 0x1p-1022 ? -1022 : ilogb(s[c]);
        S[c] = ldexp(s[c], 52 - x[c]);
      }
      double q[NC][K];
      double L[NC];
      int mytie = 0;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int32_t i = base + j;
        const Raw r = stage[stage_slot<K, Src>(me * K + j)];  // past wend: masked
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const bool act = i < wend && i >= pos[c];
          q[c][j] = act ? ldexp(src.term(r, c), 52 - x[c]) : 0.0;
          mytie |= (q[c][j] - floor(q[c][j]) == 0.5) ? 1 << c : 0;
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double a0 = 0.0, a1 = 0.0;
#pragma unroll
        for (int j = 0; j < K; j += 2) {
          a0 += rint(fmin(q[c][j], kFoldSat));
          if (j + 1 < K) a1 += rint(fmin(q[c][j + 1], kFoldSat));
        }
        L[c] = fmin(a0 + a1, kFoldSat);
      }
      double excl[NC];
      {
        double incl[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) incl[c] = L[c];
        for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const double t = __shfl_up(incl[c], off, 64);
            if (lane >= off) incl[c] = fmin(incl[c] + t, kFoldSat);
          }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          excl[c] = __shfl_up(incl[c], 1, 64);
          if (lane == 0) excl[c] = 0.0;
          if (lane == 63) sh.part[par][wv][c] = incl[c];
        }
      }
      __syncthreads();
      int32_t mine[NC];
      double Rk[NC], tk[NC], tot[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double before = 0.0, all = 0.0;
        for (int k = 0; k < (NT / 64); ++k) {
          const double v = sh.part[par][k][c];
          if (k < wv) before = fmin(before + v, kFoldSat);
          all = fmin(all + v, kFoldSat);
        }
        tot[c] = all;
        double R = fmin(S[c] + fmin(before + excl[c], kFoldSat), kFoldSat);
        mine[c] = INT32_MAX;
        Rk[c] = 0.0;
        tk[c] = 0.0;
        if (R + L[c] < 0x1p53 - 1.0 && !((mytie >> c) & 1)) continue;  // nothing special here
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const int32_t i = base + j;
          if (mine[c] == INT32_MAX && i < wend && i >= pos[c]) {
            const double qq = q[c][j];
            if (qq - floor(qq) == 0.5 || !(qq < 0x1p53 - R)) {
              mine[c] = i;
              Rk[c] = R;
              tk[c] = src.term(stage[stage_slot<K, Src>(me * K + j)], c);
            } else {
              R += rint(fmin(qq, kFoldSat));
            }
          }
        }
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        int32_t v = mine[c];
        for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
        if (lane == 0) sh.first[par][wv][c] = v;
      }
      __syncthreads();
      int32_t first[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        int32_t m = sh.first[par][0][c];
        for (int k = 1; k < (NT / 64); ++k) m = min(m, sh.first[par][k][c]);
        first[c] = m;
        if (m == INT32_MAX) {
          if (pos[c] < wend) {  // the rest of the window, none special
            s[c] = ldexp(fmin(S[c] + tot[c], kFoldSat), x[c] - 52);
            pos[c] = wend;
          }
        } else if (mine[c] == m) {
          sh.Sk[c] = Rk[c];
          sh.tk[c] = tk[c];
        }
      }
      par ^= 1;
      __syncthreads();
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (first[c] == INT32_MAX) continue;
        s[c] = ldexp(sh.Sk[c], x[c] - 52) + sh.tk[c];  // the serial step itself
        pos[c] = first[c] + 1;
        if (!(s[c] <= 0x1.fffffffffffffp1023)) pos[c] = max(pos[c], wend);  // inf absorbs
      }
      __syncthreads();  // Sk / tk / part are rewritten by the next round
    }
    return true;
  };
  for (;;) {
    if (!pass(bufA, wA, bufB, wB)) break;
    if (!pass(bufB, wB, bufA, wA)) break;
  }
  if (seen < e)  // the last serial stretch reached e
    for (int32_t i = seen + me; i < e; i += NT)
      side += src.side(src.load(i), (double)(e - i));
  return fold_sum_d(side, sh);
}

// LDS stage of one window (plus one record of padding per K): kQeK records
// per thread of quantize_error terms, kTotK per thread of leaf records.
constexpr int kQeK = 4;
constexpr int kTotK = 2;
constexpr int kQeStage = kFoldThreads * (kQeK + 1) * (int)sizeof(double);
constexpr int kTotStage = kFoldThreads * (kTotK + 1) * (int)sizeof(TotRaw);
constexpr int kStageBytes = kQeStage > kTotStage ? kQeStage : kTotStage;

// total_color over [b, e) onto tot[] (one chain per channel) and number_unique.
__device__ void fold_tot(const Terms& T, int32_t b, int32_t e, double tot[4], uint64_t* nu,
                         FoldShared& sh, TotRaw* stage) {
  *nu += (uint64_t)exact_sums<4, kTotK>(tot, b, e, TotSrc{&T}, sh, stage);
}

// One workgroup per long segment.  quantize_error is a serial chain; its share
// of the root's chain is only needed approximately: sum_i term_i * (e - i).
// The quantize_error chain and the four total_color chains are separate
// launches (kFoldQe, kFoldTot1 over loop 1's pairs; kFoldTot2 over loop 2's).
enum FoldMode { kFoldQe = 0, kFoldTot2 = 1, kFoldTot1 = 2 };

template <int Mode>
__global__ __launch_bounds__(kFoldThreads) void k_fold_long(const Frame* __restrict__ fr, Batch B) {
  constexpr bool L2 = Mode == kFoldTot2;
  __shared__ FoldShared sh;
  __shared__ __align__(16) unsigned char stage[kStageBytes];
  __shared__ int32_t s_item;
  const int32_t nh = B.ctl[L2 ? 6 : 5], nl = B.ctl[L2 ? 4 : 3];
  for (;;) {  // a queue: the huge segments first, then the others
    if (threadIdx.x == 0) s_item = atomicAdd(&B.ctl[8 + Mode], 1);
    __syncthreads();
    const int32_t it = s_item;
    __syncthreads();
    if (it >= nh + nl) break;
    const uint64_t ent = it < nh ? B.huge_list[it] : B.long_list[it - nh];
    const Frame& d = fr[(int32_t)(ent >> 32)];
    const int32_t n = (int32_t)(ent & 0xFFFFFFFFu);
    const TreeDev& t = d.t;
    const Terms& T = d.T;
    const int32_t b = d.seg_b[n], e = d.seg_e[n];
    const int level = t.level[n];
    const int32_t s = seg_loop2(d, n, b, e);
    if (Mode == kFoldQe) {
      const double* term = T.term;
      double qe = 0.0;
      const double share = exact_sums<1, kQeK>(&qe, b, e, QeSrc{term}, sh, (double*)stage);
      if (threadIdx.x == 0) {
        t.qe[n] = qe;
        t.share[n] = share;
      }
    } else if (Mode == kFoldTot1) {  // loop 1's totals: leaves only
      double tot[4] = {0.0, 0.0, 0.0, 0.0};
      uint64_t nu = 0;
      if (level == 8 && s > b) fold_tot(T, b, s, tot, &nu, sh, (TotRaw*)stage);
      if (threadIdx.x == 0) {
        for (int c = 0; c < 4; ++c) t.total[4 * n + c] = tot[c];
        t.nu[n] = nu;
      }
    } else if (s < e) {
      double tot[4];
      for (int c = 0; c < 4; ++c) tot[c] = t.total[4 * n + c];
      uint64_t nu = t.nu[n];
      fold_tot(T, s, e, tot, &nu, sh, (TotRaw*)stage);
      if (threadIdx.x == 0) {
        for (int c = 0; c < 4; ++c) t.total[4 * n + c] = tot[c];
        t.nu[n] = nu;
      }
    }
    __syncthreads();
  }
}

// The running values of every pair, stored by run-major pair index for the
// exact root chain (only run when the bound cannot separate the root's error).
__global__ void k_fold_record(const Frame* __restrict__ fr, Map m) {
  const At at = map_at(m, 64);
  const Frame& d = fr[at.f];
  const int64_t n = at.i + 1;
  if (n > d.M || !d.ok || d.info[6] != 0 || !d.info[14]) return;
  double qe = 0.0;
  for (int32_t i = d.seg_b[n]; i < d.seg_e[n]; ++i) {
    qe += d.T.term[i];
    int64_t r;
    int l;
    pair_decode(d.spair[i], d.R1, d.R, d.depth, &r, &l);
    d.record[pair_rmajor(r, l, d.R1, d.depth)] = qe;  // the root chain's order
  }
}

// ---- 5. tree logic (one workgroup, level-synchronous) -------------------------------
// PruneToCubeDepth, Reduce and DefineImageColormap are recursive walks whose
// only order-dependent state is the live-node counter and the double sums a
// merge adds into a parent.  Each walk is restated as a few level-by-level
// sweeps over the nodes (the tree is at most 9 levels deep):
//
//   * A pass prunes with a predicate S (level > depth for PruneToCubeDepth,
//     quantize_error <= threshold for Reduce).  PruneChild(a) is entered at the
//     post-order visit of every S node a and attempts (rd_octree.cpp:69-93) each
//     node still attached below a, children first.  So node x is first attempted
//     in the "group" of a(x), its lowest S ancestor-or-self, and attempts are
//     ordered by (post-order of a(x), post-order of x inside the group).  A
//     merge happens while live > maxc, i.e. for the first live - maxc attempts;
//     the rank of an attempt is a prefix sum of group sizes plus its position in
//     its group.
//   * Merged nodes form a downward-closed set, and a parent receives its merged
//     children in attempt order: children that are S themselves (merged at their
//     own visit) in index order, then the others (merged later, inside an
//     ancestor's PruneChild) in index order.  One bottom-up sweep folds them in
//     exactly that order, so every double matches the serial walk.
//   * Reduce counts a node when it is not S and its number_unique, as it stands
//     when the walk visits it (after its S children merged), is non-zero; the
//     next threshold is the minimum quantize_error over those nodes.
//   * QuantizeErrorFlatten + qsort (rd_octree.cpp:388-398) only use one order
//     statistic.  glibc 2.35's qsort is a merge sort here whose merge takes the
//     left element when cmp(left, right) <= 0; QuantizeErrorCompare returns 1
//     exactly when left > right, so the merge is an ordinary stable ascending
//     merge and the MagickEpsilon tie never changes the order.  The statistic is
//     found by radix selection on the bit patterns (errors are >= 0).
//   * DefineImageColormap numbers the colour-bearing nodes in post-order: a
//     prefix sum over the post-order.  The candidate order of the search image is
//     the same post-order, so candidate k is colour k.
constexpr int kTreeThreads = 1024;
constexpr int kTreeWaves = kTreeThreads / 64;

struct TreeShared {
  int32_t off[10];     // level l occupies lvl[off[l], off[l+1])
  int32_t cur[10];
  int64_t part_i[kTreeWaves];
  double part_d[kTreeWaves];
  int32_t part_s[kTreeWaves];
  uint32_t hist[2048];
  uint64_t sel[3];
  int64_t scal[4];
};

struct Tree {
  TreeDev t;
  TreeWork W;
  int kids;
  int64_t maxc;
  int depth;
  int64_t M;
};

template <typename Fn>
__device__ __forceinline__ void for_level(const TreeShared& S, const TreeWork& W, int l, Fn&& fn) {
  for (int32_t i = S.off[l] + (int32_t)threadIdx.x; i < S.off[l + 1]; i += kTreeThreads)
    fn(W.lvl[i]);
}

__device__ int64_t block_sum(int64_t v, TreeShared& S) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) S.part_i[threadIdx.x >> 6] = v;
  __syncthreads();
  int64_t s = 0;
  for (int k = 0; k < kTreeWaves; ++k) s += S.part_i[k];
  __syncthreads();
  return s;
}

// Exclusive prefix of v over the workgroup, in thread order.
__device__ int64_t block_excl_i64(int64_t v, TreeShared& S) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) S.part_i[wave] = incl;
  __syncthreads();
  int64_t before = 0;
  for (int k = 0; k < wave; ++k) before += S.part_i[k];
  __syncthreads();
  return before + incl - v;
}

__device__ double block_min(double v, TreeShared& S) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  if ((threadIdx.x & 63) == 0) S.part_d[threadIdx.x >> 6] = v;
  __syncthreads();
  double m = S.part_d[0];
  for (int k = 1; k < kTreeWaves; ++k) m = fmin(m, S.part_d[k]);
  __syncthreads();
  return m;
}

// Exclusive prefix sum of a[0, n) in place; returns the total.
__device__ int32_t block_scan(int32_t* a, int32_t n, TreeShared& S) {
  const int32_t per = (n + kTreeThreads - 1) / kTreeThreads;
  const int32_t lo = min(n, (int32_t)threadIdx.x * per), hi = min(n, lo + per);
  int32_t local = 0;
  for (int32_t i = lo; i < hi; ++i) local += a[i];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int32_t incl = local;
  for (int off = 1; off < 64; off <<= 1) {
    const int32_t t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) S.part_s[wv] = incl;
  __syncthreads();
  int32_t before = 0, total = 0;
  for (int k = 0; k < kTreeWaves; ++k) {
    if (k < wv) before += S.part_s[k];
    total += S.part_s[k];
  }
  int32_t run = before + incl - local;
  for (int32_t i = lo; i < hi; ++i) {
    const int32_t v = a[i];
    a[i] = run;
    run += v;
  }
  __syncthreads();
  return total;
}

// Node ids grouped by level (the order inside a level is irrelevant).
__device__ void level_lists(const Tree& T, TreeShared& S) {
  if (threadIdx.x < 10) S.off[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t n = threadIdx.x; n <= T.M; n += kTreeThreads) atomicAdd(&S.off[T.t.level[n] + 1], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int l = 1; l < 10; ++l) S.off[l] += S.off[l - 1];
    for (int l = 0; l < 10; ++l) S.cur[l] = S.off[l];
  }
  __syncthreads();
  for (int64_t n = threadIdx.x; n <= T.M; n += kTreeThreads)
    T.W.lvl[atomicAdd(&S.cur[T.t.level[n]], 1)] = (int32_t)n;
  __syncthreads();
}

enum PassMode { kPruneCube = 0, kReduce = 1, kShape = 2, kPruneLevel = 3 };

// Attachment, the S ancestor, subtree and group sizes, post-order starts and
// pre-order indices of the current tree.
__device__ void tree_shape(const Tree& T, TreeShared& S, int mode, double thr) {
  const TreeDev& t = T.t;
  const TreeWork& W = T.W;
  for (int l = 0; l <= 8; ++l) {
    for_level(S, W, l, [&](int32_t y) {
      int a = 1;
      int32_t up = -1;
      if (l > 0) {
        const int32_t p = t.parent[y];
        a = W.att[p] && t.child[16 * p + t.id[y]] == y;
        up = W.anc[p];
      }
      bool s = false;
      if (a && mode == kPruneCube) s = t.level[y] > T.depth;
      // PruneLevel (rd_octree.cpp:95-102) enters PruneChild at every node of
      // the cube depth; the nodes below it are attempted inside that call.
      if (a && mode == kPruneLevel) s = t.level[y] == T.depth;
      if (a && mode == kReduce) s = t.qe[y] <= thr;
      W.att[y] = a;
      W.anc[y] = !a ? -1 : (s ? y : up);
    });
    __syncthreads();
  }
  for (int l = 8; l >= 0; --l) {
    for_level(S, W, l, [&](int32_t y) {
      if (!W.att[y]) return;
      int32_t sz = 1, g = 1;
      for (int k = 0; k < T.kids; ++k) {
        const int32_t ch = t.child[16 * y + k];
        if (ch < 0) continue;
        sz += W.size[ch];
        if (W.anc[ch] != ch) g += W.gsz[ch];
      }
      W.size[y] = sz;
      W.gsz[y] = g;
    });
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    W.start[0] = 0;
    W.gstart[0] = 0;
    W.pre[0] = 0;
  }
  __syncthreads();
  for (int l = 0; l < 8; ++l) {
    for_level(S, W, l, [&](int32_t y) {
      if (!W.att[y]) return;
      int32_t run = W.start[y], grun = W.gstart[y], prun = W.pre[y] + 1;
      for (int k = 0; k < T.kids; ++k) {
        const int32_t ch = t.child[16 * y + k];
        if (ch < 0) continue;
        W.start[ch] = run;
        W.pre[ch] = prun;
        run += W.size[ch];
        prun += W.size[ch];
        if (W.anc[ch] != ch) {
          W.gstart[ch] = grun;
          grun += W.gsz[ch];
        } else {
          W.gstart[ch] = 0;
        }
      }
    });
    __syncthreads();
  }
}

__device__ __forceinline__ int32_t post_of(const TreeWork& W, int32_t y) {
  return W.start[y] + W.size[y] - 1;
}

// Merges a walk makes with `live` nodes: every attempt when they fit within
// live - maxc, else that many.
__device__ int64_t tree_merges(const Tree& T, TreeShared& S, int64_t live, int64_t* attempts) {
  const TreeWork& W = T.W;
  int64_t mine = 0;
  for (int64_t y = threadIdx.x; y <= T.M; y += kTreeThreads) mine += (W.att[y] && W.anc[y] >= 0);
  *attempts = block_sum(mine, S);
  const int64_t K = live - T.maxc;
  return K > 0 ? min(*attempts, K) : 0;
}

// The first half of a PruneToCubeDepth or Reduce walk: the shape of the tree and
// which nodes merge (W.merged).  Reads only the links and levels (and qe for
// Reduce); returns the number of merges.
__device__ int64_t tree_mark(const Tree& T, TreeShared& S, int mode, double thr, int64_t live) {
  const TreeWork& W = T.W;
  tree_shape(T, S, mode, thr);
  int64_t attempts = 0;
  const int64_t merged = tree_merges(T, S, live, &attempts);
  const int64_t K = live - T.maxc;
  const bool cut = K > 0 && attempts > K;
  if (cut) {
    const int32_t A = W.size[0];
    for (int32_t i = threadIdx.x; i < A; i += kTreeThreads) W.arr[i] = 0;
    __syncthreads();
    for (int64_t y = threadIdx.x; y <= T.M; y += kTreeThreads)
      if (W.att[y] && W.anc[y] == (int32_t)y) W.arr[post_of(W, (int32_t)y)] = W.gsz[y];
    __syncthreads();
    block_scan(W.arr, A, S);
  }
  for (int64_t y = threadIdx.x; y <= T.M; y += kTreeThreads) {
    int m = 0;
    if (K > 0 && W.att[y] && W.anc[y] >= 0) {
      m = 1;
      if (cut) {
        const int32_t a = W.anc[y];
        const int64_t rank = (int64_t)W.arr[post_of(W, a)] + W.gstart[y] + W.gsz[y] - 1;
        m = rank < K;
      }
    }
    W.merged[y] = m;
  }
  __syncthreads();
  return merged;
}

// The second half: the merges W.merged marks, bottom-up.  Updates *live; for
// Reduce returns the colour count and the next threshold (seeded with root
// qe - 1).
__device__ void tree_merge(const Tree& T, TreeShared& S, int mode, int64_t merged, int64_t* live,
                           int64_t* count, double* next) {
  const TreeDev& t = T.t;
  const TreeWork& W = T.W;
  int64_t cnt = 0;
  double mn = INFINITY;
  for (int l = 8; l >= 0; --l) {
    for_level(S, W, l, [&](int32_t y) {
      if (!W.att[y]) return;
      int32_t* cl = t.child + 16 * y;
      double tot[4];
      for (int c = 0; c < 4; ++c) tot[c] = t.total[4 * y + c];
      uint64_t nu = t.nu[y];
      bool any = false;
      for (int k = 0; k < T.kids; ++k) {  // S children: merged at their own visit
        const int32_t ch = cl[k];
        if (ch < 0 || !W.merged[ch] || W.anc[ch] != ch) continue;
        nu += t.nu[ch];
        for (int c = 0; c < 4; ++c) tot[c] += t.total[4 * ch + c];
        any = true;
      }
      const uint64_t nu_visit = nu;
      for (int k = 0; k < T.kids; ++k) {  // the rest: merged inside an ancestor's PruneChild
        const int32_t ch = cl[k];
        if (ch < 0 || !W.merged[ch] || W.anc[ch] == ch) continue;
        nu += t.nu[ch];
        for (int c = 0; c < 4; ++c) tot[c] += t.total[4 * ch + c];
        any = true;
      }
      if (y == 0 && W.merged[0]) {  // the root's parent is the root itself
        nu += nu;
        for (int c = 0; c < 4; ++c) tot[c] += tot[c];
        any = true;
      }
      if (any) {
        for (int c = 0; c < 4; ++c) t.total[4 * y + c] = tot[c];
        t.nu[y] = nu;
        for (int k = 0; k < T.kids; ++k)
          if (cl[k] >= 0 && W.merged[cl[k]]) cl[k] = -1;
        if (y == 0 && W.merged[0]) cl[t.id[0]] = -1;
      }
      if (mode == kReduce && W.anc[y] != y) {
        if (nu_visit > 0) ++cnt;
        mn = fmin(mn, t.qe[y]);
      }
    });
    __syncthreads();
  }
  *live -= merged;
  if (mode == kReduce) {
    *count = block_sum(cnt, S);
    *next = fmin(block_min(mn, S), t.qe[0] - 1.0);
  }
}

// tree_merge with each node's children read together and their sums loaded
// four at a time (more registers; the row-by-row kernel keeps tree_merge).
__device__ void tree_merge_batched(const Tree& T, TreeShared& S, int mode, int64_t merged,
                                   int64_t* live, int64_t* count, double* next) {
  const TreeDev& t = T.t;
  const TreeWork& W = T.W;
  int64_t cnt = 0;
  double mn = INFINITY;
  for (int l = 8; l >= 0; --l) {
    for_level(S, W, l, [&](int32_t y) {
      if (!W.att[y]) return;
      int32_t* cl = t.child + 16 * y;
      // The children's links and marks are read together, then each merged
      // child's sums four slots at a time: the walk's dependent loads per node
      // are a few batches, not a chain per child.  The additions keep the
      // reference's order (S children by slot, then the rest by slot).
      uint32_t mS = 0, mR = 0;  // merged children by slot: S nodes, the rest
      {
        int32_t cs[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) cs[k] = k < T.kids ? cl[k] : -1;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
          const bool v = cs[k] >= 0;
          const int32_t mg = v ? W.merged[cs[k]] : 0;
          const int32_t an = v ? W.anc[cs[k]] : -1;
          if (mg) (an == cs[k] ? mS : mR) |= 1u << k;
        }
      }
      double tot[4];
      for (int c = 0; c < 4; ++c) tot[c] = t.total[4 * y + c];
      uint64_t nu = t.nu[y];
      auto add4 = [&](uint32_t mask, int k0) {
        uint64_t cn[4];
        double ct[4][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const bool on = (mask >> (k0 + j)) & 1u;
          const int32_t ch = on ? cl[k0 + j] : 0;
          cn[j] = on ? t.nu[ch] : 0;
#pragma unroll
          for (int c = 0; c < 4; ++c) ct[j][c] = on ? t.total[4 * ch + c] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (!((mask >> (k0 + j)) & 1u)) continue;
          nu += cn[j];
#pragma unroll
          for (int c = 0; c < 4; ++c) tot[c] += ct[j][c];
        }
      };
#pragma unroll 1
      for (int k0 = 0; k0 < 16; k0 += 4)  // S children: merged at their own visit
        if (mS >> k0 & 15u) add4(mS, k0);
      const uint64_t nu_visit = nu;
#pragma unroll 1
      for (int k0 = 0; k0 < 16; k0 += 4)  // the rest: merged inside an ancestor's PruneChild
        if (mR >> k0 & 15u) add4(mR, k0);
      bool any = (mS | mR) != 0;
      if (y == 0 && W.merged[0]) {  // the root's parent is the root itself
        nu += nu;
        for (int c = 0; c < 4; ++c) tot[c] += tot[c];
        any = true;
      }
      if (any) {
        for (int c = 0; c < 4; ++c) t.total[4 * y + c] = tot[c];
        t.nu[y] = nu;
#pragma unroll
        for (int k = 0; k < 16; ++k)
          if ((mS | mR) >> k & 1u) cl[k] = -1;
        if (y == 0 && W.merged[0]) cl[t.id[0]] = -1;
      }
      if (mode == kReduce && W.anc[y] != y) {
        if (nu_visit > 0) ++cnt;
        mn = fmin(mn, t.qe[y]);
      }
    });
    __syncthreads();
  }
  *live -= merged;
  if (mode == kReduce) {
    *count = block_sum(cnt, S);
    *next = fmin(block_min(mn, S), t.qe[0] - 1.0);
  }
}

// One PruneToCubeDepth or Reduce walk.
__device__ void tree_pass(const Tree& T, TreeShared& S, int mode, double thr, int64_t* live,
                          int64_t* count, double* next) {
  const int64_t merged = tree_mark(T, S, mode, thr, *live);
  tree_merge(T, S, mode, merged, live, count, next);
}

// The k-th smallest (0-based) of {qe of attached nodes with pre-order < L}
// plus `zeros` zeros: MSB-first radix selection on the bit patterns.
__device__ double select_kth(const Tree& T, TreeShared& S, int64_t L, int64_t k, int64_t zeros) {
  const TreeDev& t = T.t;
  const TreeWork& W = T.W;
  if (threadIdx.x == 0) {
    S.sel[0] = 0;  // prefix
    S.sel[1] = 0;  // mask
    S.sel[2] = (uint64_t)k;
  }
  __syncthreads();
  const int shifts[6] = {53, 42, 31, 20, 9, 0};
  const int widths[6] = {11, 11, 11, 11, 11, 9};
  for (int ps = 0; ps < 6; ++ps) {
    const int sh = shifts[ps];
    const uint64_t dm = (1ull << widths[ps]) - 1;
    for (int i = threadIdx.x; i < 2048; i += kTreeThreads) S.hist[i] = 0;
    __syncthreads();
    const uint64_t prefix = S.sel[0], mask = S.sel[1];
    for (int64_t y = threadIdx.x; y <= T.M; y += kTreeThreads) {
      if (!W.att[y] || W.pre[y] >= L) continue;
      const uint64_t b = (uint64_t)__double_as_longlong(t.qe[y]);
      if ((b & mask) == prefix) atomicAdd(&S.hist[(b >> sh) & dm], 1u);
    }
    if (threadIdx.x == 0 && prefix == 0 && zeros > 0) atomicAdd(&S.hist[0], (uint32_t)zeros);
    __syncthreads();
    // The first digit d whose cumulative count passes kk (the last digit, dm,
    // when none before it does), by a prefix over the bins: thread t holds
    // bins 2t and 2t + 1.
    static_assert(kTreeThreads * 2 == 2048, "two bins per thread");
    const uint64_t kk = S.sel[2];
    const int b0 = 2 * (int)threadIdx.x;
    const uint64_t h0 = S.hist[b0], h1 = S.hist[b0 + 1];
    const int64_t before = block_excl_i64(h0 + h1, S);
    const uint64_t c0 = (uint64_t)before, c1 = c0 + h0;
    // Bin b below dm is the one when its range [lo, lo + h) holds kk; dm is
    // when kk lies past every bin before it.  Exactly one bin qualifies.
    const int idm = (int)dm;
    auto hit = [&](int b, uint64_t lo, uint64_t h) {
      return b < idm ? (kk >= lo && kk < lo + h) : (b == idm && kk >= lo);
    };
    int64_t d = -1;
    uint64_t cum = 0;
    if (hit(b0, c0, h0)) {
      d = b0;
      cum = c0;
    } else if (hit(b0 + 1, c1, h1)) {
      d = b0 + 1;
      cum = c1;
    }
    if (d >= 0) {
      S.sel[2] = kk - cum;
      S.sel[0] = prefix | ((uint64_t)d << sh);
      S.sel[1] = mask | (dm << sh);
    }
    __syncthreads();
  }
  return __longlong_as_double((long long)S.sel[0]);
}

__global__ __launch_bounds__(kTreeThreads) void k_prune_cube(const Frame* __restrict__ fr) {
  const Frame& d = fr[blockIdx.x];
  if (!d.ok || d.info[6] != 0) return;
  int32_t* info = d.info;
  const Tree T{d.t, d.W, d.kids, d.maxc, d.depth, d.M};
  __shared__ TreeShared S;
  level_lists(T, S);
  if (threadIdx.x < 10) T.W.lvl_off[threadIdx.x] = S.off[threadIdx.x];
  int64_t live = 1 + (int64_t)info[4];
  // Loop 1's MaxQNodes test (rd_octree.cpp:366): the nodes loop 1 creates before
  // its last row are bounded by its node total.
  if (threadIdx.x == 0 && live > kMaxQNodes) info[6] = 1;
  if (info[0] < info[5]) {  // loop 1 broke: PruneToCubeDepth (rd_octree.cpp:370-372)
    int64_t cnt = 0;
    double nx = 0.0;
    const int64_t merged = tree_mark(T, S, kPruneCube, 0.0, live);
    tree_merge_batched(T, S, kPruneCube, merged, &live, &cnt, &nx);
  }
  // Loop 2's nodes join the live count; its MaxQNodes test (rd_octree.cpp:380)
  // is bounded by the total.
  live += T.M - (int64_t)info[4];
  if (threadIdx.x == 0) {
    info[7] = (int32_t)live;
    if (live > kMaxQNodes) info[6] = 1;
  }
}

// ReduceImageColors (rd_octree.cpp:388-404), DefineImageColormap (:142-176) and
// the search image.  info layout in sa_octree.h.
__global__ __launch_bounds__(kTreeThreads) void k_reduce(const Frame* __restrict__ fr) {
  const Frame& d = fr[blockIdx.x];
  if (!d.ok || d.info[6] != 0) return;
  const Tree T{d.t, d.W, d.kids, d.maxc, d.depth, d.M};
  const int assoc = d.assoc;
  const double root_qe = d.dinfo[2];
  int32_t* info = d.info;
  const PaletteBuild out = d.pb;
  __shared__ TreeShared S;
  const TreeDev& t = T.t;
  const TreeWork& W = T.W;
  if (threadIdx.x < 10) S.off[threadIdx.x] = W.lvl_off[threadIdx.x];
  if (threadIdx.x == 0) t.qe[0] = root_qe;
  __syncthreads();
  int64_t live = info[7];
  int64_t count = (int64_t)info[2] + info[3];
  double next = 0.0;
  if (count > T.maxc) {
    next = 0.0;
    if (live > 0) {
      tree_shape(T, S, kShape, 0.0);
      const int64_t L = live;
      int64_t mine = 0;
      for (int64_t y = threadIdx.x; y <= T.M; y += kTreeThreads) mine += (W.att[y] && W.pre[y] < L);
      const int64_t inc = block_sum(mine, S);
      const int64_t aggressive = (110 * (T.maxc + 1)) / 100;
      if (L > aggressive) next = select_kth(T, S, L, L - aggressive, L - inc);
    }
  }
  int passes = 0;
  while (count > T.maxc) {
    // The reference loops forever if a pass prunes nothing; stop instead.
    if (++passes > 100000) {
      if (threadIdx.x == 0) info[6] = 3;
      return;
    }
    const double thr = next;
    tree_pass(T, S, kReduce, thr, &live, &count, &next);
  }
  // DefineImageColormap: colour numbers are the post-order ranks of the
  // colour-bearing nodes.  Search-image rows are the pre-order ranks of the
  // internal nodes (the root always has a row).
  tree_shape(T, S, kShape, 0.0);
  const int32_t A = W.size[0];
  for (int32_t i = threadIdx.x; i < A; i += kTreeThreads) {
    W.arr[i] = 0;
    W.arr2[i] = 0;
  }
  __syncthreads();
  for (int64_t y = threadIdx.x; y <= T.M; y += kTreeThreads) {
    if (!W.att[y]) continue;
    bool internal = (y == 0);
    for (int k = 0; k < T.kids; ++k) internal |= t.child[16 * y + k] >= 0;
    W.merged[y] = internal;  // reused as the internal flag
    if (t.nu[y] > 0) W.arr[post_of(W, (int32_t)y)] = 1;
    if (internal) W.arr2[W.pre[y]] = 1;
  }
  __syncthreads();
  const int32_t cnt = block_scan(W.arr, A, S);
  const int32_t rows = block_scan(W.arr2, A, S);
  const PaletteLayout L = palette_layout(cnt, rows, T.kids);
  if (threadIdx.x == 0) {
    info[8] = cnt;
    info[9] = (int32_t)live;
    info[10] = rows;
    info[11] = L.bytes;
  }
  if (L.bytes > out.image_cap || rows > 32767 || cnt > T.maxc) {
    if (threadIdx.x == 0) info[6] = 2;
    return;
  }
  unsigned char* img = out.image;
  double* col = (double*)(img + L.off_col);
  float* ent = (float*)(img + L.off_ent);
  float* cent = (float*)(img + L.off_cent);
  int32_t* sub = (int32_t*)(img + L.off_sub);
  int16_t* child = (int16_t*)(img + L.off_child);
  uint16_t* cand = (uint16_t*)(img + L.off_cand);
  for (int64_t y = threadIdx.x; y <= T.M; y += kTreeThreads) {
    if (!W.att[y]) {
      t.row[y] = -1;
      continue;
    }
    const int32_t post = post_of(W, (int32_t)y);
    if (t.nu[y] > 0) {
      const int32_t cn = W.arr[post];
      t.color_number[y] = (uint32_t)cn;
      double* e = out.entries + 4 * cn;
      const double alpha = safe_recip((double)t.nu[y]);
      const double* tot = t.total + 4 * y;
      if (!assoc) {
        e[0] = (double)(float)(alpha * kRange * tot[0]);
        e[1] = (double)(float)(alpha * kRange * tot[1]);
        e[2] = (double)(float)(alpha * kRange * tot[2]);
        e[3] = kRange;
      } else {
        const double opacity = alpha * kRange * tot[3];
        e[3] = (double)(float)opacity;
        if (e[3] == kRange) {
          e[0] = (double)(float)(alpha * kRange * tot[0]);
          e[1] = (double)(float)(alpha * kRange * tot[1]);
          e[2] = (double)(float)(alpha * kRange * tot[2]);
        } else {
          const double gamma = safe_recip(kScale * e[3]);
          e[0] = (double)(float)(alpha * gamma * kRange * tot[0]);
          e[1] = (double)(float)(alpha * gamma * kRange * tot[1]);
          e[2] = (double)(float)(alpha * gamma * kRange * tot[2]);
        }
      }
      for (int k = 0; k < 4; ++k) {
        ent[4 * cn + k] = (float)e[k];
        cent[4 * cn + k] = (float)e[k];
      }
      assoc_entry(assoc != 0, e, col + 4 * cn);
      cand[cn] = (uint16_t)cn;
    }
    if (!W.merged[y]) {
      t.row[y] = -1;
      continue;
    }
    const int32_t r = W.arr2[W.pre[y]];
    t.row[y] = r;
    const int32_t first = W.arr[W.start[y]];
    sub[2 * r] = first;
    sub[2 * r + 1] = W.arr[post] + (t.nu[y] > 0 ? 1 : 0) - first;
    for (int k = 0; k < T.kids; ++k) {
      const int32_t ch = t.child[16 * y + k];
      child[r * T.kids + k] =
          ch < 0 ? kNoChild : (W.merged[ch] ? (int16_t)W.arr2[W.pre[ch]] : kLeaf);
    }
  }
  // The descent table is filled by k_descent (one thread per entry).
  if (threadIdx.x == 0) {
    PaletteDev* h = out.hdr;
    h->count = cnt;
    h->assoc = assoc;
    h->kids = T.kids;
    h->rows = rows;
    h->bytes = L.bytes;
    h->off_ent = L.off_ent;
    h->off_cent = L.off_cent;
    h->off_sub = L.off_sub;
    h->off_desc = L.off_desc;
    h->off_child = L.off_child;
    h->off_cand = L.off_cand;
    h->pad_ = 0;
    h->image = out.image;
  }
}

// The three-level descent table of the search image (sa_palette.h).
__global__ void k_descent(const Frame* __restrict__ fr, Map m) {
  const At at = map_at(m, 256);
  const Frame& d = fr[at.f];
  const int64_t e = at.i;
  if (!d.ok || d.info[6] != 0) return;
  const PaletteDev h = *d.pb.hdr;
  if (e >= desc_entries(h.kids)) return;
  unsigned char* img = (unsigned char*)h.image;
  ((uint32_t*)(img + h.off_desc))[e] =
      descent3((const int16_t*)(img + h.off_child), h.kids, (uint32_t)e);
}

// Exact root chain (only when the bound cannot separate it): the running
// values of every pair, in pair order.
__global__ void k_root_exact(const Frame* __restrict__ fr) {
  const Frame& d = fr[blockIdx.x];
  if (threadIdx.x != 0 || !d.ok || d.info[6] != 0 || !d.info[14]) return;
  double s = 0.0;
  for (int64_t p = 0; p < d.P; ++p) s += d.record[p];
  d.dinfo[2] = s;
}

// dinfo[0] = the parallel root chain (sum of shares), dinfo[1] = max non-root
// qe, dinfo[2] = the root qe Reduce uses; info[14] = 1 when the serial chain
// must decide it.  The approximation differs from the serial chain by at most
// ~3 P u sum (u = 2^-53): the reference's own rounding, the running values it
// adds, and the order of the parallel sum.  A clear separation of 1 + 8 P u sum
// from every other node's error is demanded.
__global__ void k_root_bound(const Frame* __restrict__ fr) {
#pragma clang fp contract(off)
  const Frame& d = fr[blockIdx.x];
  if (!d.ok || d.info[6] != 0) return;
  const TreeDev& t = d.t;
  const int64_t M = d.M;
  __shared__ double ssum[256], smax[256];
  double s = 0.0, m = 0.0;
  for (int64_t n = 1 + threadIdx.x; n <= M; n += 256) {
    s += t.share[n];
    m = fmax(m, t.qe[n]);
  }
  ssum[threadIdx.x] = s;
  smax[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) {
      ssum[threadIdx.x] += ssum[threadIdx.x + off];
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double root_qe = ssum[0], mx = smax[0];
    const double err = (double)d.P * 2.3e-16 * fabs(root_qe) * 4.0;
    const bool exact =
        d.exact_root || !(root_qe - 1.0 - err > mx + 1e-9 * mx + 2e-12) || M == 0;
    d.dinfo[0] = root_qe;
    d.dinfo[1] = mx;
    d.dinfo[2] = root_qe;
    d.info[14] = exact ? 1 : 0;
  }
}

// ---- 6. trees that pass MaxQNodes: classification row by row ----------------------
// When the live node count has passed MaxQNodes at the start of a row, the
// reference prunes the cube depth's level (PruneLevel, then depth - 1) and the
// rows after it descend to the new depth (rd_octree.cpp:366-369, :380-383).
// Node creation, the prunes and every sum then depend on the row at which the
// count crossed, so such an image is classified again by one workgroup that
// follows the reference row by row, each row in parallel:
//
//   * Descent (a chunk of up to kRowChunk consecutive runs at a time, level by
//     level).  A run whose child is missing asks for it under (parent, child
//     index); the first run that asks creates it, the others take it.  New
//     nodes are staged while the chunk descends, then numbered in the
//     reference's creation order -- by creating run, and by level within a
//     run (a run creates consecutive levels down to its leaf).
//   * Sums (level by level).  Every run's term is formed in parallel, the
//     runs are sorted by node (then by run), and each node's chain is folded
//     in run order onto its value from the rows before.  The root's chain of
//     running totals is added serially, run by run and level by level, so the
//     root's quantize_error is exact here.
//   * At the start of a row PruneLevel runs as a tree pass (kPruneLevel); after
//     a loop-1 row whose colours passed the palette, PruneToCubeDepth.
//
// The pass leaves the node arrays, the live and colour counts and the root's
// error where k_reduce takes them.  A launch stops before a row whose new
// nodes could pass the node capacity (status 1); the host grows the arrays and
// launches again from that row.
constexpr int kRowChunk = 2048;
constexpr int kRowLevels = 8;
constexpr int kStageMax = kRowChunk * kRowLevels;
constexpr int kRowHash = 4096;

struct RowsState {
  int64_t y, r, M, live, colors;
  double root_qe;
  int32_t loop, depth, status, pad;
};

struct RowsArgs {
  const float4* px;
  const int32_t* count;
  const uint32_t* pack;
  const int32_t* row_first;  // first run of each row, row_first[h] = R
  int64_t h;
  int32_t assoc, kids, maxc, px_pack;  // px_pack: px is rebuilt from the pack (run_px)
  TreeDev t;
  TreeWork W;
  int64_t cap;               // node capacity of t and W
  int32_t* at;               // node of each run of the chunk at each level, 9 x kRowChunk
  double* rv;                // running value of each (run, level) pair, run-major, limit
                             // per run (the root's chain order)
  double* lt;                // leaf terms of total_color of each run, kRowChunk x 4
  double* tmg;               // quantize_error term of each (level, run), 8 x kRowChunk
  uint64_t* skey;            // each level's runs sorted by (node, run), 8 x kRowChunk
  int32_t* seg;              // chain starts: level << 16 | sorted position
  double* ts;                // the terms in sorted order, 8 x kRowChunk
  double* rvs;               // running values in sorted order, 8 x kRowChunk
  double* lts;               // leaf terms in sorted order, kRowChunk x 4
  int32_t* scnt;             // run lengths in sorted order (leaf level)
  int32_t* sp;               // staged nodes: parent, child index | level << 8, creating run
  int32_t* sp_child;         // 16 per staged node
  int32_t* fid;              // node id of each staged node
  RowsState* st;
  int32_t* info;
  double* dinfo;
};

// A node reference during a chunk's descent: an id (>= 0), none (-1), or
// staged node s (-2 - s).
__device__ __forceinline__ int32_t rows_child(const RowsArgs& a, int32_t p, int cid) {
  return p >= 0 ? a.t.child[16 * (int64_t)p + cid] : a.sp_child[16 * (-2 - p) + cid];
}

__device__ __forceinline__ void rows_pixel(const RowsArgs& a, int64_t r, double av[4]) {
  const float4 f = run_px(a.px, a.pack, a.px_pack, r);
  assoc_pixel(a.assoc != 0, f.x, f.y, f.z, f.w, av);
}

// One chunk of a row: runs [c0, c0 + n), levels 1..limit.  Everything uniform
// is passed by reference and updated identically in every thread.
// One node's chain of a level, from sorted position k: the terms in run order
// onto the node's value from the rows before, B at a time (a batch's loads
// overlap; the additions stay in order).  Leaf levels also fold the totals.
template <bool LEAF, int B>
struct FoldBatch {
  uint32_t nd[B];
  double tv[B];
  double lv[LEAF ? B : 1][4];
  int32_t cv[LEAF ? B : 1];
};

template <bool LEAF, int B>
__device__ __forceinline__ void rows_load(const RowsArgs& a, const uint64_t* sk, const double* ts,
                                          int j, int n, FoldBatch<LEAF, B>& f) {
#pragma unroll
  for (int q = 0; q < B; ++q) {
    const bool in = j + q < n;
    f.nd[q] = in ? (uint32_t)(sk[j + q] >> 32) : ~0u;
    f.tv[q] = in ? ts[j + q] : 0.0;
    if (LEAF) {
      for (int c = 0; c < 4; ++c) f.lv[q][c] = in ? a.lts[4 * (j + q) + c] : 0.0;
      f.cv[q] = in ? a.scnt[j + q] : 0;
    }
  }
}

template <bool LEAF, int B>
__device__ __forceinline__ void rows_fold(const RowsArgs& a, int l, int k, int n) {
  const TreeDev& t = a.t;
  const uint64_t* sk = a.skey + (l - 1) * kRowChunk;
  const double* ts = a.ts + (l - 1) * kRowChunk;
  double* rvs = a.rvs + (l - 1) * kRowChunk;
  const uint32_t node = (uint32_t)(sk[k] >> 32);
  double v = t.qe[node];
  double tot[4] = {0.0, 0.0, 0.0, 0.0};
  uint64_t nu = 0;
  if (LEAF) {
    for (int c = 0; c < 4; ++c) tot[c] = t.total[4 * (int64_t)node + c];
    nu = t.nu[node];
  }
  FoldBatch<LEAF, B> cur;
  rows_load<LEAF, B>(a, sk, ts, k, n, cur);
  for (int j = k;; j += B) {
    bool end = false;
#pragma unroll
    for (int q = 0; q < B; ++q) {
      if (end || cur.nd[q] != node) {
        end = true;
        continue;
      }
      v += cur.tv[q];
      rvs[j + q] = v;
      if (LEAF) {
        for (int c = 0; c < 4; ++c) tot[c] += cur.lv[q][c];
        nu += (uint64_t)cur.cv[q];
      }
    }
    if (end) break;
    rows_load<LEAF, B>(a, sk, ts, j + B, n, cur);
  }
  t.qe[node] = v;
  if (LEAF) {
    for (int c = 0; c < 4; ++c) t.total[4 * (int64_t)node + c] = tot[c];
    t.nu[node] = nu;
  }
}

constexpr int kRootK = 8;  // root-chain terms per thread and window (8192 per window)

__device__ void rows_chunk(const RowsArgs& a, TreeShared& S, unsigned char* raw, int64_t c0,
                           int n, int limit, int colors_level, int64_t& M, int64_t& live,
                           int64_t& colors, double& root_qe) {
  const int tid = threadIdx.x;
  const bool assoc = a.assoc != 0;
  const TreeDev& t = a.t;
  int32_t* cur = (int32_t*)raw;                               // kRowChunk
  uint64_t* hkey = (uint64_t*)(raw + 4 * kRowChunk);          // kRowHash
  uint32_t* hmin = (uint32_t*)(hkey + kRowHash);              // kRowHash
  int32_t* hval = (int32_t*)(hmin + kRowHash);                // kRowHash
  int32_t* nstage = (int32_t*)&S.scal[0];
  constexpr int kPer = kRowChunk / kTreeThreads;
  for (int i = tid; i < n; i += kTreeThreads) {
    cur[i] = 0;
    a.at[i] = 0;
  }
  if (tid == 0) *nstage = 0;
  __syncthreads();
  // -- descent ------------------------------------------------------------------
  for (int l = 1; l <= limit; ++l) {
    for (int k = tid; k < kRowHash; k += kTreeThreads) {
      hkey[k] = kEmpty64;
      hmin[k] = kEmpty32;
    }
    __syncthreads();
    int32_t nxt[kPer], par[kPer];
    uint32_t slot[kPer];
    int cidv[kPer];
    for (int q = 0; q < kPer; ++q) {
      const int i = tid + q * kTreeThreads;
      nxt[q] = -1;
      slot[q] = kEmpty32;
      if (i >= n) continue;
      const int32_t p = cur[i];
      const uint32_t pk = a.pack[c0 + i];
      const int cid = node_id(assoc, pk & 0xFFu, (pk >> 8) & 0xFFu, (pk >> 16) & 0xFFu, pk >> 24,
                              8 - l);
      par[q] = p;
      cidv[q] = cid;
      const int32_t c = rows_child(a, p, cid);
      nxt[q] = c;
      if (c != -1) continue;
      const uint64_t key = ((((uint64_t)(uint32_t)p) << 8) | (uint64_t)cid) + 1ull;
      uint32_t sl = hash32(key) & (kRowHash - 1);
      for (;;) {  // at most kRowChunk keys in kRowHash slots: a free slot exists
        const unsigned long long prev =
            atomicCAS((unsigned long long*)&hkey[sl], (unsigned long long)kEmpty64,
                      (unsigned long long)key);
        if (prev == kEmpty64 || prev == key) break;
        sl = (sl + 1) & (kRowHash - 1);
      }
      atomicMin(&hmin[sl], (uint32_t)i);
      slot[q] = sl;
    }
    __syncthreads();
    for (int q = 0; q < kPer; ++q) {
      const int i = tid + q * kTreeThreads;
      if (slot[q] == kEmpty32 || hmin[slot[q]] != (uint32_t)i) continue;
      const int32_t s = atomicAdd(nstage, 1);  // < kStageMax: one per (run, level)
      a.sp[3 * s] = par[q];
      a.sp[3 * s + 1] = cidv[q] | (l << 8);
      a.sp[3 * s + 2] = i;
      for (int k = 0; k < 16; ++k) a.sp_child[16 * s + k] = -1;
      if (par[q] >= 0)
        a.t.child[16 * (int64_t)par[q] + cidv[q]] = -2 - s;
      else
        a.sp_child[16 * (-2 - par[q]) + cidv[q]] = -2 - s;
      hval[slot[q]] = -2 - s;
    }
    __syncthreads();
    for (int q = 0; q < kPer; ++q) {
      const int i = tid + q * kTreeThreads;
      if (i >= n) continue;
      const int32_t v = slot[q] == kEmpty32 ? nxt[q] : hval[slot[q]];
      cur[i] = v;
      a.at[l * kRowChunk + i] = v;
    }
    __syncthreads();
  }
  // -- number the staged nodes in creation order -----------------------------------
  const int32_t ns = *nstage;
  int32_t* made = (int32_t*)raw;                    // per run: nodes it created
  int32_t* base = made + kRowChunk;                 // their exclusive scan
  for (int i = tid; i < n; i += kTreeThreads) made[i] = 0;
  __syncthreads();
  for (int s = tid; s < ns; s += kTreeThreads) atomicAdd(&made[a.sp[3 * s + 2]], 1);
  __syncthreads();
  for (int i = tid; i < n; i += kTreeThreads) base[i] = made[i];
  __syncthreads();
  block_scan(base, n, S);
  for (int s = tid; s < ns; s += kTreeThreads) {
    const int i = a.sp[3 * s + 2];
    const int lev = a.sp[3 * s + 1] >> 8;
    a.fid[s] = (int32_t)(M + 1 + base[i] + (lev - (limit - made[i] + 1)));
  }
  __syncthreads();
  int64_t ncol = 0;
  for (int s = tid; s < ns; s += kTreeThreads) {
    const int32_t f = a.fid[s];
    const int32_t p = a.sp[3 * s];
    const int cid = a.sp[3 * s + 1] & 0xFF, lev = a.sp[3 * s + 1] >> 8;
    const int i = a.sp[3 * s + 2];
    t.key[f] = node_key(a.pack[c0 + i], lev, assoc);
    t.level[f] = lev;
    t.id[f] = cid;
    t.parent[f] = p >= 0 ? p : a.fid[-2 - p];
    t.qe[f] = 0.0;
    t.share[f] = 0.0;
    for (int c = 0; c < 4; ++c) t.total[4 * (int64_t)f + c] = 0.0;
    t.nu[f] = 0;
    t.color_number[f] = 0;
    t.row[f] = -1;
    for (int k = 0; k < 16; ++k) {
      const int32_t ch = a.sp_child[16 * s + k];
      t.child[16 * (int64_t)f + k] = ch >= -1 ? ch : a.fid[-2 - ch];
    }
    if (p >= 0) t.child[16 * (int64_t)p + cid] = f;
    if (lev == colors_level) ++ncol;
  }
  for (int l = 1; l <= limit; ++l)
    for (int i = tid; i < n; i += kTreeThreads) {
      const int32_t v = a.at[l * kRowChunk + i];
      if (v < -1) a.at[l * kRowChunk + i] = a.fid[-2 - v];
    }
  ncol = block_sum(ncol, S);  // synchronises
  M += ns;
  live += ns;
  colors += ncol;
  // -- sums ---------------------------------------------------------------------------
  // Level by level, every run's term is formed and the runs are sorted by node
  // -- a stable block radix sort, so each node's runs stay in run order -- into
  // global arrays in sorted order; then the chains of every level are folded
  // at once (the levels' nodes are distinct, so the longest chain sets the
  // time, not the sum of the levels' longest), and the running values go back
  // to run-major order for the root's chain.
  using RowSort = rocprim::block_radix_sort<uint32_t, kTreeThreads, kRowChunk / kTreeThreads,
                                            uint32_t>;
  static_assert(sizeof(typename RowSort::storage_type) <= 4 * kRowChunk + 16 * kRowHash,
                "the sort's storage fits the chunk's LDS");
  typename RowSort::storage_type& sort_st = *reinterpret_cast<typename RowSort::storage_type*>(raw);
  int32_t* nseg = (int32_t*)&S.scal[2];
  const int end_bit = 64 - __builtin_clzll((unsigned long long)(M + 2));  // every id + padding
  const uint32_t pad = (uint32_t)(((uint64_t)1 << end_bit) - 1);          // above every id
  if (tid == 0) *nseg = 0;
  for (int l = 1; l <= limit; ++l) {
    double* tm = a.tmg + (l - 1) * kRowChunk;
    constexpr int kPer2 = kRowChunk / kTreeThreads;
    uint32_t sk[kPer2], sv[kPer2];
    for (int q = 0; q < kPer2; ++q) {  // blocked: thread t holds runs kPer2 * t + q
      const int i = kPer2 * tid + q;
      sv[q] = (uint32_t)i;
      if (i >= n) {
        sk[q] = pad;
        continue;
      }
      const int32_t node = a.at[l * kRowChunk + i];
      double av[4], mid[4];
      rows_pixel(a, c0 + i, av);
      node_mid(t.key[node], mid);
      const double e0 = kScale * (av[0] - mid[0]);
      const double e1 = kScale * (av[1] - mid[1]);
      const double e2 = kScale * (av[2] - mid[2]);
      const double e3 = assoc ? kScale * (av[3] - mid[3]) : 0.0;
      double dd = e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3;
      if (dd != dd) dd = 0.0;
      const int32_t cnt = a.count[c0 + i];
      tm[i] = (double)cnt * sqrt(dd);
      sk[q] = (uint32_t)node;
      if (l == limit) {  // this run's leaf terms, folded below in run order
        double* lt = a.lt + 4 * i;
        lt[0] = tot_term(cnt, clamp_pixel(av[0]));
        lt[1] = tot_term(cnt, clamp_pixel(av[1]));
        lt[2] = tot_term(cnt, clamp_pixel(av[2]));
        lt[3] = tot_term(cnt, clamp_pixel(assoc ? av[3] : kRange));
      }
    }
    __syncthreads();  // the terms are written; the sort reuses the chunk's LDS
    RowSort().sort(sk, sv, sort_st, 0, (unsigned)end_bit);
    __syncthreads();
    uint32_t* lnode = (uint32_t*)raw;  // the sorted nodes, to find where chains start
    for (int q = 0; q < kPer2; ++q) lnode[kPer2 * tid + q] = sk[q];
    __syncthreads();
    uint64_t* skl = a.skey + (l - 1) * kRowChunk;
    double* ts = a.ts + (l - 1) * kRowChunk;
    for (int q = 0; q < kPer2; ++q) {
      const int k = kPer2 * tid + q;
      if (k >= n) continue;
      const int i = (int)sv[q];
      skl[k] = ((uint64_t)sk[q] << 32) | (uint32_t)i;
      ts[k] = tm[i];
      if (l == limit) {
        for (int c = 0; c < 4; ++c) a.lts[4 * k + c] = a.lt[4 * i + c];
        a.scnt[k] = a.count[c0 + i];
      }
      if (k == 0 || lnode[k - 1] != sk[q]) a.seg[atomicAdd(nseg, 1)] = (l << 16) | k;
    }
    __syncthreads();
  }
  const int segs = *nseg;
  for (int e = tid; e < segs; e += kTreeThreads) {
    const int l = a.seg[e] >> 16, k = a.seg[e] & 0xFFFF;
    if (l == limit)
      rows_fold<true, 4>(a, l, k, n);
    else
      rows_fold<false, 8>(a, l, k, n);
  }
  __syncthreads();
  for (int e = tid; e < limit * n; e += kTreeThreads) {  // running values, run-major
    const int l = 1 + e / n, k = e % n;
    const int i = (int)(uint32_t)a.skey[(l - 1) * kRowChunk + k];
    a.rv[(int64_t)limit * i + (l - 1)] = a.rvs[(l - 1) * kRowChunk + k];
  }
  __syncthreads();
  // The root's chain: run by run, level by level (rd_octree.cpp:341): the
  // running values lie in that order, limit per run.  Summed exactly in
  // parallel (exact_sums: the running values are finite and non-negative),
  // its windows staged in the chunk's LDS, which is free by now.
  __shared__ FoldSharedT<kTreeWaves> fsh;
  static_assert(kTreeThreads * (kRootK + 1) * sizeof(double) <= 4 * kRowChunk + 16 * kRowHash,
                "the root chain's stage fits the chunk's LDS");
  double q = root_qe;
  (void)exact_sums<1, kRootK, QeSrc, kTreeThreads>(&q, 0, n * limit, QeSrc{a.rv}, fsh,
                                                   (double*)raw);
  root_qe = q;
}

__global__ __launch_bounds__(kTreeThreads) void k_rows(RowsArgs a) {
  __shared__ TreeShared S;
  __shared__ __align__(16) unsigned char raw[4 * kRowChunk + 16 * kRowHash];
  const int tid = threadIdx.x;
  const RowsState st = *a.st;
  int64_t y = st.y, r = st.r, M = st.M, live = st.live, colors = st.colors;
  double root_qe = st.root_qe;
  int loop = st.loop, depth = st.depth, status = 0;
  const TreeDev& t = a.t;
  if (y == 0 && M == 0 && tid == 0) {  // the root (rd_octree.cpp:282-285)
    t.key[0] = 0;
    t.level[0] = 0;
    t.id[0] = 0;
    t.parent[0] = 0;
    t.qe[0] = 0.0;
    t.share[0] = 0.0;
    for (int c = 0; c < 4; ++c) t.total[c] = 0.0;
    t.nu[0] = 0;
    t.color_number[0] = 0;
    t.row[0] = -1;
    for (int k = 0; k < 16; ++k) t.child[k] = -1;
  }
  __syncthreads();
  while (y < a.h) {
    const int64_t re = a.row_first[y + 1];
    int limit = loop == 1 ? kMaxDepth : depth;
    if (M + 1 + (re - r) * limit > a.cap) {
      status = 1;
      break;
    }
    if (live > kMaxQNodes) {  // PruneLevel, then depth - 1
      const Tree T{a.t, a.W, a.kids, a.maxc, depth, M};
      level_lists(T, S);
      int64_t cnt = 0;
      double nx = 0.0;
      tree_pass(T, S, kPruneLevel, 0.0, &live, &cnt, &nx);
      --depth;
      if (loop == 2) limit = depth;
    }
    for (int64_t c0 = r; c0 < re; c0 += kRowChunk) {
      const int n = (int)(re - c0 < kRowChunk ? re - c0 : kRowChunk);
      rows_chunk(a, S, raw, c0, n, limit, limit, M, live, colors, root_qe);
    }
    if (loop == 1 && colors > a.maxc) {  // PruneToCubeDepth, then loop 2
      const Tree T{a.t, a.W, a.kids, a.maxc, depth, M};
      level_lists(T, S);
      int64_t cnt = 0;
      double nx = 0.0;
      tree_pass(T, S, kPruneCube, 0.0, &live, &cnt, &nx);
      loop = 2;
    }
    ++y;
    r = re;
  }
  if (status == 0) {
    const Tree T{a.t, a.W, a.kids, a.maxc, depth, M};
    level_lists(T, S);
    if (tid < 10) a.W.lvl_off[tid] = S.off[tid];
    if (tid == 0) {
      a.info[2] = (int32_t)colors;
      a.info[3] = 0;
      a.info[6] = 0;
      a.info[7] = (int32_t)live;
      a.info[14] = 1;
      a.dinfo[0] = root_qe;
      a.dinfo[2] = root_qe;
      t.qe[0] = root_qe;
    }
  }
  if (tid == 0) {
    RowsState o;
    o.y = y;
    o.r = r;
    o.M = M;
    o.live = live;
    o.colors = colors;
    o.root_qe = root_qe;
    o.loop = loop;
    o.depth = depth;
    o.status = status == 0 ? 2 : status;
    o.pad = 0;
    *a.st = o;
  }
}

// row_first[y] = the first run of row y (every row starts a run), row_first[h] = R.
__global__ void k_row_first(const int32_t* __restrict__ row, int64_t R, int64_t h,
                            int32_t* __restrict__ row_first) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i == 0) row_first[h] = (int32_t)R;
  if (i >= R) return;
  if (i == 0 || row[i] != row[i - 1]) row_first[row[i]] = (int32_t)i;
}

// k_descent for one image.
__global__ void k_descent_one(const Frame* __restrict__ fr) {
  const Frame& d = *fr;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (!d.ok || d.info[6] != 0) return;
  const PaletteDev h = *d.pb.hdr;
  if (e >= desc_entries(h.kids)) return;
  unsigned char* img = (unsigned char*)h.image;
  ((uint32_t*)(img + h.off_desc))[e] =
      descent3((const int16_t*)(img + h.off_child), h.kids, (uint32_t)e);
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned blocks(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

// ---------------------------------------------------------------------------------
// Host orchestration.  A batch runs in four phases, each closed by one
// synchronisation that reads back the scalars sizing the next phase (runs per
// image; the loop-1 extent; nodes per image; the results).  Each phase's
// arrays are one stream-ordered allocation; descriptors go up and results come
// back through a pinned buffer kept per host thread.
// ---------------------------------------------------------------------------------
namespace {
std::atomic<int> g_cells{1};       // the octree.cells knob
std::atomic<int> g_exact_root{0};  // the octree.exact_root knob
}  // namespace

void octree_set_cells(int on) { g_cells.store(on ? 1 : 0); }
int octree_cells() { return g_cells.load(); }
void octree_set_exact_root(int on) { g_exact_root.store(on ? 1 : 0); }
int octree_exact_root() { return g_exact_root.load(); }

int octree_depth(int maxc, bool assoc, bool gray) {  // rd_octree.cpp:255-275
  int depth = 1;
  for (int64_t c = maxc; c != 0; ++depth) c >>= 2;
  if (depth > 2) --depth;
  if (assoc && depth > 5) --depth;
  if (gray) depth = 8;
  if (depth > 8) depth = 8;
  if (depth < 2) depth = 2;
  return depth;
}

namespace {

// The two sorts of the build (node entries by first pair, intervals by node)
// hold a few hundred thousand keys on a video batch; rocprim's default sorts
// that many by a block sort and eleven merge passes (22 launches each).
// Onesweep at every size above one block: a pass per 8 key bits.
using SortCfg = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                           rocprim::default_config, 0>;

// Images are grouped into sub-batches of at most this many pixels (a larger
// image is a sub-batch of its own), which bounds the scratch of one phase and
// keeps every batch-wide index within 32 bits.
constexpr int64_t kBatchPixels = (int64_t)1 << 26;

enum MapKind { kMapNT, kMapR, kMapP, kMapCap, kMapM256, kMapM64, kMapDesc, kMapDC, kMapDT, kMaps };

uint32_t pow2_at_least(int64_t v) {
  uint32_t c = 1;
  while ((int64_t)c < v) c <<= 1;
  return c;
}

int bits_for(int64_t n) {  // radix bits that hold every key in [0, n)
  int b = 1;
  while (b < 32 && ((int64_t)1 << b) < n) ++b;
  return b;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// One stream-ordered allocation carved into 256-byte aligned arrays.
struct Carve {
  struct Req {
    void** p;
    size_t bytes;
  };
  std::vector<Req> reqs;
  template <typename T>
  void want(T** p, int64_t n) {
    reqs.push_back({(void**)p, align256((size_t)std::max<int64_t>(n, 1) * sizeof(T))});
  }
  hipError_t take(std::vector<void*>& owned, hipStream_t s) {
    size_t total = 0;
    for (const Req& r : reqs) total += r.bytes;
    void* base = nullptr;
    const hipError_t e = scratch_alloc(&base, total, s);
    if (e != hipSuccess) return e;
    owned.push_back(base);
    size_t off = 0;
    for (const Req& r : reqs) {
      *r.p = (char*)base + off;
      off += r.bytes;
    }
    reqs.clear();
    return hipSuccess;
  }
};

// Pinned staging per host thread (grown on demand, kept for the process).
struct Staging {
  void* p = nullptr;
  size_t n = 0;
};
thread_local Staging tl_staging;

hipError_t staging(size_t bytes, char** out) {
  Staging& st = tl_staging;
  if (st.n < bytes) {
    if (st.p) (void)hipHostFree(st.p);
    st.p = nullptr;
    st.n = 0;
    const size_t want = std::max<size_t>(bytes, 64 << 10);
    const hipError_t e = hipHostMalloc(&st.p, want, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    st.n = want;
  }
  *out = (char*)st.p;
  return hipSuccess;
}

class SubBuild {
 public:
  SubBuild(const OctreeReq* rq, int F, hipStream_t s, OctreeOut* outs, OctreeResult* res)
      : rq_(rq), F_(F), s_(s), out_(outs), res_(res), H_(F) {}

  void run() {
    hipError_t e = hipSuccess;
    const char* where = "";
    if ((e = phase_runs()) != hipSuccess) where = "runs";
    if (e == hipSuccess && (e = phase_extent()) != hipSuccess) where = "loop-1 extent";
    if (e == hipSuccess && (e = phase_nodes()) != hipSuccess) where = "nodes";
    if (e == hipSuccess && (e = phase_tree()) != hipSuccess) where = "tree";
    if (e != hipSuccess) {
      (void)hipStreamSynchronize(s_);
      for (int f = 0; f < F_; ++f)
        if (res_[f].status == 0) fail(f, 3, "octree build (%s): %s", where, hipGetErrorString(e));
    }
    for (int f = 0; f < F_; ++f)
      if (res_[f].status != 0) drop_outputs(f);
    for (void* q : owned_) (void)scratch_free(q, s_);
    owned_.clear();
  }

 private:
  const OctreeReq* rq_;
  const int F_;
  hipStream_t s_;
  OctreeOut* out_;
  OctreeResult* res_;
  std::vector<Frame> H_;
  std::vector<void*> owned_;
  char* pin_ = nullptr;
  Frame* d_fr_ = nullptr;
  int64_t* d_tab_ = nullptr;
  char* d_rb_ = nullptr;
  int64_t grid_[kMaps] = {};
  Batch B_{};
  // host copies of what the device reports
  const int32_t* hinfo(int f) const { return (const int32_t*)(pin_ + up_bytes()) + 16 * f; }
  const double* hdinfo(int f) const {
    return (const double*)(pin_ + up_bytes() + 64 * (size_t)F_) + 4 * f;
  }
  const int32_t* hctl() const { return (const int32_t*)(pin_ + up_bytes() + 96 * (size_t)F_); }
  static constexpr int kCtl = 16;
  size_t up_bytes() const { return sizeof(Frame) * F_ + sizeof(int64_t) * kMaps * (F_ + 1); }
  size_t rb_bytes() const { return 96 * (size_t)F_ + 4 * kCtl; }
  Map map(int kind) const { return Map{d_tab_ + (size_t)kind * (F_ + 1), F_}; }

  void fail(int f, int status, const char* fmt, ...) __attribute__((format(printf, 4, 5))) {
    res_[f].status = status;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(res_[f].msg, sizeof(res_[f].msg), fmt, ap);
    va_end(ap);
    H_[f].ok = 0;
  }

  void drop_outputs(int f) {
    OctreeOut& o = out_[f];
    if (o.block) (void)hipFree(o.block);
    if (o.tree_block) (void)hipFree(o.tree_block);
    o.block = o.tree_block = nullptr;
    o.image = nullptr;
    o.hdr = nullptr;
    o.entries = nullptr;
    o.tree = TreeDev{};
  }

  // Element counts per image for every launch shape, from what is known so far.
  void fill_maps(int64_t* tab) {
    for (int k = 0; k < kMaps; ++k) {
      int64_t* first = tab + (size_t)k * (F_ + 1);
      const int per = k == kMapM64 ? 64 : (k == kMapDT || k == kMapNT) ? 1 : 256;
      int64_t run = 0;
      for (int f = 0; f < F_; ++f) {
        const Frame& d = H_[f];
        int64_t n = 0;
        switch (k) {
          case kMapNT: n = d.ok ? (d.N + kRunTile - 1) / kRunTile : 0; break;
          case kMapR: n = d.ok ? d.R : 0; break;
          case kMapP: n = d.P_ins; break;  // failed images keep their pairs in the sort
          case kMapDC: n = d.dense ? d.dcells : 0; break;
          case kMapDT: n = d.dense && d.ok ? d.dtiles : 0; break;
          case kMapCap: n = d.ok ? d.ncap : 0; break;
          case kMapM256:
          case kMapM64: n = d.ok ? d.M : 0; break;
          case kMapDesc: n = d.ok && d.pb.hdr ? desc_entries(d.kids) : 0; break;
        }
        first[f] = run;
        run += blocks(n, per);
      }
      first[F_] = run;
      grid_[k] = run;
    }
  }

  hipError_t upload() {
    memcpy(pin_, H_.data(), sizeof(Frame) * F_);
    fill_maps((int64_t*)(pin_ + sizeof(Frame) * F_));
    return hipMemcpyAsync(d_fr_, pin_, up_bytes(), hipMemcpyHostToDevice, s_);
  }

  hipError_t readback() {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(pin_ + up_bytes(), d_rb_, rb_bytes(), hipMemcpyDeviceToHost, s_);
    if (e == hipSuccess) e = hipStreamSynchronize(s_);
    return e;
  }

  template <typename T>
  void fill(T* p, int64_t n, T v);

  // 1. runs: run starts per tile, one scan over the batch's tiles, runs per image.
  hipError_t phase_runs() {
    hipError_t e = staging(up_bytes() + rb_bytes(), &pin_);
    if (e != hipSuccess) return e;
    int64_t sumT = 0;  // quarters of run tiles
    for (int f = 0; f < F_; ++f) sumT += 4 * ((rq_[f].w * rq_[f].h + kRunTile - 1) / kRunTile);
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, tcnt_, toff_, (int)sumT, s_);
    if (e != hipSuccess) return e;
    char* up = nullptr;
    Carve c;
    c.want(&up, (int64_t)up_bytes());  // descriptors, then the block tables
    c.want(&d_rb_, (int64_t)rb_bytes());
    c.want(&tcnt_, sumT);
    c.want(&toff_, sumT);
    c.want((char**)&tmp, (int64_t)tmp_bytes);
    if ((e = c.take(owned_, s_)) != hipSuccess) return e;
    d_fr_ = (Frame*)up;
    d_tab_ = (int64_t*)(up + sizeof(Frame) * F_);
    int32_t* info = (int32_t*)d_rb_;
    double* dinfo = (double*)(d_rb_ + 64 * (size_t)F_);
    B_.ctl = (int32_t*)(d_rb_ + 96 * (size_t)F_);
    for (int f = 0; f < F_; ++f) {
      const OctreeReq& q = rq_[f];
      Frame& d = H_[f];
      memset(&d, 0, sizeof(Frame));
      d.src = q.src;
      d.fmt = q.fmt;
      d.assoc = q.assoc ? 1 : 0;
      d.px_pack = !d.assoc && (q.fmt == kRgb8 || q.fmt == kRgba8);
      d.exact_root = g_exact_root.load();
      d.maxc = q.maxc <= 0 ? 65536 : std::min(q.maxc, 65536);
      d.depth = octree_depth(d.maxc, d.assoc != 0, q.gray != 0);
      d.kids = d.assoc ? 16 : 8;
      d.ok = 1;
      d.w = q.w;
      d.h = q.h;
      d.N = q.w * q.h;
      d.info = info + 16 * f;
      d.dinfo = dinfo + 4 * f;
    }
    if ((e = upload()) != hipSuccess) return e;
    k_frame_init<<<F_, 64, 0, s_>>>(d_fr_);
    if (grid_[kMapNT]) {
      k_run_tiles<<<grid_[kMapNT], 256, 0, s_>>>(d_fr_, map(kMapNT), tcnt_);
      e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, tcnt_, toff_, (int)sumT, s_);
      if (e != hipSuccess) return e;
    }
    k_run_count<<<F_, 64, 0, s_>>>(d_fr_, map(kMapNT), tcnt_, toff_);
    return readback();
  }

  // 2. run records and loop 1's extent (one workgroup per image).
  hipError_t phase_extent() {
    int64_t sumR = 0, sumPx = 0, sumB = 0;
    for (int f = 0; f < F_; ++f) {
      Frame& d = H_[f];
      d.R = hinfo(f)[12];
      d.bcap = pow2_at_least(2 * ((int64_t)d.maxc + 1 + 1024));
      sumR += d.R;
      sumPx += d.px_pack ? 0 : d.R;
      sumB += d.bcap;
    }
    int32_t *start = nullptr, *count = nullptr, *row = nullptr;
    float4* px = nullptr;
    uint32_t *pack = nullptr, *bfirst = nullptr;
    uint64_t* bkeys = nullptr;
    Carve c;
    c.want(&start, sumR);
    c.want(&count, sumR);
    c.want(&row, sumR);
    c.want(&px, sumPx);
    c.want(&pack, sumR);
    c.want(&bkeys, sumB);
    c.want(&bfirst, sumB);
    hipError_t e = c.take(owned_, s_);
    if (e != hipSuccess) return e;
    int64_t ro = 0, po = 0, bo = 0;
    for (int f = 0; f < F_; ++f) {
      Frame& d = H_[f];
      d.start = start + ro;
      d.count = count + ro;
      d.row = row + ro;
      d.px = d.px_pack ? nullptr : px + po;
      d.pack = pack + ro;
      d.bkeys = bkeys + bo;
      d.bfirst = bfirst + bo;
      ro += d.R;
      po += d.px_pack ? 0 : d.R;
      bo += d.bcap;
    }
    if ((e = upload()) != hipSuccess) return e;
    if (grid_[kMapNT]) {
      k_run_emit<<<grid_[kMapNT], 256, 0, s_>>>(d_fr_, map(kMapNT), toff_);
      k_run_tail<<<blocks(4 * grid_[kMapNT], 256), 256, 0, s_>>>(d_fr_, map(kMapNT), tcnt_, toff_,
                                                                4 * grid_[kMapNT]);
    }
    fill(bkeys, sumB, kEmpty64);
    fill(bfirst, sumB, kEmpty32);
    k_break<<<F_, 1024, 0, s_>>>(d_fr_);
    return readback();
  }

  // 3. the node hash tables and the list of node entries.
  hipError_t phase_nodes() {
    int64_t sumP = 0, sumC = 0;
    for (int f = 0; f < F_; ++f) {
      Frame& d = H_[f];
      if (!d.ok) continue;
      if (hinfo(f)[6] == 4) {
        fail(f, 3, "internal: leaf hash table overflow");
        continue;
      }
      d.R1 = hinfo(f)[1];
      const int64_t P = 8 * d.R1 + (int64_t)d.depth * (d.R - d.R1);
      if (P >= ((int64_t)1 << 31)) {
        fail(f, 5, "too many runs");
        continue;
      }
      int64_t bound_nodes = 8 * ((int64_t)d.maxc + 1 + d.w) + 1;
      int64_t pw = 1;
      for (int l = 1; l <= d.depth; ++l) {
        pw = std::min<int64_t>(pw * d.kids, (int64_t)1 << 40);
        bound_nodes += std::min<int64_t>(d.R, pw);
      }
      bound_nodes = std::min<int64_t>(bound_nodes, P + 1);
      bound_[f] = bound_nodes;
      d.P = P;
      d.P_ins = P;
      int64_t cells = 0;
      for (int64_t l = 1, pw = 1; l <= d.depth; ++l) cells += (pw *= d.kids);
      int64_t cmax = 1;
      for (int l = 1; l <= d.depth; ++l) cmax *= d.kids;
      // Tiles of 8192 runs (4096 when the deepest level has 4096 cells), so
      // the scatter's LDS stays within 64 KiB; at most 2^22 table entries.
      const int sh = cmax <= 512 ? 13 : 12;
      const int64_t tiles = (d.R - d.R1 + ((int64_t)1 << sh) - 1) >> sh;
      if (g_cells.load() && cells <= kDenseCellsMax && d.R > d.R1 &&
          tiles * cells <= ((int64_t)1 << 22)) {
        d.dense = 1;
        d.dcells = (int32_t)cells;
        d.dtshift = sh;
        d.dtiles = (int32_t)((d.R - d.R1 + ((int64_t)1 << sh) - 1) >> sh);
        d.P_ins = 8 * d.R1;
        d.exp1 = d.P_ins <= 65536;  // loop 1 of a few rows: its intervals one per thread
      }
      d.ncap = pow2_at_least(2 * bound_nodes + 64);
      d.pair_base = sumP;
      sumP += P;
      sumC += d.ncap;
    }
    if (sumP >= ((int64_t)1 << 32) - 1) {  // one image per sub-batch is bounded above
      for (int f = 0; f < F_; ++f)
        if (H_[f].ok) fail(f, 5, "too many runs");
      return hipSuccess;
    }
    sumP_ = sumP;
    // Every image's pair range at its running offset (empty for the images
    // that failed), so the ranges are ascending and tile [0, sumP).
    nwg_ = 0;
    for (int64_t f = 0, run = 0; f < F_; ++f) {
      H_[f].pair_base = run;
      run += H_[f].P;
      nwg_ += blocks(H_[f].P_ins, 256);  // k_node_insert's workgroups (the kMapP grid)
    }
    int64_t sumDC = 0, sumDT = 0;
    max_dc_ = 0;
    max_scatter_lds_ = 0;
    for (int f = 0; f < F_; ++f) {
      Frame& d = H_[f];
      if (!d.dense) continue;
      d.dbase = sumDC;
      sumDC += d.dcells;
      sumDT += (int64_t)d.dtiles * d.dcells;
      max_dc_ = std::max<int64_t>(max_dc_, d.dcells);
      int64_t cmax = 1;
      for (int l = 1; l <= d.depth; ++l) cmax *= d.kids;
      max_scatter_lds_ = std::max<int64_t>(max_scatter_lds_,
                                           dense_scatter_lds((int64_t)1 << d.dtshift, cmax));
    }
    sumDC_ = sumDC;
    uint64_t* nkeys = nullptr;
    uint32_t *nfirst = nullptr, *pair_slot = nullptr;
    Carve c;
    c.want(&nkeys, sumC);
    c.want(&nfirst, sumC);
    size_t tmp_red = 0;
    hipError_t e = hipcub::DeviceReduce::Sum(nullptr, tmp_red, (int32_t*)nullptr,
                                             (int32_t*)nullptr, (int)nwg_, s_);
    if (e != hipSuccess) return e;
    char* red_tmp = nullptr;
    c.want(&pair_slot, sumP);
    c.want(&wgc_, nwg_);
    c.want(&red_tmp, (int64_t)tmp_red);
    c.want(&B_.ekey, sumC);
    c.want(&B_.eslot, sumC);
    c.want(&slot_node_, sumC);
    int32_t *dtab = nullptr, *dtot = nullptr;
    uint32_t *dfirst = nullptr, *dslot = nullptr;
    c.want(&dtab, sumDT);
    c.want(&dfirst, sumDC);
    c.want(&dtot, sumDC);
    c.want(&dslot, sumDC);
    if ((e = c.take(owned_, s_)) != hipSuccess) return e;
    pair_slot_ = pair_slot;
    for (int64_t f = 0, to = 0; f < F_; ++f) {
      Frame& d = H_[f];
      if (!d.dense) continue;
      d.dtab = dtab + to;
      d.dfirst = dfirst + d.dbase;
      d.dtot = dtot + d.dbase;
      d.dslot = dslot + d.dbase;
      to += (int64_t)d.dtiles * d.dcells;
    }
    int64_t co = 0;
    for (int f = 0; f < F_; ++f) {
      Frame& d = H_[f];
      if (!d.ok) continue;
      d.nkeys = nkeys + co;
      d.nfirst = nfirst + co;
      d.slot_node = slot_node_ + co;
      d.pair_slot = pair_slot + d.pair_base;
      co += d.ncap;
    }
    if ((e = upload()) != hipSuccess) return e;
    fill(B_.ctl, kCtl, 0);
    fill(nkeys, sumC, kEmpty64);
    fill(nfirst, sumC, kEmpty32);
    if (grid_[kMapP]) {
      k_node_insert<<<grid_[kMapP], 256, 0, s_>>>(d_fr_, map(kMapP), wgc_);
      if ((e = hipcub::DeviceReduce::Sum(red_tmp, tmp_red, wgc_, B_.ctl + 12, (int)nwg_, s_)) !=
          hipSuccess)
        return e;
    }
    if (grid_[kMapDT]) {  // loop 2 of the images placed by cell
      fill(dfirst, sumDC, kEmpty32);
      k_dense_count<<<grid_[kMapDT], 256, dense_count_lds(max_dc_), s_>>>(d_fr_, map(kMapDT));
      k_dense_scan<<<grid_[kMapDC], 256, 0, s_>>>(d_fr_, map(kMapDC));
      k_dense_insert<<<grid_[kMapDC], 256, 0, s_>>>(d_fr_, map(kMapDC));
    }
    if (grid_[kMapP]) k_node_collect<<<grid_[kMapCap], 256, 0, s_>>>(d_fr_, map(kMapCap), B_);
    // The outputs outlive the batch (plain allocations owned by the caller):
    // allocated while the node kernels run, sized by the node bound.
    for (int f = 0; f < F_; ++f) {
      Frame& d = H_[f];
      if (!d.ok) continue;
      const PaletteLayout cap =
          palette_layout(d.maxc, (int)std::min<int64_t>(bound_[f] + 1, 32767), d.kids);
      const size_t img = align256((size_t)cap.bytes), hdr = align256(sizeof(PaletteDev));
      void* blk = nullptr;
      if ((e = hipMalloc(&blk, img + hdr + sizeof(double) * 4 * (size_t)d.maxc)) != hipSuccess)
        return e;
      OctreeOut& o = out_[f];
      o.block = blk;
      o.image = (unsigned char*)blk;
      o.hdr = (PaletteDev*)((char*)blk + img);
      o.entries = (double*)((char*)blk + img + hdr);
      d.pb.image = o.image;
      d.pb.image_cap = cap.bytes;
      d.pb.hdr = o.hdr;
      d.pb.entries = o.entries;
    }
    return readback();
  }

  // 4. node order, ordered sums, PruneToCubeDepth, the root, Reduce, colormap.
  hipError_t phase_tree() {
    const int64_t E = hctl()[0];
    const int64_t sumP = sumP_;
    int64_t eb = 0, sumNodes = 0, sumW = 0;
    for (int f = 0; f < F_; ++f) {
      Frame& d = H_[f];
      if (d.P == 0) continue;  // took no part in phase 3
      d.entry_base = eb;
      eb += hinfo(f)[13];
      if (d.ok && hinfo(f)[6] == 4) fail(f, 3, "internal: node hash table overflow");
      d.M = d.ok ? hinfo(f)[13] : 0;
      d.node_base = sumNodes;
      sumNodes += d.ok ? d.M + 1 : 1;
    }
    if (sumP == 0) return readback();
    for (int f = 0; f < F_; ++f)
      if (H_[f].ok) sumW += H_[f].M + 2;
    const int64_t NI = hctl()[12];  // intervals (k_node_insert)
    const int64_t NT = NI + sumDC_;  // and one per cell of the images placed by cell
    const int bitsA = bits_for(sumP), bitsB = bits_for(sumNodes + 1);
    size_t tmpA = 0, tmpB = 0, tmpSel = 0, tmpSum = 0;
    hipError_t e = rocprim::radix_sort_pairs<SortCfg>(nullptr, tmpA, B_.ekey, (uint32_t*)nullptr,
                                                       B_.eslot, (uint32_t*)nullptr, (int)E, 0,
                                                       bitsA, s_);
    if (e != hipSuccess) return e;
    e = rocprim::radix_sort_pairs<SortCfg>(nullptr, tmpB, (int32_t*)nullptr, (int32_t*)nullptr,
                                           (uint32_t*)nullptr, (uint32_t*)nullptr, (int)NT, 0,
                                           bitsB, s_);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmpSel, (int32_t*)nullptr, (int32_t*)nullptr,
                                         (int)nwg_, s_);
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmpSum, (int32_t*)nullptr, (int32_t*)nullptr,
                                         (int)NT, s_);
    if (e != hipSuccess) return e;
    const size_t tmpBytes = std::max({tmpA, tmpB, tmpSel, tmpSum});
    TreeDev t{};
    uint32_t *skey = nullptr, *sslot = nullptr, *spair = nullptr;
    int32_t *pp = nullptr, *seg_b = nullptr, *seg_e = nullptr, *lcnt = nullptr;
    double* term = nullptr;
    float4* lval = nullptr;
    int32_t* w[11] = {};
    int32_t* lvl_off = nullptr;
    char* tmp = nullptr;
    Carve c;
    c.want(&t.key, sumNodes);
    c.want(&t.child, 16 * sumNodes);
    c.want(&t.parent, sumNodes);
    c.want(&t.id, sumNodes);
    c.want(&t.level, sumNodes);
    c.want(&t.qe, sumNodes);
    c.want(&t.share, sumNodes);
    c.want(&t.total, 4 * sumNodes);
    c.want(&t.nu, sumNodes);
    c.want(&t.color_number, sumNodes);
    c.want(&t.row, sumNodes);
    c.want(&skey, E);
    c.want(&sslot, E);
    c.want(&pp, 2 * sumP);  // the running values of the exact root chain
    c.want(&spair, sumP);
    uint32_t *istart = nullptr, *ival = nullptr, *sival = nullptr;
    int32_t *ikey = nullptr, *sikey = nullptr, *slen = nullptr, *spos = nullptr, *wgo = nullptr;
    c.want(&istart, NT);
    c.want(&ival, NT);
    c.want(&sival, NT);
    c.want(&ikey, NT);
    c.want(&sikey, NT);
    c.want(&slen, NT);
    c.want(&spos, NT);
    c.want(&wgo, nwg_);
    int32_t* vlen = nullptr;
    uint32_t* dpos = nullptr;
    c.want(&vlen, sumDC_);
    c.want(&dpos, sumDC_);
    int32_t* l2s = nullptr;
    c.want(&l2s, sumNodes);
    const int64_t ntiles = (sumP + kExpandTile - 1) / kExpandTile;
    int32_t* tile_first = nullptr;
    c.want(&tile_first, ntiles);
    int32_t* tiles = nullptr;  // the tiles k_expand takes
    c.want(&tiles, ntiles);
    bool any_exp1 = false;
    for (int f = 0; f < F_; ++f) any_exp1 |= H_[f].P > 0 && H_[f].exp1;
    c.want(&seg_b, sumNodes);
    c.want(&seg_e, sumNodes);
    c.want(&term, sumP);
    c.want(&lval, sumP);
    c.want(&lcnt, sumP);
    c.want(&B_.mid_list, sumNodes);
    c.want(&B_.long_list, sumNodes);
    c.want(&B_.huge_list, sumNodes);
    for (int k = 0; k < 11; ++k) c.want(&w[k], sumW);
    c.want(&lvl_off, 10 * (int64_t)F_);
    c.want(&tmp, (int64_t)tmpBytes);
    if ((e = c.take(owned_, s_)) != hipSuccess) return e;
    B_.skey = skey;
    B_.sslot = sslot;
    int64_t wo = 0;
    for (int f = 0; f < F_; ++f) {
      Frame& d = H_[f];
      if (d.P == 0) continue;
      const int64_t nb = d.node_base, pb = d.pair_base;
      TreeDev& td = d.t;
      td.key = t.key + nb;
      td.child = t.child + 16 * nb;
      td.parent = t.parent + nb;
      td.id = t.id + nb;
      td.level = t.level + nb;
      td.qe = t.qe + nb;
      td.share = t.share + nb;
      td.total = t.total + 4 * nb;
      td.nu = t.nu + nb;
      td.color_number = t.color_number + nb;
      td.row = t.row + nb;
      d.seg_b = seg_b + nb;
      d.seg_e = seg_e + nb;
      d.spair = spair + pb;
      d.T.term = term + pb;
      d.T.lval = lval + pb;
      d.T.lcnt = lcnt + pb;
      d.record = (double*)pp + pb;
      d.dpos = dpos;
      d.l2s = l2s + nb;
      if (d.ok) {
        int32_t** wp[11] = {&d.W.lvl, &d.W.att, &d.W.anc, &d.W.size, &d.W.gsz, &d.W.start,
                            &d.W.gstart, &d.W.pre, &d.W.merged, &d.W.arr, &d.W.arr2};
        for (int k = 0; k < 11; ++k) *wp[k] = w[k] + wo;
        d.W.lvl_off = lvl_off + 10 * f;
        wo += d.M + 2;
      }
    }
    if ((e = upload()) != hipSuccess) return e;
    fill(B_.ctl, kCtl, 0);
    if (E > 0) {
      e = rocprim::radix_sort_pairs<SortCfg>(tmp, tmpA, B_.ekey, skey, B_.eslot, sslot, (int)E, 0,
                                             bitsA, s_);
      if (e != hipSuccess) return e;
    }
    fill(t.child, 16 * sumNodes, -1);
    fill(t.color_number, sumNodes, 0u);
    fill(seg_b, sumNodes, 0);
    fill(seg_e, sumNodes, 0);
    if (sumDC_) fill(l2s, sumNodes, (int32_t)INT32_MAX);
    k_root_init<<<F_, 64, 0, s_>>>(d_fr_);
    if (grid_[kMapM256]) k_node_assign<<<grid_[kMapM256], 256, 0, s_>>>(d_fr_, map(kMapM256), B_);
    // Pairs into node order through their intervals: the intervals' first pairs
    // in pair order, a stable sort of the intervals by node, each sorted
    // interval's first position, then every position from its tile's intervals.
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tmpSel, wgc_, wgo, (int)nwg_, s_);
    if (e != hipSuccess) return e;
    k_compact<<<grid_[kMapP], 256, 0, s_>>>(d_fr_, map(kMapP), wgo, pair_slot_, istart, ikey,
                                            ival);
    if (grid_[kMapDC])
      k_dense_virtual<<<grid_[kMapDC], 256, 0, s_>>>(d_fr_, map(kMapDC), NI, (int32_t)sumNodes,
                                                    istart, ikey, ival, vlen);
    e = rocprim::radix_sort_pairs<SortCfg>(tmp, tmpB, ikey, sikey, ival, sival, (int)NT, 0, bitsB,
                                           s_);
    if (e != hipSuccess) return e;
    const unsigned gi = (unsigned)blocks(NT, 256);
    k_interval_len<<<gi, 256, 0, s_>>>(d_fr_, F_, sival, istart, NI, NT, sumP, vlen, slen);
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tmpSum, slen, spos, (int)NT, s_);
    if (e != hipSuccess) return e;
    k_interval_ends<<<gi, 256, 0, s_>>>(d_fr_, F_, sikey, sival, spos, slen, NI, NT,
                                        (int32_t)sumNodes, dpos, tile_first);
    // Tiles of the sorted range that touch an image not expanded per interval
    // (every tile when no image is).
    std::vector<int32_t> tl;
    for (int f = 0; f < F_ && any_exp1; ++f) {
      const Frame& d = H_[f];
      if (d.P == 0 || d.exp1) continue;
      const int64_t t0 = d.pair_base / kExpandTile, t1 = (d.pair_base + d.P - 1) / kExpandTile;
      for (int64_t t = std::max<int64_t>(t0, tl.empty() ? 0 : tl.back() + 1); t <= t1; ++t)
        tl.push_back((int32_t)t);
    }
    if (!any_exp1) {
      k_expand<<<(unsigned)ntiles, 256, 0, s_>>>(d_fr_, F_, sumP, nullptr, NT, tile_first, sikey,
                                                 sival, istart, spos, spair);
    } else if (!tl.empty()) {
      if ((e = hipMemcpyAsync(tiles, tl.data(), sizeof(int32_t) * tl.size(), hipMemcpyHostToDevice,
                              s_)) != hipSuccess)
        return e;
      k_expand<<<(unsigned)tl.size(), 256, 0, s_>>>(d_fr_, F_, sumP, tiles, NT, tile_first, sikey,
                                                    sival, istart, spos, spair);
    }
    if (any_exp1)
      k_expand_short<<<gi, 256, 0, s_>>>(d_fr_, F_, NI, NT, sikey, sival, istart, spos, slen,
                                          spair);
    if (grid_[kMapDT])
      k_dense_scatter<false><<<grid_[kMapDT], kScatterThreads, max_scatter_lds_, s_>>>(
          d_fr_, map(kMapDT), spair);
    if (grid_[kMapM256]) k_links<<<grid_[kMapM256], 256, 0, s_>>>(d_fr_, map(kMapM256));
    // At most sumP / kShortSeg segments are medium, sumP / kMidSeg long.
    const int mid_grid = (int)std::min<int64_t>(std::max<int64_t>(sumP / kShortSeg, 1), 8192);
    const int long_grid = (int)std::min<int64_t>(std::max<int64_t>(sumP / kMidSeg, 1), 2048);
    if (grid_[kMapM64]) k_fold_short<false><<<grid_[kMapM64], 64, 0, s_>>>(d_fr_, map(kMapM64), B_);
    k_fold_long<kFoldQe><<<long_grid, kFoldThreads, 0, s_>>>(d_fr_, B_);
    k_fold_long<kFoldTot1><<<long_grid, kFoldThreads, 0, s_>>>(d_fr_, B_);
    k_fold_mid<false><<<mid_grid, 64, 0, s_>>>(d_fr_, B_);
    k_prune_cube<<<F_, kTreeThreads, 0, s_>>>(d_fr_);
    if (grid_[kMapM64]) k_fold_short<true><<<grid_[kMapM64], 64, 0, s_>>>(d_fr_, map(kMapM64), B_);
    k_fold_long<kFoldTot2><<<long_grid, kFoldThreads, 0, s_>>>(d_fr_, B_);
    k_fold_mid<true><<<mid_grid, 64, 0, s_>>>(d_fr_, B_);
    k_root_bound<<<F_, 256, 0, s_>>>(d_fr_);
    if (grid_[kMapDT])  // pair indices of the cell-placed images whose root is summed serially
      k_dense_scatter<true><<<grid_[kMapDT], kScatterThreads, max_scatter_lds_, s_>>>(
          d_fr_, map(kMapDT), spair);
    if (grid_[kMapM64]) k_fold_record<<<grid_[kMapM64], 64, 0, s_>>>(d_fr_, map(kMapM64));
    k_root_exact<<<F_, 64, 0, s_>>>(d_fr_);
    k_reduce<<<F_, kTreeThreads, 0, s_>>>(d_fr_);
    if (grid_[kMapDesc]) k_descent<<<grid_[kMapDesc], 256, 0, s_>>>(d_fr_, map(kMapDesc));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int f = 0; f < F_; ++f)
      if ((e = keep_tree(f)) != hipSuccess) return e;
    if ((e = readback()) != hipSuccess) return e;
    // Trees that may pass MaxQNodes are classified again row by row.
    std::vector<int> rows;
    for (int f = 0; f < F_; ++f)
      if (H_[f].ok && hinfo(f)[6] == 1) rows.push_back(f);
    for (int f : rows)
      if ((e = rows_build(f)) != hipSuccess) return e;
    if (!rows.empty() && (e = readback()) != hipSuccess) return e;
    for (int f = 0; f < F_; ++f) {
      Frame& d = H_[f];
      if (!d.ok) continue;
      const int32_t* in = hinfo(f);
      OctreeResult& r = res_[f];
      if (in[6] == 4) {
        fail(f, 3, "internal: node hash table overflow");
      } else if (in[6] == 1) {
        fail(f, 3, "internal: the row-by-row classification did not finish");
      } else if (in[6] == 2) {
        fail(f, 5, "search image above its capacity (%d rows, %d bytes)", in[10], in[11]);
      } else if (in[6] == 3) {
        fail(f, 2, "colour reduction does not converge");
      } else {
        r.count = in[8];
        r.rows = in[10];
        r.image_bytes = in[11];
        r.nodes = d.M + 1;
        r.depth = d.depth;
        r.live = in[9];
        r.root_qe = hdinfo(f)[2];
        r.exact_root = in[14];
        out_[f].tree_nodes = out_[f].keep_tree ? d.M + 1 : 0;
      }
    }
    return hipSuccess;
  }

  // A copy of the node arrays for an image that keeps its tree (replacing an
  // earlier copy).
  hipError_t keep_tree(int f) {
    {
      Frame& d = H_[f];
      OctreeOut& o = out_[f];
      if (!d.ok || !o.keep_tree) return hipSuccess;
      if (o.tree_block) {
        hipError_t e = hipStreamSynchronize(s_);
        if (e == hipSuccess) e = hipFree(o.tree_block);
        if (e != hipSuccess) return e;
        o.tree_block = nullptr;
        o.tree = TreeDev{};
      }
      const size_t n = (size_t)d.M + 1;
      const size_t sz[11] = {8 * n, 64 * n, 4 * n, 4 * n, 4 * n, 8 * n,
                             8 * n, 32 * n, 8 * n, 4 * n, 4 * n};
      void* src[11] = {d.t.key, d.t.child, d.t.parent, d.t.id, d.t.level, d.t.qe,
                       d.t.share, d.t.total, d.t.nu, d.t.color_number, d.t.row};
      size_t total = 0;
      for (size_t b : sz) total += align256(b);
      char* blk = nullptr;
      hipError_t e = hipMalloc((void**)&blk, total);
      if (e != hipSuccess) return e;
      o.tree_block = blk;
      void** dst[11] = {(void**)&o.tree.key, (void**)&o.tree.child, (void**)&o.tree.parent,
                        (void**)&o.tree.id, (void**)&o.tree.level, (void**)&o.tree.qe,
                        (void**)&o.tree.share, (void**)&o.tree.total, (void**)&o.tree.nu,
                        (void**)&o.tree.color_number, (void**)&o.tree.row};
      size_t off = 0;
      for (int k = 0; k < 11; ++k) {
        *dst[k] = blk + off;
        e = hipMemcpyAsync(blk + off, src[k], sz[k], hipMemcpyDeviceToDevice, s_);
        if (e != hipSuccess) return e;
        off += align256(sz[k]);
      }
    }
    return hipSuccess;
  }

  // Node and sweep arrays of capacity `cap` for the row-by-row classification.
  hipError_t rows_arrays(RowsArgs& a, int64_t cap) {
    Carve c;
    c.want(&a.t.key, cap);
    c.want(&a.t.child, 16 * cap);
    c.want(&a.t.parent, cap);
    c.want(&a.t.id, cap);
    c.want(&a.t.level, cap);
    c.want(&a.t.qe, cap);
    c.want(&a.t.share, cap);
    c.want(&a.t.total, 4 * cap);
    c.want(&a.t.nu, cap);
    c.want(&a.t.color_number, cap);
    c.want(&a.t.row, cap);
    int32_t** wp[11] = {&a.W.lvl, &a.W.att, &a.W.anc, &a.W.size, &a.W.gsz, &a.W.start,
                        &a.W.gstart, &a.W.pre, &a.W.merged, &a.W.arr, &a.W.arr2};
    for (int k = 0; k < 11; ++k) c.want(wp[k], cap + 2);
    a.cap = cap;
    return c.take(owned_, s_);
  }

  // The classification of an image whose live node count may pass MaxQNodes,
  // row by row (k_rows), then Reduce, the colormap and the search image.
  hipError_t rows_build(int f) {
    Frame& d = H_[f];
    const int64_t R = d.R, h = d.h;
    const int64_t worst = 1 + (int64_t)kMaxDepth * R;  // every (run, level) creating a node
    int64_t cap = std::min<int64_t>(worst, std::max<int64_t>((int64_t)1 << 20, 4 * R));
    RowsArgs a{};
    a.px = d.px;
    a.count = d.count;
    a.pack = d.pack;
    a.h = h;
    a.assoc = d.assoc;
    a.kids = d.kids;
    a.maxc = d.maxc;
    a.px_pack = d.px_pack;
    a.info = d.info;
    a.dinfo = d.dinfo;
    int32_t *row_first = nullptr, *lvl_off = nullptr;
    Carve c;
    c.want(&row_first, h + 1);
    c.want(&a.at, (int64_t)(kRowLevels + 1) * kRowChunk);
    c.want(&a.rv, (int64_t)kRowLevels * kRowChunk);
    c.want(&a.lt, 4 * (int64_t)kRowChunk);
    c.want(&a.tmg, (int64_t)kRowLevels * kRowChunk);
    c.want(&a.skey, (int64_t)kRowLevels * kRowChunk);
    c.want(&a.seg, (int64_t)kRowLevels * kRowChunk);
    c.want(&a.ts, (int64_t)kRowLevels * kRowChunk);
    c.want(&a.rvs, (int64_t)kRowLevels * kRowChunk);
    c.want(&a.lts, 4 * (int64_t)kRowChunk);
    c.want(&a.scnt, (int64_t)kRowChunk);
    c.want(&a.sp, 3 * (int64_t)kStageMax);
    c.want(&a.sp_child, 16 * (int64_t)kStageMax);
    c.want(&a.fid, (int64_t)kStageMax);
    c.want(&lvl_off, 10);
    c.want(&a.st, 1);
    hipError_t e = c.take(owned_, s_);
    if (e != hipSuccess) return e;
    if ((e = rows_arrays(a, cap)) != hipSuccess) return e;
    a.row_first = row_first;
    a.W.lvl_off = lvl_off;
    RowsState st{};
    st.live = 1;
    st.loop = 1;
    st.depth = d.depth;
    if ((e = hipMemcpyAsync(a.st, &st, sizeof st, hipMemcpyHostToDevice, s_)) != hipSuccess) return e;
    k_row_first<<<blocks(R, 256), 256, 0, s_>>>(d.row, R, h, row_first);
    for (;;) {
      k_rows<<<1, kTreeThreads, 0, s_>>>(a);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if ((e = hipMemcpyAsync(&st, a.st, sizeof st, hipMemcpyDeviceToHost, s_)) != hipSuccess)
        return e;
      if ((e = hipStreamSynchronize(s_)) != hipSuccess) return e;
      if (st.status != 1 || cap >= worst) break;
      // Out of node capacity before row st.y: grow and go on from there.
      RowsArgs b = a;
      if ((e = rows_arrays(b, std::min<int64_t>(worst, 2 * cap))) != hipSuccess) return e;
      const size_t n = (size_t)st.M + 1;
      const struct {
        void* dst;
        const void* src;
        size_t bytes;
      } mv[11] = {{b.t.key, a.t.key, 8 * n},          {b.t.child, a.t.child, 64 * n},
                  {b.t.parent, a.t.parent, 4 * n},    {b.t.id, a.t.id, 4 * n},
                  {b.t.level, a.t.level, 4 * n},      {b.t.qe, a.t.qe, 8 * n},
                  {b.t.share, a.t.share, 8 * n},      {b.t.total, a.t.total, 32 * n},
                  {b.t.nu, a.t.nu, 8 * n},            {b.t.color_number, a.t.color_number, 4 * n},
                  {b.t.row, a.t.row, 4 * n}};
      for (const auto& m : mv)
        if ((e = hipMemcpyAsync(m.dst, m.src, m.bytes, hipMemcpyDeviceToDevice, s_)) != hipSuccess)
          return e;
      b.W.lvl_off = lvl_off;
      a = b;
      cap = a.cap;
    }
    if (st.status != 2) return hipSuccess;  // info[6] stays 1: reported as unfinished
    d.t = a.t;
    d.W = a.W;
    d.M = st.M;
    d.depth = st.depth;
    if ((e = upload()) != hipSuccess) return e;
    k_reduce<<<1, kTreeThreads, 0, s_>>>(d_fr_ + f);
    k_descent_one<<<blocks(desc_entries(d.kids), 256), 256, 0, s_>>>(d_fr_ + f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return keep_tree(f);
  }

  int64_t sumP_ = 0;
  int32_t* tcnt_ = nullptr;        // run starts per quarter of a tile of kRunTile pixels
  int32_t* toff_ = nullptr;        // their exclusive scan over the batch
  int32_t* slot_node_ = nullptr;
  uint32_t* pair_slot_ = nullptr;  // batch pair array: slot of an interval's first pair
  int32_t* wgc_ = nullptr;         // intervals per workgroup of k_node_insert
  int64_t nwg_ = 0;
  int64_t sumDC_ = 0;              // cells of the images placed by cell
  int64_t max_dc_ = 0;
  int64_t max_scatter_lds_ = 0;
  std::vector<int64_t> bound_ = std::vector<int64_t>(F_, 0);  // node bound per image
};

template <typename T>
void SubBuild::fill(T* p, int64_t n, T v) {
  if (n <= 0) return;
  if constexpr (sizeof(T) == 8)
    k_fill_u64<<<blocks(n, 256), 256, 0, s_>>>((uint64_t*)p, n, (uint64_t)v);
  else if constexpr (std::is_same<T, int32_t>::value)
    k_fill_i32<<<blocks(n, 256), 256, 0, s_>>>(p, n, v);
  else
    k_fill_u32<<<blocks(n, 256), 256, 0, s_>>>((uint32_t*)p, n, (uint32_t)v);
}

}  // namespace

namespace {

// A second host thread with its own stream, one per calling thread and kept
// for its life: a call with several sub-batches hands it every other one, so
// one sub-batch's narrow tree kernels and read-backs overlap the next one's
// wide kernels (32 + 32 1080p frames: 14.5 ms against 17.4 ms in turn).
class SubHelper {
 public:
  ~SubHelper() {
    if (!th_.joinable()) return;
    {
      std::lock_guard<std::mutex> g(mu_);
      quit_ = true;
    }
    cv_.notify_all();
    th_.join();
  }

  // Runs job(stream) on the helper thread, its stream on `dev` ordered after
  // `after` (an event already recorded).  Returns at once.
  void start(int dev, hipEvent_t after, std::function<void(hipStream_t)> job) {
    if (!th_.joinable()) th_ = std::thread([this] { loop(); });
    {
      std::lock_guard<std::mutex> g(mu_);
      dev_ = dev;
      after_ = after;
      job_ = std::move(job);
      busy_ = true;
    }
    cv_.notify_all();
  }

  void wait() {
    std::unique_lock<std::mutex> g(mu_);
    cv_.wait(g, [this] { return !busy_; });
  }

 private:
  void loop() {
    hipStream_t st = nullptr;
    int st_dev = -1;
    for (;;) {
      std::function<void(hipStream_t)> job;
      int dev;
      hipEvent_t after;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [this] { return quit_ || (busy_ && job_); });
        if (quit_) break;
        job = std::move(job_);
        job_ = nullptr;
        dev = dev_;
        after = after_;
      }
      (void)hipSetDevice(dev);
      if (st_dev != dev) {
        if (st) (void)hipStreamDestroy(st);
        st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) st = nullptr;
        st_dev = st ? dev : -1;
      }
      if (st) (void)hipStreamWaitEvent(st, after, 0);
      job(st);
      {
        std::lock_guard<std::mutex> g(mu_);
        busy_ = false;
      }
      cv_.notify_all();
    }
    if (st) (void)hipStreamDestroy(st);
  }

  std::thread th_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::function<void(hipStream_t)> job_;
  int dev_ = 0;
  hipEvent_t after_ = nullptr;
  bool busy_ = false, quit_ = false;
};
thread_local SubHelper tl_helper;

void run_group(const OctreeReq* rq, const std::vector<int>& group, hipStream_t s, OctreeOut* outs,
               OctreeResult* res) {
  std::vector<OctreeReq> q;
  std::vector<OctreeOut> o;
  std::vector<OctreeResult> r(group.size());
  for (int i : group) {
    q.push_back(rq[i]);
    o.push_back(outs[i]);
  }
  SubBuild(q.data(), (int)group.size(), s, o.data(), r.data()).run();
  for (size_t k = 0; k < group.size(); ++k) {
    outs[group[k]] = o[k];
    res[group[k]] = r[k];
  }
}

}  // namespace

void build_octree_batch(const OctreeReq* rq, int n, hipStream_t s, OctreeOut* outs,
                        OctreeResult* res) {
  std::vector<std::vector<int>> groups(1);
  int64_t px = 0;
  for (int i = 0; i < n; ++i) {
    res[i] = OctreeResult{};
    res[i].msg[0] = 0;
    const int64_t N = rq[i].w * rq[i].h;
    if (rq[i].w < 1 || rq[i].h < 1) {
      res[i].status = 1;
      snprintf(res[i].msg, sizeof(res[i].msg), "empty image");
      continue;
    }
    if (N > ((int64_t)1 << 28)) {
      res[i].status = 5;
      snprintf(res[i].msg, sizeof(res[i].msg), "image above 2^28 pixels");
      continue;
    }
    if (!groups.back().empty() && px + N > kBatchPixels) {
      groups.emplace_back();
      px = 0;
    }
    groups.back().push_back(i);
    px += N;
  }
  if (groups.back().empty()) groups.pop_back();
  // Sub-batches 1, 3, 5, ... on the helper's stream (after everything queued
  // on s so far), 0, 2, 4, ... here; each ends synchronised with its stream.
  int dev = 0;
  hipEvent_t ev = nullptr;
  const bool two = groups.size() > 1 && hipGetDevice(&dev) == hipSuccess &&
                   hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess &&
                   hipEventRecord(ev, s) == hipSuccess;
  if (two) {
    tl_helper.start(dev, ev, [&](hipStream_t hs) {
      for (size_t g = 1; g < groups.size(); g += 2) {
        if (hs) {
          run_group(rq, groups[g], hs, outs, res);
        } else {
          for (int i : groups[g]) {
            res[i].status = 3;
            snprintf(res[i].msg, sizeof(res[i].msg), "octree build: no second stream");
          }
        }
      }
    });
  }
  for (size_t g = 0; g < groups.size(); g += two ? 2 : 1) run_group(rq, groups[g], s, outs, res);
  if (two) tl_helper.wait();
  if (ev) (void)hipEventDestroy(ev);
}

OctreeResult build_octree(const void* src, int fmt, int64_t w, int64_t h, int maxc, int assoc,
                          int gray, hipStream_t s, OctreeOut* out) {
  const OctreeReq q{src, fmt, w, h, maxc, assoc, gray};
  OctreeResult r;
  build_octree_batch(&q, 1, s, out, &r);
  return r;
}

}  // namespace sa
