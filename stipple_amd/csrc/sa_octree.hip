// sa_octree.hip -- ColorTree::Build on the device, bit for bit.
//
// Restates src/rd_octree.cpp:219-409 (quantize.c's Classify / PruneToCubeDepth /
// Reduce / DefineImageColormap).  The O(pixels) classification is data-parallel
// integer work; the O(nodes) tree walks run as level-synchronous sweeps of one
// workgroup.  Every double the reference accumulates is accumulated here in the
// reference's order.
//
//   1. Runs.  quantize.c collapses runs of identical adjacent pixels per row and
//      weights them by their length (rd_octree.cpp:305-313).  A flag per pixel,
//      an exclusive scan, and one record per run.
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
//      are sums over its pairs in run order.  A stable radix sort of the pairs
//      by node puts each node's pairs in that order; every pair's term is
//      computed in parallel, then each node's chain is folded in order over
//      contiguous memory.  The totals of loop 2 are folded after
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
// Not handled here (the build reports SA_EUNSUPPORTED): trees whose live node
// count would exceed MaxQNodes (266817) during classification, where the
// reference starts pruning levels mid-image (rd_octree.cpp:366-369, 380-383).
#include <hipcub/hipcub.hpp>

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

// ---- 1. runs ------------------------------------------------------------------
__global__ void k_flags(const void* src, int fmt, int64_t w, int64_t n, int32_t* flags) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  int f = 1;
  if (p % w != 0) f = same_px(load_pixel(fmt, src, p), load_pixel(fmt, src, p - 1)) ? 0 : 1;
  flags[p] = f;
}

__global__ void k_run_starts(const int32_t* flags, const int32_t* rid, int64_t n, int32_t* start) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  if (flags[p]) start[rid[p]] = (int32_t)p;
}

__global__ void k_run_data(const void* src, int fmt, int64_t w, int64_t n, int assoc,
                           const int32_t* start, int64_t R, int32_t* count, double4* px,
                           uint32_t* pack, int32_t* row) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  const int64_t p = start[r];
  const int64_t e = (r + 1 < R) ? start[r + 1] : n;
  count[r] = (int32_t)(e - p);
  const float4 f = load_pixel(fmt, src, p);
  double a[4];
  assoc_pixel(assoc != 0, f.x, f.y, f.z, f.w, a);
  px[r] = make_double4(a[0], a[1], a[2], a[3]);
  const uint32_t c0 = scale_char(clamp_pixel(a[0])), c1 = scale_char(clamp_pixel(a[1]));
  const uint32_t c2 = scale_char(clamp_pixel(a[2])), c3 = scale_char(clamp_pixel(a[3]));
  pack[r] = c0 | (c1 << 8) | (c2 << 16) | (c3 << 24);
  row[r] = (int32_t)(p / w);
}

// ---- 2. loop-1 extent -----------------------------------------------------------
// info[0] <- y1 (the row after which the distinct leaf count exceeds maxc, or h)
// info[1] <- R1 (first run of loop 2, R when loop 1 covers the image)
__global__ __launch_bounds__(1024) void k_break(const uint32_t* pack, const int32_t* row,
                                                int64_t R, int64_t h, int assoc, int maxc,
                                                uint64_t* hkeys, uint32_t* hfirst, uint32_t cap,
                                                int32_t* info) {
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
      // First run of row y1+1.
      int64_t r = s_hit;
      while (r < R && row[r] == y1) ++r;
      info[1] = (int32_t)r;
    }
  }
}

// ---- 3. nodes ---------------------------------------------------------------------
__device__ __forceinline__ void pair_decode(uint32_t p, int64_t R1, int depth, int64_t* r,
                                            int* l) {
  const uint64_t l1 = 8ull * (uint64_t)R1;
  if (p < l1) {
    *r = p >> 3;
    *l = (int)(p & 7) + 1;
  } else {
    const uint64_t q = p - l1;
    *r = R1 + (int64_t)(q / depth);
    *l = (int)(q % depth) + 1;
  }
}

// Pair (r-1, l) always exists and precedes (r, l) in pair order (both loops
// classify every run down to level depth <= 8).  When the two runs share the
// level-l node, (r, l) cannot be its node's first pair, so it skips the insert
// (kNoSlot) and k_node_lookup finds its slot later without atomics.  This
// keeps the smooth regions of an image, where thousands of consecutive runs
// share their low-level nodes, from serialising on a few hash slots.
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

__global__ void k_node_insert(const uint32_t* pack, int64_t R1, int depth, int64_t P, int assoc,
                              uint64_t* hkeys, uint32_t* hfirst, uint32_t cap, uint32_t* pair_slot,
                              int32_t* info) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  int64_t r;
  int l;
  pair_decode((uint32_t)p, R1, depth, &r, &l);
  const uint64_t key = node_key(pack[r], l, assoc != 0);
  if (r > 0 && node_key(pack[r - 1], l, assoc != 0) == key) {
    pair_slot[p] = kNoSlot;
    return;
  }
  uint32_t slot = hash32(key) & (cap - 1);
  for (uint32_t probe = 0;; ++probe) {
    if (probe >= cap) {  // table full: report, never spin
      info[6] = 4;
      pair_slot[p] = 0;
      return;
    }
    const unsigned long long prev =
        atomicCAS((unsigned long long*)&hkeys[slot], (unsigned long long)kEmpty64,
                  (unsigned long long)key);
    if (prev == kEmpty64 || prev == key) break;
    slot = (slot + 1) & (cap - 1);
  }
  atomicMin(&hfirst[slot], (uint32_t)p);
  pair_slot[p] = slot;
}

// Slots of the pairs k_node_insert skipped (their key is in the table).
__global__ void k_node_lookup(const uint32_t* pack, int64_t R1, int depth, int64_t P, int assoc,
                              const uint64_t* hkeys, uint32_t cap, uint32_t* pair_slot,
                              int32_t* info) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= P || pair_slot[p] != kNoSlot) return;
  int64_t r;
  int l;
  pair_decode((uint32_t)p, R1, depth, &r, &l);
  const uint64_t key = node_key(pack[r], l, assoc != 0);
  uint32_t slot = hash32(key) & (cap - 1);
  uint32_t probe = 0;
  while (hkeys[slot] != key && probe < cap) {
    slot = (slot + 1) & (cap - 1);
    ++probe;
  }
  if (probe >= cap) {  // never inserted: report
    info[6] = 4;
    pair_slot[p] = 0;
    return;
  }
  pair_slot[p] = slot;
}

__global__ void k_node_collect(const uint64_t* hkeys, const uint32_t* hfirst, uint32_t cap,
                               int32_t* ncount, uint32_t* lfirst, uint32_t* lslot) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= cap) return;
  if (hkeys[s] == kEmpty64) return;
  const int k = atomicAdd(ncount, 1);
  lfirst[k] = hfirst[s];
  lslot[k] = (uint32_t)s;
}

// Node ids in creation order (root is 0); creation counters for the colour count.
__global__ void k_node_assign(const uint32_t* sfirst, const uint32_t* sslot, int64_t M,
                              const uint64_t* hkeys, int64_t R1, int depth, int32_t* slot_node,
                              uint64_t* nkey, int32_t* info) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= M) return;
  const int32_t node = (int32_t)(k + 1);
  const uint32_t slot = sslot[k];
  slot_node[slot] = node;
  const uint64_t key = hkeys[slot];
  nkey[node] = key;
  const int level = (int)(key >> 40);
  const bool loop1 = sfirst[k] < 8ull * (uint64_t)R1;
  if (loop1 && level == 8) atomicAdd(&info[2], 1);        // leaves counted in loop 1
  if (!loop1 && level == depth) atomicAdd(&info[3], 1);   // level-depth nodes new in loop 2
  if (loop1) atomicAdd(&info[4], 1);                      // nodes created by loop 1
}

__global__ void k_pair_nodes(const uint32_t* pair_slot, const int32_t* slot_node, int64_t P,
                             int32_t* pnode, uint32_t* pidx) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  pnode[p] = slot_node[pair_slot[p]];
  pidx[p] = (uint32_t)p;
}

__global__ void k_segments(const int32_t* snode, int64_t P, int32_t* seg_begin, int32_t* seg_end) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  const int32_t n = snode[i];
  if (i == 0 || snode[i - 1] != n) seg_begin[n] = (int32_t)i;
  if (i == P - 1 || snode[i + 1] != n) seg_end[n] = (int32_t)(i + 1);
}

__global__ void k_links(const uint64_t* nkey, int64_t M, const uint64_t* hkeys,
                        const int32_t* slot_node, uint32_t cap, TreeDev t, int32_t* info) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x + 1;
  if (k > M) return;
  const uint64_t key = nkey[k];
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
struct Runs {
  const int32_t* count;
  const double4* px;
};

struct Terms {
  double* term;   // count * |pixel - mid| per sorted pair
  float4* lval;   // clamped leaf channels (leaf pairs only)
  int32_t* lcnt;  // run length per sorted pair
};

__device__ __forceinline__ bool leaf_pair(uint32_t p, int64_t R1, int depth, int l) {
  return p < 8ull * (uint64_t)R1 ? l == 8 : l == depth;
}

__global__ void k_pair_terms(const uint64_t* nkey, Runs R, const uint32_t* spair,
                             const int32_t* snode, int64_t P, int64_t R1, int depth, int assoc,
                             Terms T) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  const uint32_t p = spair[i];
  int64_t r;
  int l;
  pair_decode(p, R1, depth, &r, &l);
  double mid[4];
  node_mid(nkey[snode[i]], mid);
  const double4 v = R.px[r];
  const int32_t cnt = R.count[r];
  const double e0 = kScale * (v.x - mid[0]);
  const double e1 = kScale * (v.y - mid[1]);
  const double e2 = kScale * (v.z - mid[2]);
  const double e3 = assoc ? kScale * (v.w - mid[3]) : 0.0;
  double d = e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3;
  if (d != d) d = 0.0;
  T.term[i] = (double)cnt * sqrt(d);
  T.lcnt[i] = cnt;
  if (leaf_pair(p, R1, depth, l))
    T.lval[i] = make_float4(clamp_pixel(v.x), clamp_pixel(v.y), clamp_pixel(v.z),
                            clamp_pixel(assoc ? v.w : kRange));
}

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

// One leaf term of total_color, as the reference forms it: count * scale * value.
__device__ __forceinline__ double tot_term(int32_t cnt, float v) {
  return (double)cnt * kScale * (double)v;
}

constexpr int kShortSeg = 256;

// L2 = false: quantize_error (+ the node's share of the root chain) over all
// pairs, and loop 1's leaf totals.  L2 = true: loop 2's leaf totals, folded onto
// what PruneToCubeDepth left.  Segments longer than kShortSeg are queued for
// k_fold_long.
template <bool L2>
__global__ void k_fold_short(TreeDev t, int64_t M, Terms T, const uint32_t* spair,
                             const int32_t* seg_b, const int32_t* seg_e, int64_t R1, int depth,
                             int32_t* long_list, int32_t* long_count) {
  const int64_t n = (int64_t)blockIdx.x * 64 + threadIdx.x + 1;
  if (n > M) return;
  const int32_t b = seg_b[n], e = seg_e[n];
  const int level = t.level[n];
  if (L2 && level != depth) return;
  if (e - b > kShortSeg) {
    long_list[atomicAdd(long_count, 1)] = (int32_t)n;
    return;
  }
  const uint32_t l1 = (uint32_t)(8 * R1);
  const int32_t s = loop2_start(spair, b, e, l1);
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

constexpr int kChunk = 512;           // terms staged per step
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

template <bool L2>
__global__ __launch_bounds__(64) void k_fold_long(TreeDev t, Terms T, const uint32_t* spair,
                                                  const int32_t* seg_b, const int32_t* seg_e,
                                                  int64_t R1, int depth, const int32_t* long_list,
                                                  const int32_t* long_count) {
  __shared__ double s_buf[2 * 4 * kChunk];
  const int32_t items = *long_count;
  const uint32_t l1 = (uint32_t)(8 * R1);
  for (int32_t it = blockIdx.x; it < items; it += gridDim.x) {
    const int32_t n = long_list[it];
    const int32_t b = seg_b[n], e = seg_e[n];
    const int level = t.level[n];
    const int32_t s = loop2_start(spair, b, e, l1);
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

// The running values of every pair, stored by pair index for the exact root
// chain (only run when the bound cannot separate the root's error).
__global__ void k_fold_record(int64_t M, const double* term, const uint32_t* spair,
                              const int32_t* seg_b, const int32_t* seg_e, double* record) {
  const int64_t n = (int64_t)blockIdx.x * 64 + threadIdx.x + 1;
  if (n > M) return;
  double qe = 0.0;
  for (int32_t i = seg_b[n]; i < seg_e[n]; ++i) {
    qe += term[i];
    record[spair[i]] = qe;
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

enum PassMode { kPruneCube = 0, kReduce = 1, kShape = 2 };

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

// One PruneToCubeDepth or Reduce walk.  Updates *live; for Reduce returns the
// colour count and the next threshold (seeded with root qe - 1).
__device__ void tree_pass(const Tree& T, TreeShared& S, int mode, double thr, int64_t* live,
                          int64_t* count, double* next) {
  const TreeDev& t = T.t;
  const TreeWork& W = T.W;
  tree_shape(T, S, mode, thr);
  int64_t mine = 0;
  for (int64_t y = threadIdx.x; y <= T.M; y += kTreeThreads) mine += (W.att[y] && W.anc[y] >= 0);
  const int64_t attempts = block_sum(mine, S);
  const int64_t K = *live - T.maxc;
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
  const int64_t merged = K > 0 ? min(attempts, K) : 0;
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
    if (threadIdx.x == 0) {
      uint64_t kk = S.sel[2], cum = 0;
      uint64_t d = 0;
      for (; d < dm; ++d) {
        if (cum + S.hist[d] > kk) break;
        cum += S.hist[d];
      }
      S.sel[2] = kk - cum;
      S.sel[0] = prefix | (d << sh);
      S.sel[1] = mask | (dm << sh);
    }
    __syncthreads();
  }
  return __longlong_as_double((long long)S.sel[0]);
}

__global__ __launch_bounds__(kTreeThreads) void k_prune_cube(Tree T, int32_t* info) {
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
    tree_pass(T, S, kPruneCube, 0.0, &live, &cnt, &nx);
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
__global__ __launch_bounds__(kTreeThreads) void k_reduce(Tree T, int assoc, double root_qe,
                                                         int32_t* info, PaletteBuild out) {
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
__global__ void k_descent(const PaletteDev* hdr, int32_t* info) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (info[6] != 0) return;
  const PaletteDev h = *hdr;
  if (e >= desc_entries(h.kids)) return;
  unsigned char* img = (unsigned char*)h.image;
  ((uint32_t*)(img + h.off_desc))[e] =
      descent3((const int16_t*)(img + h.off_child), h.kids, (uint32_t)e);
}

// Exact root chain (only when the bound cannot separate it): the running
// values of every pair, in pair order.
__global__ void k_root_exact(const double* record, int64_t P, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int64_t p = 0; p < P; ++p) s += record[p];
  *out = s;
}

__global__ void k_root_bound(TreeDev t, int64_t M, double* out) {
  // out[0] = sum of shares, out[1] = max non-root qe (one workgroup).
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
    out[0] = ssum[0];
    out[1] = smax[0];
  }
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
// Host orchestration.  Allocations are stream-ordered; the four scalar reads
// (runs, loop-1 extent, node count, root bound) synchronise the stream.
// ---------------------------------------------------------------------------------
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

#define OT_TRY(expr)                                          \
  do {                                                        \
    hipError_t e_ = (expr);                                   \
    if (e_ != hipSuccess) {                                   \
      res.status = 3;                                         \
      snprintf(res.msg, sizeof(res.msg), "%s: %s", #expr,     \
               hipGetErrorString(e_));                        \
      goto done;                                              \
    }                                                         \
  } while (0)

struct Arena {
  std::vector<void*> ptrs;
  hipStream_t s;
  template <typename T>
  hipError_t get(T** p, int64_t n) {
    void* q = nullptr;
    hipError_t e = hipMallocAsync(&q, (size_t)std::max<int64_t>(n, 1) * sizeof(T), s);
    if (e == hipSuccess) ptrs.push_back(q);
    *p = (T*)q;
    return e;
  }
  void release() {
    for (void* q : ptrs) (void)hipFreeAsync(q, s);
    ptrs.clear();
  }
};

static uint32_t pow2_at_least(int64_t v) {
  uint32_t c = 1;
  while ((int64_t)c < v) c <<= 1;
  return c;
}

OctreeResult build_octree(const void* src, int fmt, int64_t w, int64_t h, int maxc_in, int assoc_in,
                          int gray, hipStream_t s, OctreeOut* outp) {
  OctreeResult res;
  res.status = 0;
  res.msg[0] = 0;
  const bool assoc = assoc_in != 0;
  const int64_t maxc = maxc_in <= 0 ? 65536 : std::min(maxc_in, 65536);
  const int depth = octree_depth((int)maxc, assoc, gray != 0);
  const int kids = assoc ? 16 : 8;
  const int64_t N = w * h;
  Arena A{{}, s};
  int32_t *flags = nullptr, *rid = nullptr, *start = nullptr, *count = nullptr, *row = nullptr;
  double4* px = nullptr;
  uint32_t* pack = nullptr;
  int32_t* info = nullptr;
  int32_t hinfo[16] = {0};
  int64_t R = 0, R1 = 0, P = 0, M = 0;
  uint32_t bcap = 0, ncap = 0;
  uint64_t* bkeys = nullptr;
  uint32_t* bfirst = nullptr;
  uint64_t* nkeys_h = nullptr;
  uint32_t *nfirst_h = nullptr, *pair_slot = nullptr, *lfirst = nullptr, *lslot = nullptr;
  uint32_t *sfirst = nullptr, *sslot = nullptr, *pidx = nullptr, *spair = nullptr;
  int32_t *slot_node = nullptr, *pnode = nullptr, *snode = nullptr, *seg_b = nullptr, *seg_e = nullptr;
  int32_t* ncount = nullptr;
  void* cub_tmp = nullptr;
  size_t cub_bytes = 0;
  double* bound = nullptr;
  double hbound[2] = {0, 0};
  double root_qe = 0.0;
  double *flat = nullptr, *record = nullptr;
  Terms T{};
  int32_t *long_list = nullptr, *long_count = nullptr;
  TreeWork W{};
  int long_grid = 0;
  TreeDev t{};
  OctreeOut& out = *outp;
  if (N < 1) {
    res.status = 1;
    snprintf(res.msg, sizeof(res.msg), "empty image");
    return res;
  }
  if (N > ((int64_t)1 << 28)) {
    res.status = 5;
    snprintf(res.msg, sizeof(res.msg), "image above 2^28 pixels");
    return res;
  }
  // 1. runs
  OT_TRY(A.get(&flags, N));
  OT_TRY(A.get(&rid, N));
  OT_TRY(A.get(&info, 16));
  OT_TRY(hipMemsetAsync(info, 0, 16 * sizeof(int32_t), s));
  k_flags<<<blocks(N, 256), 256, 0, s>>>(src, fmt, w, N, flags);
  OT_TRY(hipGetLastError());
  OT_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, cub_bytes, flags, rid, (int)N, s));
  OT_TRY(hipMallocAsync(&cub_tmp, cub_bytes, s));
  OT_TRY(hipcub::DeviceScan::ExclusiveSum(cub_tmp, cub_bytes, flags, rid, (int)N, s));
  OT_TRY(hipFreeAsync(cub_tmp, s));
  cub_tmp = nullptr;
  {
    int32_t last[2];
    OT_TRY(hipMemcpyAsync(&last[0], rid + N - 1, 4, hipMemcpyDeviceToHost, s));
    OT_TRY(hipMemcpyAsync(&last[1], flags + N - 1, 4, hipMemcpyDeviceToHost, s));
    OT_TRY(hipStreamSynchronize(s));
    R = (int64_t)last[0] + last[1];
  }
  OT_TRY(A.get(&start, R));
  OT_TRY(A.get(&count, R));
  OT_TRY(A.get(&row, R));
  OT_TRY(A.get(&px, R));
  OT_TRY(A.get(&pack, R));
  k_run_starts<<<blocks(N, 256), 256, 0, s>>>(flags, rid, N, start);
  k_run_data<<<blocks(R, 256), 256, 0, s>>>(src, fmt, w, N, assoc, start, R, count, px, pack, row);
  OT_TRY(hipGetLastError());
  // 2. loop-1 extent
  bcap = pow2_at_least(2 * (maxc + 1 + 1024));
  OT_TRY(A.get(&bkeys, bcap));
  OT_TRY(A.get(&bfirst, bcap));
  k_fill_u64<<<blocks(bcap, 256), 256, 0, s>>>(bkeys, bcap, kEmpty64);
  k_fill_u32<<<blocks(bcap, 256), 256, 0, s>>>(bfirst, bcap, kEmpty32);
  k_break<<<1, 1024, 0, s>>>(pack, row, R, h, assoc, (int)maxc, bkeys, bfirst, bcap, info);
  OT_TRY(hipGetLastError());
  OT_TRY(hipMemcpyAsync(hinfo, info, 16 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  OT_TRY(hipStreamSynchronize(s));
  if (hinfo[6] == 4) {
    res.status = 3;
    snprintf(res.msg, sizeof(res.msg), "internal: leaf hash table overflow");
    goto done;
  }
  R1 = hinfo[1];
  P = 8 * R1 + (int64_t)depth * (R - R1);
  if (P >= ((int64_t)1 << 32) - 1) {
    res.status = 5;
    snprintf(res.msg, sizeof(res.msg), "too many runs");
    goto done;
  }
  // 3. nodes
  {
    int64_t bound_nodes = 8 * (maxc + 1 + w) + 1;
    int64_t pw = 1;
    for (int l = 1; l <= depth; ++l) {
      pw = std::min<int64_t>(pw * kids, (int64_t)1 << 40);
      bound_nodes += std::min<int64_t>(R, pw);
    }
    bound_nodes = std::min<int64_t>(bound_nodes, P + 1);
    ncap = pow2_at_least(2 * bound_nodes + 64);
  }
  OT_TRY(A.get(&nkeys_h, ncap));
  OT_TRY(A.get(&nfirst_h, ncap));
  OT_TRY(A.get(&pair_slot, P));
  k_fill_u64<<<blocks(ncap, 256), 256, 0, s>>>(nkeys_h, ncap, kEmpty64);
  k_fill_u32<<<blocks(ncap, 256), 256, 0, s>>>(nfirst_h, ncap, kEmpty32);
  k_node_insert<<<blocks(P, 256), 256, 0, s>>>(pack, R1, depth, P, assoc, nkeys_h, nfirst_h, ncap,
                                               pair_slot, info);
  k_node_lookup<<<blocks(P, 256), 256, 0, s>>>(pack, R1, depth, P, assoc, nkeys_h, ncap, pair_slot,
                                               info);
  OT_TRY(hipGetLastError());
  OT_TRY(A.get(&ncount, 1));
  OT_TRY(hipMemsetAsync(ncount, 0, 4, s));
  OT_TRY(A.get(&lfirst, ncap));
  OT_TRY(A.get(&lslot, ncap));
  k_node_collect<<<blocks(ncap, 256), 256, 0, s>>>(nkeys_h, nfirst_h, ncap, ncount, lfirst, lslot);
  OT_TRY(hipGetLastError());
  {
    int32_t m = 0;
    OT_TRY(hipMemcpyAsync(&m, ncount, 4, hipMemcpyDeviceToHost, s));
    OT_TRY(hipStreamSynchronize(s));
    M = m;
  }
  OT_TRY(A.get(&sfirst, M));
  OT_TRY(A.get(&sslot, M));
  cub_bytes = 0;
  OT_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, cub_bytes, lfirst, sfirst, lslot, sslot, (int)M,
                                            0, 32, s));
  OT_TRY(hipMallocAsync(&cub_tmp, cub_bytes, s));
  OT_TRY(hipcub::DeviceRadixSort::SortPairs(cub_tmp, cub_bytes, lfirst, sfirst, lslot, sslot, (int)M,
                                            0, 32, s));
  OT_TRY(hipFreeAsync(cub_tmp, s));
  cub_tmp = nullptr;
  // Node arrays (index 0 is the root).
  OT_TRY(A.get(&t.key, M + 1));
  OT_TRY(A.get(&t.child, 16 * (M + 1)));
  OT_TRY(A.get(&t.parent, M + 1));
  OT_TRY(A.get(&t.id, M + 1));
  OT_TRY(A.get(&t.level, M + 1));
  OT_TRY(A.get(&t.qe, M + 1));
  OT_TRY(A.get(&t.share, M + 1));
  OT_TRY(A.get(&t.total, 4 * (M + 1)));
  OT_TRY(A.get(&t.nu, M + 1));
  OT_TRY(A.get(&t.color_number, M + 1));
  OT_TRY(A.get(&t.row, M + 1));
  OT_TRY(A.get(&slot_node, ncap));
  k_fill_i32<<<blocks(16 * (M + 1), 256), 256, 0, s>>>(t.child, 16 * (M + 1), -1);
  k_fill_u32<<<blocks(M + 1, 256), 256, 0, s>>>(t.color_number, M + 1, 0u);
  k_node_assign<<<blocks(M, 256), 256, 0, s>>>(sfirst, sslot, M, nkeys_h, R1, depth, slot_node,
                                               t.key, info);
  OT_TRY(hipGetLastError());
  {
    // Root: key 0, its own parent, level 0.
    const uint64_t zero = 0;
    const int32_t z = 0;
    const double dz = 0.0;
    OT_TRY(hipMemcpyAsync(t.key, &zero, 8, hipMemcpyHostToDevice, s));
    OT_TRY(hipMemcpyAsync(t.parent, &z, 4, hipMemcpyHostToDevice, s));
    OT_TRY(hipMemcpyAsync(t.id, &z, 4, hipMemcpyHostToDevice, s));
    OT_TRY(hipMemcpyAsync(t.level, &z, 4, hipMemcpyHostToDevice, s));
    OT_TRY(hipMemcpyAsync(t.qe, &dz, 8, hipMemcpyHostToDevice, s));
    OT_TRY(hipMemcpyAsync(t.share, &dz, 8, hipMemcpyHostToDevice, s));
    OT_TRY(hipMemsetAsync(t.total, 0, 32, s));
    OT_TRY(hipMemsetAsync(t.nu, 0, 8, s));
  }
  // 4. ordered sums
  OT_TRY(A.get(&pnode, P));
  OT_TRY(A.get(&pidx, P));
  OT_TRY(A.get(&snode, P));
  OT_TRY(A.get(&spair, P));
  k_pair_nodes<<<blocks(P, 256), 256, 0, s>>>(pair_slot, slot_node, P, pnode, pidx);
  OT_TRY(hipGetLastError());
  {
    int end_bit = 1;
    while (end_bit < 32 && ((int64_t)1 << end_bit) <= M) ++end_bit;
    cub_bytes = 0;
    OT_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, cub_bytes, pnode, snode, pidx, spair, (int)P,
                                              0, end_bit, s));
    OT_TRY(hipMallocAsync(&cub_tmp, cub_bytes, s));
    OT_TRY(hipcub::DeviceRadixSort::SortPairs(cub_tmp, cub_bytes, pnode, snode, pidx, spair, (int)P,
                                              0, end_bit, s));
    OT_TRY(hipFreeAsync(cub_tmp, s));
    cub_tmp = nullptr;
  }
  OT_TRY(A.get(&seg_b, M + 1));
  OT_TRY(A.get(&seg_e, M + 1));
  OT_TRY(hipMemsetAsync(seg_b, 0, 4 * (M + 1), s));
  OT_TRY(hipMemsetAsync(seg_e, 0, 4 * (M + 1), s));
  k_segments<<<blocks(P, 256), 256, 0, s>>>(snode, P, seg_b, seg_e);
  k_links<<<blocks(M, 256), 256, 0, s>>>(t.key, M, nkeys_h, slot_node, ncap, t, info);
  OT_TRY(hipGetLastError());
  {
    Runs RR{count, px};
    OT_TRY(A.get(&T.term, P));
    OT_TRY(A.get(&T.lval, P));
    OT_TRY(A.get(&T.lcnt, P));
    k_pair_terms<<<blocks(P, 256), 256, 0, s>>>(t.key, RR, spair, snode, P, R1, depth, assoc, T);
    OT_TRY(hipGetLastError());
    OT_TRY(A.get(&long_list, M + 1));
    OT_TRY(A.get(&long_count, 2));
    OT_TRY(hipMemsetAsync(long_count, 0, 8, s));
    long_grid = (int)std::min<int64_t>(std::max<int64_t>(P / kShortSeg, 1), 2048);
    k_fold_short<false><<<blocks(M, 64), 64, 0, s>>>(t, M, T, spair, seg_b, seg_e, R1, depth,
                                                     long_list, long_count);
    k_fold_long<false><<<long_grid, 64, 0, s>>>(t, T, spair, seg_b, seg_e, R1, depth, long_list,
                                                long_count);
    OT_TRY(hipGetLastError());
    // 5a. PruneToCubeDepth, then loop 2's totals.
    {
      const int32_t hh = (int32_t)h;
      OT_TRY(hipMemcpyAsync(info + 5, &hh, 4, hipMemcpyHostToDevice, s));
    }
    for (int32_t** pp : {&W.lvl, &W.att, &W.anc, &W.size, &W.gsz, &W.start, &W.gstart, &W.pre,
                         &W.merged, &W.arr, &W.arr2})
      OT_TRY(A.get(pp, M + 2));
    OT_TRY(A.get(&W.lvl_off, 10));
    {
      const Tree TT{t, W, kids, maxc, depth, M};
      k_prune_cube<<<1, kTreeThreads, 0, s>>>(TT, info);
    }
    OT_TRY(hipMemsetAsync(long_count + 1, 0, 4, s));
    k_fold_short<true><<<blocks(M, 64), 64, 0, s>>>(t, M, T, spair, seg_b, seg_e, R1, depth,
                                                    long_list, long_count + 1);
    k_fold_long<true><<<long_grid, 64, 0, s>>>(t, T, spair, seg_b, seg_e, R1, depth, long_list,
                                               long_count + 1);
    OT_TRY(hipGetLastError());
    // The root's chain: bound, or exact when the bound is not enough.
    OT_TRY(A.get(&bound, 2));
    k_root_bound<<<1, 256, 0, s>>>(t, M, bound);
    OT_TRY(hipGetLastError());
    OT_TRY(hipMemcpyAsync(hbound, bound, 16, hipMemcpyDeviceToHost, s));
    OT_TRY(hipMemcpyAsync(hinfo, info, 16 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    OT_TRY(hipStreamSynchronize(s));
    if (hinfo[6] == 4) {
      res.status = 3;
      snprintf(res.msg, sizeof(res.msg), "internal: node hash table overflow");
      goto done;
    }
    if (hinfo[6] != 0) {
      res.status = 5;
      snprintf(res.msg, sizeof(res.msg),
               "the octree would exceed MaxQNodes (266817) during classification; the level "
               "pruning quantize.c performs then is not implemented on the device");
      goto done;
    }
    root_qe = hbound[0];
    // The approximation differs from the serial chain by at most ~3 P u sum
    // (u = 2^-53): the reference's own rounding, the running values it adds,
    // and the order of the parallel sum.  Demand a clear separation of 1 + 8 P u sum.
    const double err = (double)P * 2.3e-16 * std::fabs(root_qe) * 4.0;
    if (!(root_qe - 1.0 - err > hbound[1] + 1e-9 * hbound[1] + 2e-12) || M == 0) {
      OT_TRY(A.get(&record, P));
      k_fold_record<<<blocks(M, 64), 64, 0, s>>>(M, T.term, spair, seg_b, seg_e, record);
      OT_TRY(A.get(&flat, 1));
      k_root_exact<<<1, 1, 0, s>>>(record, P, flat);
      OT_TRY(hipGetLastError());
      OT_TRY(hipMemcpyAsync(&root_qe, flat, 8, hipMemcpyDeviceToHost, s));
      OT_TRY(hipStreamSynchronize(s));
      res.exact_root = 1;
    }
  }
  // 5b. Reduce, colormap, search image.
  {
    // The palette outlives this call: plain allocations, owned by the caller.
    const PaletteLayout cap = palette_layout((int)maxc, (int)(M + 1), kids);
    PaletteBuild pb;
    pb.image_cap = cap.bytes;
    OT_TRY(hipMalloc((void**)&pb.image, (size_t)cap.bytes));
    out.image = pb.image;
    OT_TRY(hipMalloc((void**)&pb.hdr, sizeof(PaletteDev)));
    out.hdr = pb.hdr;
    OT_TRY(hipMalloc((void**)&pb.entries, sizeof(double) * 4 * (size_t)maxc));
    out.entries = pb.entries;
    {
      const Tree TT{t, W, kids, maxc, depth, M};
      k_reduce<<<1, kTreeThreads, 0, s>>>(TT, assoc, root_qe, info, pb);
    }
    k_descent<<<blocks(desc_entries(kids), 256), 256, 0, s>>>(pb.hdr, info);
    OT_TRY(hipGetLastError());
    OT_TRY(hipMemcpyAsync(hinfo, info, 16 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    OT_TRY(hipStreamSynchronize(s));
    if (hinfo[6] == 2) {
      res.status = 5;
      snprintf(res.msg, sizeof(res.msg), "search image above its capacity (%d rows, %d bytes)",
               hinfo[10], hinfo[11]);
      goto done;
    }
    if (hinfo[6] == 3) {
      res.status = 2;
      snprintf(res.msg, sizeof(res.msg), "colour reduction does not converge");
      goto done;
    }
    res.count = hinfo[8];
    res.rows = hinfo[10];
    res.image_bytes = hinfo[11];
    res.nodes = M + 1;
    res.depth = depth;
    res.live = hinfo[9];
    res.root_qe = root_qe;
    if (out.keep_tree) {
      // Hand the node arrays to the caller (they are freed with the palette).
      out.tree = t;
      out.tree_nodes = M + 1;
      for (void** pp : {(void**)&t.key, (void**)&t.child, (void**)&t.parent, (void**)&t.id,
                        (void**)&t.level, (void**)&t.qe, (void**)&t.share, (void**)&t.total,
                        (void**)&t.nu, (void**)&t.color_number, (void**)&t.row}) {
        auto it = std::find(A.ptrs.begin(), A.ptrs.end(), *pp);
        if (it != A.ptrs.end()) A.ptrs.erase(it);
      }
    }
  }
done:
  A.release();
  return res;
}

}  // namespace sa
