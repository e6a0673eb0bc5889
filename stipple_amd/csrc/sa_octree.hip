// sa_octree.hip -- ColorTree::Build on the device, bit for bit.
//
// Restates src/rd_octree.cpp:219-409 (quantize.c's Classify / PruneToCubeDepth /
// Reduce / DefineImageColormap).  The O(pixels) classification is data-parallel
// integer work; the O(nodes) tree logic is a short sequential program on one
// device thread.  Every double the reference accumulates is accumulated here
// in the reference's order.
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
//      by node gives each node its pairs in that order; one thread per node then
//      folds them sequentially, so every double matches the reference's.  The
//      totals of loop 2 are folded after PruneToCubeDepth has merged loop 1's
//      deep leaves, exactly when the reference adds them.
//   5. Tree logic on one thread: the gated PruneToCubeDepth, the reduction
//      threshold (a restatement of glibc 2.35 qsort's merge sort, because the
//      reference's comparator is not a strict weak order), the Reduce passes,
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

__global__ void k_node_insert(const uint32_t* pack, int64_t R1, int depth, int64_t P, int assoc,
                              uint64_t* hkeys, uint32_t* hfirst, uint32_t cap, uint32_t* pair_slot,
                              int32_t* info) {
  const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= P) return;
  int64_t r;
  int l;
  pair_decode((uint32_t)p, R1, depth, &r, &l);
  const uint64_t key = node_key(pack[r], l, assoc != 0);
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
struct Runs {
  const int32_t* count;
  const double4* px;
};

__device__ __forceinline__ double pair_term(const Runs& R, int64_t r, const double mid[4],
                                            bool assoc) {
  const double4 v = R.px[r];
  const double e0 = kScale * (v.x - mid[0]);
  const double e1 = kScale * (v.y - mid[1]);
  const double e2 = kScale * (v.z - mid[2]);
  const double e3 = assoc ? kScale * (v.w - mid[3]) : 0.0;
  double d = e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3;
  if (d != d) d = 0.0;
  return (double)R.count[r] * sqrt(d);
}

// Loop 1 and loop 2 quantize_error (one chain per node), the approximate sum of
// its running values (the node's share of the root's chain), and the totals of
// loop 1's leaves.  When `record` is set, each pair's running value is stored
// for the exact root chain.
__global__ void k_node_sums(TreeDev t, int64_t M, Runs R, const uint32_t* spair,
                            const int32_t* seg_begin, const int32_t* seg_end, int64_t R1, int depth,
                            int assoc, double* record) {
  // With `record` set this only stores the running values (the node arrays
  // already hold their final, post-prune contents).
  const int64_t n = (int64_t)blockIdx.x * 64 + threadIdx.x + 1;
  if (n > M) return;
  double mid[4];
  node_mid(t.key[n], mid);
  const int level = t.level[n];
  double qe = 0.0, share = 0.0;
  double tot[4] = {0.0, 0.0, 0.0, 0.0};
  uint64_t nu = 0;
  for (int32_t i = seg_begin[n]; i < seg_end[n]; ++i) {
    const uint32_t p = spair[i];
    int64_t r;
    int l;
    pair_decode(p, R1, depth, &r, &l);
    qe += pair_term(R, r, mid, assoc != 0);
    share += qe;
    if (record) record[p] = qe;
    if (r < R1 && level == 8) {
      const double4 v = R.px[r];
      const double cnt = (double)R.count[r];
      tot[0] += cnt * kScale * (double)clamp_pixel(v.x);
      tot[1] += cnt * kScale * (double)clamp_pixel(v.y);
      tot[2] += cnt * kScale * (double)clamp_pixel(v.z);
      tot[3] += cnt * kScale * (double)clamp_pixel(assoc ? v.w : kRange);
      nu += (uint64_t)R.count[r];
    }
  }
  if (record) return;
  t.qe[n] = qe;
  t.share[n] = share;
  for (int c = 0; c < 4; ++c) t.total[4 * n + c] = tot[c];
  t.nu[n] = nu;
}

// Loop 2's leaf totals, folded onto whatever PruneToCubeDepth left.
__global__ void k_node_sums_l2(TreeDev t, int64_t M, Runs R, const uint32_t* spair,
                               const int32_t* seg_begin, const int32_t* seg_end, int64_t R1,
                               int depth, int assoc) {
  const int64_t n = (int64_t)blockIdx.x * 64 + threadIdx.x + 1;
  if (n > M) return;
  if (t.level[n] != depth) return;
  double tot[4];
  for (int c = 0; c < 4; ++c) tot[c] = t.total[4 * n + c];
  uint64_t nu = t.nu[n];
  bool any = false;
  for (int32_t i = seg_begin[n]; i < seg_end[n]; ++i) {
    int64_t r;
    int l;
    pair_decode(spair[i], R1, depth, &r, &l);
    if (r < R1) continue;
    any = true;
    const double4 v = R.px[r];
    const double cnt = (double)R.count[r];
    tot[0] += cnt * kScale * (double)clamp_pixel(v.x);
    tot[1] += cnt * kScale * (double)clamp_pixel(v.y);
    tot[2] += cnt * kScale * (double)clamp_pixel(v.z);
    tot[3] += cnt * kScale * (double)clamp_pixel(assoc ? v.w : kRange);
    nu += (uint64_t)R.count[r];
  }
  if (!any) return;
  for (int c = 0; c < 4; ++c) t.total[4 * n + c] = tot[c];
  t.nu[n] = nu;
}

// ---- 5. tree logic (one thread) -------------------------------------------------------
struct Ctl {
  TreeDev t;
  int kids;
  int64_t live;
  int64_t maxc;
  int depth;
  int count;
  double prune_thr;
  double next_thr;
};

// rd_octree.cpp:69-93 PruneChild: post-order over the current subtree, each
// node merged into its parent while the live counter exceeds the palette size.
__device__ void prune_child(Ctl& c, int32_t top, int32_t* stk_n, int8_t* stk_i) {
  int sp = 0;
  stk_n[0] = top;
  stk_i[0] = 0;
  while (sp >= 0) {
    const int32_t n = stk_n[sp];
    if (stk_i[sp] < c.kids) {
      const int32_t ch = c.t.child[16 * n + stk_i[sp]];
      ++stk_i[sp];
      if (ch >= 0) {
        ++sp;
        stk_n[sp] = ch;
        stk_i[sp] = 0;
      }
      continue;
    }
    if (c.live > c.maxc) {
      const int32_t p = c.t.parent[n];
      if (p >= 0) {
        c.t.nu[p] += c.t.nu[n];
        for (int k = 0; k < 4; ++k) c.t.total[4 * p + k] += c.t.total[4 * n + k];
        c.t.child[16 * p + c.t.id[n]] = -1;
      }
      --c.live;
    }
    --sp;
  }
}

// Post-order walk from the root calling `visit(node)` after the children; the
// child links are re-read as the walk advances, as the recursion does.
template <typename Fn>
__device__ void post_order(const Ctl& c, int32_t* stk_n, int8_t* stk_i, Fn&& visit) {
  int sp = 0;
  stk_n[0] = 0;
  stk_i[0] = 0;
  while (sp >= 0) {
    const int32_t n = stk_n[sp];
    if (stk_i[sp] < c.kids) {
      const int32_t ch = c.t.child[16 * n + stk_i[sp]];
      ++stk_i[sp];
      if (ch >= 0) {
        ++sp;
        stk_n[sp] = ch;
        stk_i[sp] = 0;
      }
      continue;
    }
    --sp;
    visit(n);
  }
}

__global__ void k_prune_cube(TreeDev t, int kids, int64_t maxc, int depth, int64_t M,
                             int32_t* info) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Ctl c{t, kids, 1 + (int64_t)info[4], maxc, depth, 0, 0.0, 0.0};
  // Loop 1's MaxQNodes test (rd_octree.cpp:366): the nodes loop 1 creates before
  // its last row are bounded by its node total.
  if (1 + (int64_t)info[4] > kMaxQNodes) info[6] = 1;
  int32_t stk_n[12], stk_x[12];
  int8_t stk_i[12], stk_j[12];
  if (info[0] < info[5]) {  // loop 1 broke: PruneToCubeDepth
    post_order(c, stk_n, stk_i, [&](int32_t n) {
      if (c.t.level[n] > c.depth) prune_child(c, n, stk_x, stk_j);
    });
  }
  // Loop 2's nodes join the live count; its MaxQNodes test (rd_octree.cpp:380)
  // is bounded by the total.
  const int64_t live = c.live + (M - (int64_t)info[4]);
  info[7] = (int32_t)live;
  if (live > kMaxQNodes) info[6] = 1;
}

// glibc 2.35 qsort for 8-byte elements: top-down merge sort, left half n/2,
// merge takes the left element when cmp(left, right) <= 0.  Restated because
// the comparator (rd_octree.cpp:23-29) is not a strict weak order, so only the
// same algorithm gives the same permutation.
__device__ int qe_cmp(double p, double q) {
  if (p > q) return 1;
  if (fabs(q - p) <= kEps) return 0;
  return -1;
}

__device__ void msort(double* b, double* tmp, int64_t n) {
  // Iterative form of the recursion: (lo, n, state) frames.
  struct F {
    int64_t lo, n;
    int st;
  };
  F stk[64];
  int sp = 0;
  stk[0] = {0, n, 0};
  while (sp >= 0) {
    F& f = stk[sp];
    if (f.n <= 1) {
      --sp;
      continue;
    }
    const int64_t n1 = f.n / 2;
    if (f.st == 0) {
      f.st = 1;
      stk[sp + 1] = {f.lo, n1, 0};
      ++sp;
      continue;
    }
    if (f.st == 1) {
      f.st = 2;
      stk[sp + 1] = {f.lo + n1, f.n - n1, 0};
      ++sp;
      continue;
    }
    // merge [lo, lo+n1) and [lo+n1, lo+n)
    int64_t i = f.lo, j = f.lo + n1, k = 0;
    const int64_t ie = f.lo + n1, je = f.lo + f.n;
    while (i < ie && j < je) {
      if (qe_cmp(b[i], b[j]) <= 0)
        tmp[k++] = b[i++];
      else
        tmp[k++] = b[j++];
    }
    while (i < ie) tmp[k++] = b[i++];
    for (int64_t q = 0; q < k; ++q) b[f.lo + q] = tmp[q];
    --sp;
  }
}

// Reduce, DefineImageColormap and the search image.  info layout in sa_octree.h.
__global__ void k_reduce(TreeDev t, int kids, int64_t maxc, int depth, int64_t M, int assoc,
                         double root_qe, double* flat, double* tmp, int32_t* info,
                         PaletteBuild out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Ctl c{t, kids, (int64_t)info[7], maxc, depth, info[2] + info[3], 0.0, 0.0};
  t.qe[0] = root_qe;
  int32_t stk_n[12], stk_x[12];
  int8_t stk_i[12], stk_j[12];
  // ReduceImageColors (rd_octree.cpp:388-404).
  if (c.count > c.maxc) {
    c.next_thr = 0.0;
    if (c.live > 0) {
      const int64_t L = c.live;
      for (int64_t k = 0; k < L; ++k) flat[k] = 0.0;
      // Pre-order flattening, cut off at the live counter (rd_octree.cpp:129-140).
      int64_t off = 0;
      int sp = 0;
      stk_n[0] = 0;
      stk_i[0] = -1;
      while (sp >= 0 && off < L) {
        const int32_t n = stk_n[sp];
        if (stk_i[sp] < 0) {
          flat[off++] = t.qe[n];
          stk_i[sp] = 0;
          continue;
        }
        if (stk_i[sp] < kids) {
          const int32_t ch = t.child[16 * n + stk_i[sp]];
          ++stk_i[sp];
          if (ch >= 0) {
            ++sp;
            stk_n[sp] = ch;
            stk_i[sp] = -1;
          }
          continue;
        }
        --sp;
      }
      msort(flat, tmp, L);
      const int64_t aggressive = (110 * (maxc + 1)) / 100;
      if (L > aggressive) c.next_thr = flat[L - aggressive];
    }
  }
  int passes = 0;
  while (c.count > c.maxc) {
    // The reference loops forever if a pass prunes nothing; stop instead.
    if (++passes > 100000) {
      info[6] = 3;
      return;
    }
    c.prune_thr = c.next_thr;
    c.next_thr = t.qe[0] - 1.0;
    c.count = 0;
    post_order(c, stk_n, stk_i, [&](int32_t n) {
      if (c.t.qe[n] <= c.prune_thr) {
        prune_child(c, n, stk_x, stk_j);
      } else {
        if (c.t.nu[n] > 0) ++c.count;
        if (c.t.qe[n] < c.next_thr) c.next_thr = c.t.qe[n];
      }
    });
  }
  // DefineImageColormap (rd_octree.cpp:142-176).
  int cnt = 0;
  post_order(c, stk_n, stk_i, [&](int32_t n) {
    if (t.nu[n] == 0) return;
    double* e = out.entries + 4 * cnt;
    const double alpha = safe_recip((double)t.nu[n]);
    const double* tot = t.total + 4 * n;
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
    t.color_number[n] = (uint32_t)cnt;
    ++cnt;
  });
  info[8] = cnt;
  info[9] = (int32_t)c.live;
  // Search image: rows for internal reachable nodes (pre-order), candidate
  // ranges from the post-order of colour-bearing nodes (see sa_palette.h).
  int32_t rows = 0;
  {
    int sp = 0;
    stk_n[0] = 0;
    stk_i[0] = -1;
    while (sp >= 0) {
      const int32_t n = stk_n[sp];
      if (stk_i[sp] < 0) {
        bool internal = (n == 0);
        for (int k = 0; k < kids; ++k) internal |= t.child[16 * n + k] >= 0;
        t.row[n] = internal ? rows++ : -1;
        stk_i[sp] = 0;
        continue;
      }
      if (stk_i[sp] < kids) {
        const int32_t ch = t.child[16 * n + stk_i[sp]];
        ++stk_i[sp];
        if (ch >= 0) {
          ++sp;
          stk_n[sp] = ch;
          stk_i[sp] = -1;
        }
        continue;
      }
      --sp;
    }
  }
  const PaletteLayout L = palette_layout(cnt, rows, kids);
  info[10] = rows;
  info[11] = L.bytes;
  if (L.bytes > out.image_cap || rows > 32767) {
    info[6] = 2;
    return;
  }
  unsigned char* img = out.image;
  double* col = (double*)(img + L.off_col);
  float* ent = (float*)(img + L.off_ent);
  float* cent = (float*)(img + L.off_cent);
  int32_t* sub = (int32_t*)(img + L.off_sub);
  int16_t* child = (int16_t*)(img + L.off_child);
  uint16_t* cand = (uint16_t*)(img + L.off_cand);
  for (int i = 0; i < cnt; ++i) {
    const double* e = out.entries + 4 * i;
    for (int k = 0; k < 4; ++k) ent[4 * i + k] = (float)e[k];
    assoc_entry(assoc != 0, e, col + 4 * i);
  }
  // Post-order with subtree starts.
  int32_t ncand = 0;
  {
    int32_t first[12];
    int sp = 0;
    stk_n[0] = 0;
    stk_i[0] = 0;
    first[0] = 0;
    while (sp >= 0) {
      const int32_t n = stk_n[sp];
      if (stk_i[sp] < kids) {
        const int32_t ch = t.child[16 * n + stk_i[sp]];
        ++stk_i[sp];
        if (ch >= 0) {
          ++sp;
          stk_n[sp] = ch;
          stk_i[sp] = 0;
          first[sp] = ncand;
        }
        continue;
      }
      if (t.nu[n] != 0) {
        const uint32_t cn = t.color_number[n];
        for (int k = 0; k < 4; ++k) cent[4 * ncand + k] = ent[4 * cn + k];
        cand[ncand++] = (uint16_t)cn;
      }
      const int32_t r = t.row[n];
      if (r >= 0) {
        sub[2 * r] = first[sp];
        sub[2 * r + 1] = ncand - first[sp];
        for (int k = 0; k < kids; ++k) {
          const int32_t ch = t.child[16 * n + k];
          child[r * kids + k] = ch < 0 ? kNoChild : (t.row[ch] >= 0 ? (int16_t)t.row[ch] : kLeaf);
        }
      }
      --sp;
    }
  }
  // The descent table is filled by k_descent (one thread per entry).
  PaletteDev* h = out.hdr;
  h->count = cnt;
  h->assoc = assoc;
  h->kids = kids;
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
  (void)M;
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
  double *flat = nullptr, *tmp = nullptr, *record = nullptr;
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
    k_node_sums<<<blocks(M, 64), 64, 0, s>>>(t, M, RR, spair, seg_b, seg_e, R1, depth, assoc,
                                             nullptr);
    OT_TRY(hipGetLastError());
    // 5a. PruneToCubeDepth, then loop 2's totals.
    {
      const int32_t hh = (int32_t)h;
      OT_TRY(hipMemcpyAsync(info + 5, &hh, 4, hipMemcpyHostToDevice, s));
    }
    k_prune_cube<<<1, 1, 0, s>>>(t, kids, maxc, depth, M, info);
    k_node_sums_l2<<<blocks(M, 64), 64, 0, s>>>(t, M, RR, spair, seg_b, seg_e, R1, depth, assoc);
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
    // |exact - approx| <= P * 2^-52 * sum; demand a clear separation of 1 + that.
    const double err = (double)P * 2.3e-16 * std::fabs(root_qe) * 4.0;
    if (!(root_qe - 1.0 - err > hbound[1] + 1e-9 * hbound[1] + 2e-12) || M == 0) {
      OT_TRY(A.get(&record, P));
      k_node_sums<<<blocks(M, 64), 64, 0, s>>>(t, M, RR, spair, seg_b, seg_e, R1, depth, assoc,
                                               record);
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
    const int64_t live = hinfo[7];
    OT_TRY(A.get(&flat, std::max<int64_t>(live, 1)));
    OT_TRY(A.get(&tmp, std::max<int64_t>(live, 1)));
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
    k_reduce<<<1, 1, 0, s>>>(t, kids, maxc, depth, M, assoc, root_qe, flat, tmp, info, pb);
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
