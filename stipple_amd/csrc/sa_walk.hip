// sa_walk.hip -- the exact sequential Riemersma walk, batched.
//
// Restates RiemersmaWalkCpu (src/rd_riemersma_cpu.cpp:39-113, :359-417), which
// is the walk quantize.c:RiemersmaDither performs.  Bit-exact: every double
// operation is the reference's, in the reference's order, one rounding each
// (built with -ffp-contract=off).
//
// Mapping to CDNA4.  The walk is a serial chain, so one wavefront walks one
// image and many images walk at once: a persistent grid pulls images from a
// device queue, largest first.  Inside the wavefront the 64 lanes are four DPP
// rows of 16; row c holds channel c (r, g, b, a) and lane j of a row holds the
// running sum of visit t+j.  Visit t's diffusion is
//
//     p_t = x_t + w0*e_{t-16} + w1*e_{t-15} + ... + w15*e_{t-1}
//
// summed left to right.  Only the last term depends on the visit just
// finished, so when e_{t-1} is known every in-flight visit adds its own tap of
// it -- lane j adds w_{15-j} -- in ONE v_mul_f64 + ONE v_add_f64 for all
// channels.  Lane 0 is then complete; after the visit every row rotates left by
// one lane (DPP row_ror:15) and lane 15 restarts with x_{t+16}, taken from a
// register chunk of the next inputs that rotates in step (prefetched two
// chunks ahead with coalesced loads).  The serial chain per visit is one mul,
// one add, the clamp, the memo probe and the residual.
//
// Memo (quantize.c's cache, src/rd_riemersma_cpu.cpp:75-94).  First-touch
// semantics are observable, so it is reproduced exactly: a zeroed per-image
// table in HBM (entry = index+1, 0 = untouched) is the authority, and an LDS
// cache of its 64-entry lines (direct-mapped, dirty lines written back on
// eviction) serves the walk.  Entries are stored under memo_store_key, so a
// line is a 4x4x4 cube of colour codes and the targets of neighbouring visits
// share lines.  A line whose slot has never held another line is known to be
// all zero and is filled without touching HBM.
//
// Palette search on a miss: the search image (sa_palette.h) is staged in LDS
// once per image; the first three descent levels are one table read, the
// rest are uniform LDS reads, and the candidate range is scored across lanes
// with a DPP min-reduction (the last index at the minimum wins, which is the
// reference's `<=` rule).
#include "sa_internal.h"
#include "sa_math.h"

namespace sa {
namespace {

constexpr int kWave = 64;

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ double rl(double v, int lane) {
  return __hiloint2double(rl(__double2hiint(v), lane), rl(__double2loint(v), lane));
}
__device__ __forceinline__ int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <int CTRL>
__device__ __forceinline__ double dpp(double v, double fill) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

constexpr int kRowBcast0 = 0x150;  // row_newbcast:0
constexpr int kRowRor15 = 0x12F;   // lane j <- lane (j+1) mod 16
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;

// Minimum over the wavefront of a per-lane double (lanes outside the
// candidate range pass +inf).  Rows by DPP shift-scans, then the four row
// minima.
__device__ __forceinline__ double wave_min(double v, bool one_row) {
  const double inf = __builtin_inf();
  v = fmin(v, dpp<kRowShr1>(v, inf));
  v = fmin(v, dpp<kRowShr2>(v, inf));
  v = fmin(v, dpp<kRowShr4>(v, inf));
  v = fmin(v, dpp<kRowShr8>(v, inf));
  if (one_row) return rl(v, 15);
  return fmin(fmin(rl(v, 15), rl(v, 31)), fmin(rl(v, 47), rl(v, 63)));
}

struct Search {
  const double* col;     // 4 per entry
  const float4* cent;    // candidates in post-order
  const int2* sub;
  const uint32_t* desc;
  const int16_t* child;
  const uint16_t* cand;
  int kids;
};

// Descent + ordered scan for one clamped target (src/rd_riemersma_cpu.cpp:82-92,
// src/rd_octree.cpp:178-217); uniform result.
template <bool ASSOC>
__device__ __forceinline__ int lookup(const Search& s, unsigned c0, unsigned c1, unsigned c2,
                                      unsigned c3, double t0, double t1, double t2, double t3) {
  const int lane = threadIdx.x;
  uint32_t e = (c0 >> 5) | ((c1 >> 5) << 3) | ((c2 >> 5) << 6);
  if (ASSOC) e |= (c3 >> 5) << 9;
  const uint32_t d3 = (uint32_t)rfl((int)s.desc[e]);
  int prev = (int)((d3 >> 16) & 0x7FFFu);
  if (!(d3 & kDescStop)) {
    int row = (int)(d3 & 0x7FFFu);
    for (int lev = 4; lev > 0; --lev) {
      const int ch = rfl((int)s.child[row * s.kids + node_id(ASSOC, c0, c1, c2, c3, lev)]);
      if (ch == kNoChild) break;
      prev = row;
      if (ch == kLeaf) break;
      row = ch;
    }
  }
  const int2 r = s.sub[prev];
  const int off = rfl(r.x), len = rfl(r.y);
  const double t[4] = {t0, t1, t2, t3};
  double best = kFarDistance;
  int bidx = 0;
  for (int b = 0; b < len; b += kWave) {
    const int k = b + lane;
    const bool in = k < len;
    double d = __builtin_inf();
    int ci = 0;
    if (in) {
      const float4 e4 = s.cent[off + k];
      ci = (int)s.cand[off + k];
      const float ev[4] = {e4.x, e4.y, e4.z, e4.w};
      d = candidate_distance(ASSOC, ev, t);
    }
    const double m = wave_min(d, len - b <= 16);
    if (m <= best) {
      const unsigned long long mask = __ballot(in && d == m);
      const int last = 63 - __builtin_clzll(mask);
      best = m;
      bidx = rl(ci, last);
    }
  }
  return bidx;
}

// Global-address-space views: loads and stores through them are
// global_load/global_store (vmcnt only), so they never hold up an LDS wait the
// way a flat access would.
typedef const double __attribute__((address_space(1))) gdouble;
typedef uint16_t __attribute__((address_space(1))) gu16;

// LDS cache of the HBM memo table's 64-entry lines.
struct Memo {
  uint32_t* tag;    // per slot: bit 31 empty, bit 30 dirty, bits 0..29 line id
  uint16_t* line;   // 64 entries per slot: index+1, 0 = untouched
  uint16_t* table;  // HBM, 1 << K entries, zeroed before the walk
  int bits;         // log2(slots)
  // This lane's channel part of the key (memo_key_part for its row):
  // ((c8 >> k_hi_shr) & k_hi_mask) << k_hi_shl | ((c8 >> 2) & k_lo_mask) << k_lo_shl.
  uint32_t k_hi_shr, k_hi_mask, k_hi_shl, k_lo_mask, k_lo_shl;
};

// Rows 0..2 (r, g, b) give their top 4 bits to the line and bits 2..3 to the
// offset; row 3 (alpha) gives its top 6 bits to the line when associated.
template <bool ASSOC>
__device__ __forceinline__ void memo_key_consts(Memo& m, int ch) {
  m.k_hi_shr = ch == 3 ? 2u : 4u;
  m.k_hi_mask = ch == 3 ? (ASSOC ? 63u : 0u) : 15u;
  m.k_hi_shl = ch == 3 ? 18u : (uint32_t)(6 + 4 * ch);
  m.k_lo_mask = ch == 3 ? 0u : 3u;
  m.k_lo_shl = ch == 3 ? 0u : (uint32_t)(2 * ch);
}

constexpr uint32_t kEmptyTag = 0x80000000u;
constexpr uint32_t kDirty = 0x40000000u;

__device__ __forceinline__ uint32_t memo_slot(const Memo& m, uint32_t line) {
  return (line * 0x9E3779B1u) >> (32 - m.bits);
}

// Makes `line` resident in `slot` (uniform).  Returns 1 when it read HBM.
__device__ __forceinline__ uint32_t memo_fill(const Memo& m, uint32_t slot, uint32_t line,
                                              uint32_t old) {
  const int lane = threadIdx.x;
  uint16_t* dst = m.line + 64 * slot;
  uint32_t loaded = 0;
  if (old & kEmptyTag) {
    // Never held a line: this line was never written, so it is all zero.
    dst[lane] = 0;
  } else {
    if (old & kDirty) ((gu16*)m.table)[64 * (old & 0x3FFFFFFFu) + lane] = dst[lane];
    dst[lane] = ((const gu16*)m.table)[64 * line + lane];
    loaded = 1;
  }
  if (lane == 0) m.tag[slot] = line;
  __builtin_amdgcn_wave_barrier();
  return loaded;
}

struct Counters {
  uint32_t misses = 0, fills = 0, loads = 0;
};


// Inputs of visit v (clamped into the image, so the load never branches and
// the wave does not wait for it until the value is used).
__device__ __forceinline__ double load_x(const double* xs, int64_t v, int64_t n, int ch) {
  return ((gdouble*)xs)[4 * (v < n ? v : n - 1) + ch];
}

struct WalkState {
  double acc;     // running sum of this lane's in-flight visit
  double e;       // residual of the visit completed last (lane 0 of each row)
  uint32_t obuf;  // palette indices of the current 64-visit window
};

// One visit (visit t); on return lane j holds visit t+1+j.  `xq` is the
// current chunk's input register, rotating in step with the accumulators.
template <bool ASSOC, bool CACHE>
__device__ __forceinline__ void visit(WalkState& w, double& xq, double W, int64_t t, int64_t n,
                                      uint16_t* idx_out, const Memo& m, const Search& s,
                                      Counters& cnt) {
  const int lane = threadIdx.x;
  const int j = lane & 15;
  const int ch = lane >> 4;
  // Every in-flight visit adds its tap of e_{t-1} (held by lane 0).
  w.acc = w.acc + W * dpp<kRowBcast0>(w.e, 0.0);
  const float pc = clamp_pixel(w.acc);
  const double pd = (double)pc;
  const unsigned c8 = scale_char_clamped(pc);
  // The channel bytes of visit t sit in lanes 0, 16, 32, 48.
  auto chan = [&]() {
    struct C4 {
      unsigned c0, c1, c2, c3;
    };
    return C4{(unsigned)rl((int)c8, 0), (unsigned)rl((int)c8, 16), (unsigned)rl((int)c8, 32),
              ASSOC ? (unsigned)rl((int)c8, 48) : 0u};
  };
  int idx;
  if (CACHE) {
    // Each row forms its channel's part of the memo key (memo_key_part with
    // this lane's constants: no branch); four lane reads and three ORs join
    // them (the channel bytes themselves are read only on a miss).
    const uint32_t kc = (((c8 >> m.k_hi_shr) & m.k_hi_mask) << m.k_hi_shl) |
                        (((c8 >> 2) & m.k_lo_mask) << m.k_lo_shl);
    uint32_t key = (uint32_t)rl((int)kc, 0) | (uint32_t)rl((int)kc, 16) | (uint32_t)rl((int)kc, 32);
    if (ASSOC) key |= (uint32_t)rl((int)kc, 48);
    const uint32_t line = key >> 6;
    const uint32_t slot = memo_slot(m, line);
    const uint32_t tg = (uint32_t)rfl((int)m.tag[slot]);
    uint32_t ent = (uint32_t)rfl((int)m.line[64 * slot + (key & 63)]);
    if (__builtin_expect((tg & ~kDirty) != line, 0)) {
      ++cnt.fills;
      cnt.loads += memo_fill(m, slot, line, tg);
      ent = (uint32_t)rfl((int)m.line[64 * slot + (key & 63)]);
    }
    if (__builtin_expect(ent != 0, 1)) {
      idx = (int)ent - 1;
    } else {
      ++cnt.misses;
      const auto c = chan();
      idx = lookup<ASSOC>(s, c.c0, c.c1, c.c2, c.c3, rl(pd, 0), rl(pd, 16), rl(pd, 32),
                          ASSOC ? rl(pd, 48) : 65535.0);
      if (lane == 0) {
        m.line[64 * slot + (key & 63)] = (uint16_t)(idx + 1);
        m.tag[slot] = line | kDirty;
      }
    }
  } else {
    ++cnt.misses;
    const auto c = chan();
    idx = lookup<ASSOC>(s, c.c0, c.c1, c.c2, c.c3, rl(pd, 0), rl(pd, 16), rl(pd, 32),
                        ASSOC ? rl(pd, 48) : 65535.0);
  }
  // Residual against the associated palette colour, per row's channel.
  w.e = pd - s.col[4 * idx + ch];
  w.obuf = lane == (int)(t & 63) ? (uint32_t)idx : w.obuf;
  if ((t & 63) == 63 || t + 1 == n) {
    const int64_t v = (t & ~(int64_t)63) + lane;
    if (v <= t) ((gu16*)idx_out)[v] = (uint16_t)w.obuf;
  }
  // Rotate: lane j takes visit t+1+j; lane 15 restarts with x_{t+16}.  (+0.0
  // is that visit's zero-error taps: it turns -0.0 into +0.0 as the
  // reference's first adds do.)
  w.acc = dpp<kRowRor15>(w.acc, 0.0);
  xq = dpp<kRowRor15>(xq, 0.0);
  w.acc = j == 15 ? xq + 0.0 : w.acc;
}

// Chunk c (visits 16c .. 16c+15) restarts from its register `xq`, loaded two
// chunks earlier; afterwards `xq` is refilled for chunk c+3.  Returns false
// when `stop` (the image's end, or where this launch's slice ends) is reached.
template <bool ASSOC, bool CACHE>
__device__ __forceinline__ bool chunk(WalkState& w, double& xq, double W, int64_t base, int64_t n,
                                      int64_t stop, const SeqItem& it, const Memo& m,
                                      const Search& s, Counters& cnt) {
  const int lane = threadIdx.x;
  const int64_t end = base + 16 < stop ? base + 16 : stop;
  for (int64_t t = base; t < end; ++t) visit<ASSOC, CACHE>(w, xq, W, t, n, it.idx, m, s, cnt);
  xq = load_x(it.xs, base + 64 + (lane & 15), n, lane >> 4);
  return end < stop;
}

// Visits [t0, stop).  t0 and stop - t0 are multiples of 48 unless stop is the
// end, so the three chunk registers restart in their usual roles.  A walk that
// stops early leaves its queue (the in-flight sums and residuals) in it.state
// and flushes the indices of its last, partial 64-visit window; resuming loads
// them back, so a sliced walk computes exactly what an uninterrupted one does.
// `deadline` (a s_memrealtime tick, 0: none) ends the walk at the first 48-visit
// boundary past it.  Returns where the walk stopped.
template <bool ASSOC, bool CACHE>
__device__ int64_t walk_image(const SeqItem& it, double W, const Memo& m, const Search& s,
                              Counters& cnt, int64_t t0, int64_t stop, uint64_t deadline) {
  const int lane = threadIdx.x;
  const int j = lane & 15;
  const int ch = lane >> 4;
  const int64_t n = it.n;
  WalkState w;
  if (t0 == 0) {
    w.e = 0.0;
    w.obuf = 0;
    w.acc = load_x(it.xs, j, n, ch) + 0.0;
  } else {
    w.acc = ((gdouble*)it.state->acc)[lane];
    w.e = ((gdouble*)it.state->e)[lane];
    const int64_t v = (t0 & ~(int64_t)63) + lane;
    w.obuf = v < t0 ? (uint32_t)((const gu16*)it.idx)[v] : 0u;
    // Settle these loads here, before the prefetches below: left pending, the
    // compiler waits for them at the top of every visit of the first chunk,
    // which then also drains the prefetches once per 48 visits.
    __builtin_amdgcn_s_waitcnt(0);
  }
  double x0 = load_x(it.xs, t0 + 16 + j, n, ch);
  double x1 = load_x(it.xs, t0 + 32 + j, n, ch);
  double x2 = load_x(it.xs, t0 + 48 + j, n, ch);
  // Three chunk registers used round-robin: no register copy ever waits on
  // an outstanding load.
  for (int64_t base = t0;; base += 48) {
    if (deadline && base > t0 && __builtin_amdgcn_s_memrealtime() >= deadline) {
      stop = base;
      break;
    }
    if (!chunk<ASSOC, CACHE>(w, x0, W, base, n, stop, it, m, s, cnt)) break;
    if (!chunk<ASSOC, CACHE>(w, x1, W, base + 16, n, stop, it, m, s, cnt)) break;
    if (!chunk<ASSOC, CACHE>(w, x2, W, base + 32, n, stop, it, m, s, cnt)) break;
  }
  if (stop < n) {
    const int64_t last = stop - 1;
    if ((last & 63) != 63) {
      const int64_t v = (last & ~(int64_t)63) + lane;
      if (v <= last) ((gu16*)it.idx)[v] = (uint16_t)w.obuf;
    }
    it.state->acc[lane] = w.acc;
    it.state->e[lane] = w.e;
  }
  return stop;
}

// Wavefront copy of `words` 16-byte words with eight loads in flight per
// lane (the palette and memo snapshot moves at the start and end of a walk).
__device__ __forceinline__ void copy16(uint4* dst, const uint4* src, uint32_t words) {
  const uint32_t lane = threadIdx.x;
  for (uint32_t k = lane; k < words; k += 8 * kWave) {
    uint4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t q = k + u * kWave;
      v[u] = q < words ? src[q] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint32_t q = k + u * kWave;
      if (q < words) dst[q] = v[u];
    }
  }
}

template <bool ASSOC, bool CACHE>
__global__ __launch_bounds__(kWave) void seq_walk_kernel(const SeqItem* __restrict__ items,
                                                         int nitems, int32_t* queue,
                                                         const double* __restrict__ sw,
                                                         int64_t slice, uint64_t budget,
                                                         uint32_t* stats) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = threadIdx.x;
  // Lane j adds tap 15-j of the newest residual.
  const double W = sw[15 - (lane & 15)];
  Counters cnt;
  for (;;) {
    int q = 0;
    if (lane == 0) q = atomicAdd(queue, 1);
    q = rfl(__shfl(q, 0, kWave));  // uniform: item fields load into SGPRs
    if (q >= nitems) break;
    const SeqItem it = items[q];
    const int64_t t0 = it.state ? it.state->t : 0;
    if (t0 >= it.n) continue;
    int64_t stop = (slice > 0 && it.state && t0 + slice < it.n) ? t0 + slice : it.n;
    Memo m;
    m.bits = it.memo_bits;
    memo_key_consts<ASSOC>(m, lane >> 4);
    const uint32_t slots = 1u << m.bits;
    const uint32_t memo_bytes = CACHE ? 132u * slots : 0u;
    m.tag = (uint32_t*)lds;
    m.line = (uint16_t*)(lds + 4 * slots);
    unsigned char* img = lds + memo_bytes;
    const PaletteDev pd = *it.pal;
    {
      copy16((uint4*)img, (const uint4*)pd.image, (uint32_t)pd.bytes / 16);
      if (CACHE) {
        if (t0 == 0) {
          for (uint32_t k = lane; k < slots; k += kWave) m.tag[k] = kEmptyTag;
        } else {  // resume: the memo cache exactly as the last launch left it
          copy16((uint4*)lds, (const uint4*)it.snap, memo_bytes / 16);
        }
      }
      __syncthreads();
    }
    Search s;
    s.col = (const double*)img;
    s.cent = (const float4*)(img + pd.off_cent);
    s.sub = (const int2*)(img + pd.off_sub);
    s.desc = (const uint32_t*)(img + pd.off_desc);
    s.child = (const int16_t*)(img + pd.off_child);
    s.cand = (const uint16_t*)(img + pd.off_cand);
    s.kids = pd.kids;
    m.table = it.memo;
    const uint64_t deadline =
        (budget && it.state) ? __builtin_amdgcn_s_memrealtime() + budget : 0;
    stop = walk_image<ASSOC, CACHE>(it, W, m, s, cnt, t0, stop, deadline);
    __syncthreads();
    if (it.state) {
      if (CACHE && stop < it.n) copy16((uint4*)it.snap, (const uint4*)lds, memo_bytes / 16);
      if (lane == 0) it.state->t = stop;
    }
    __syncthreads();
  }
  if (stats && lane == 0) {
    atomicAdd(&stats[0], cnt.misses);
    atomicAdd(&stats[1], cnt.fills);
    atomicAdd(&stats[2], cnt.loads);
  }
}

// Progress of resumable walks, gathered for one copy to the host.
__global__ void seq_progress_kernel(const SeqItem* __restrict__ items, int nitems, int64_t* out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i < nitems) out[i] = items[i].state ? items[i].state->t : items[i].n;
}

// Resumable walks start at visit 0.
__global__ void seq_state_init_kernel(const SeqItem* __restrict__ items, int nitems) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i < nitems && items[i].state) items[i].state->t = 0;
}

// Zeroes each image's HBM memo table (a fill kernel, so the sequence can be
// captured in a graph).
__global__ __launch_bounds__(256) void memo_clear_kernel(const SeqItem* __restrict__ items,
                                                         int64_t words) {
  const SeqItem it = items[blockIdx.y];
  uint4* t = (uint4*)it.memo;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < words; k += (int64_t)gridDim.x * 256)
    t[k] = make_uint4(0, 0, 0, 0);
}

// Gather: visit order, associated, +0.0 (the zero taps a visit near the start
// adds turn -0.0 into +0.0, exactly as the reference's first adds do).
__global__ __launch_bounds__(256) void seq_gather_kernel(const SeqItem* __restrict__ items) {
  const SeqItem it = items[blockIdx.y];
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= it.n) return;
  const bool assoc = it.pal->assoc != 0;
  const float4 p = load_pixel(it.src_fmt, it.src, it.curve[v]);
  double x[4];
  assoc_pixel(assoc, p.x, p.y, p.z, p.w, x);
  double4 o;
  o.x = x[0] + 0.0;
  o.y = x[1] + 0.0;
  o.z = x[2] + 0.0;
  o.w = x[3] + 0.0;
  ((double4*)it.xs)[v] = o;
}

// Pack: palette index per visit -> output pixel at the visited position.
// RgbaF32 output follows RiemersmaWalkCpu: r,g,b = (float)entry, alpha only
// when associated (otherwise the source alpha stays).
__global__ __launch_bounds__(256) void seq_pack_kernel(const SeqItem* __restrict__ items) {
  const SeqItem it = items[blockIdx.y];
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= it.n) return;
  const PaletteDev* pal = it.pal;
  const int idx = it.idx[v];
  const int64_t p = it.curve[v];
  const float4 e = ((const float4*)((const unsigned char*)pal->image + pal->off_ent))[idx];
  const bool assoc = pal->assoc != 0;
  switch (it.dst_fmt) {
    case kIndex8: ((uint8_t*)it.dst)[p] = (uint8_t)idx; break;
    case kIndex16: ((uint16_t*)it.dst)[p] = (uint16_t)idx; break;
    case kRgb8: {
      uint8_t* d = (uint8_t*)it.dst + 3 * p;
      d[0] = (uint8_t)scale_char(e.x);
      d[1] = (uint8_t)scale_char(e.y);
      d[2] = (uint8_t)scale_char(e.z);
      break;
    }
    case kRgba8: {
      const float a = assoc ? e.w : load_pixel(it.src_fmt, it.src, p).w;
      ((uint32_t*)it.dst)[p] = scale_char(e.x) | (scale_char(e.y) << 8) |
                               (scale_char(e.z) << 16) | (scale_char(a) << 24);
      break;
    }
    case kRgb16: {
      uint16_t* d = (uint16_t*)it.dst + 3 * p;
      d[0] = (uint16_t)e.x;
      d[1] = (uint16_t)e.y;
      d[2] = (uint16_t)e.z;
      break;
    }
    case kRgba16: {
      const float a = assoc ? e.w : load_pixel(it.src_fmt, it.src, p).w;
      ushort4 o;
      o.x = (uint16_t)e.x;
      o.y = (uint16_t)e.y;
      o.z = (uint16_t)e.z;
      o.w = (uint16_t)a;
      ((ushort4*)it.dst)[p] = o;
      break;
    }
    default: {
      float4* d = (float4*)it.dst + p;
      float a = assoc ? e.w : load_pixel(it.src_fmt, it.src, p).w;
      *d = make_float4(e.x, e.y, e.z, a);
      break;
    }
  }
}

}  // namespace

hipError_t launch_seq_gather(const SeqItem* items, int nitems, int64_t max_n, hipStream_t s) {
  if (nitems <= 0 || max_n <= 0) return hipSuccess;
  seq_gather_kernel<<<dim3((unsigned)((max_n + 255) / 256), (unsigned)nitems), dim3(256), 0, s>>>(
      items);
  return hipGetLastError();
}

hipError_t launch_seq_pack(const SeqItem* items, int nitems, int64_t max_n, hipStream_t s) {
  if (nitems <= 0 || max_n <= 0) return hipSuccess;
  seq_pack_kernel<<<dim3((unsigned)((max_n + 255) / 256), (unsigned)nitems), dim3(256), 0, s>>>(
      items);
  return hipGetLastError();
}

hipError_t launch_seq_progress(const SeqItem* items, int nitems, int64_t* out, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  seq_progress_kernel<<<(unsigned)((nitems + 63) / 64), 64, 0, s>>>(items, nitems, out);
  return hipGetLastError();
}

hipError_t launch_seq_state_init(const SeqItem* items, int nitems, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  seq_state_init_kernel<<<(unsigned)((nitems + 63) / 64), 64, 0, s>>>(items, nitems);
  return hipGetLastError();
}

hipError_t launch_memo_clear(const SeqItem* items, int nitems, int key_bits, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const int64_t words = ((int64_t)2 << key_bits) / 16;
  memo_clear_kernel<<<dim3(64, (unsigned)nitems), 256, 0, s>>>(items, words);
  return hipGetLastError();
}

hipError_t launch_seq_walk(const SeqItem* items, int nitems, int32_t* queue, const double* sw,
                           int use_cache, size_t lds_bytes, int waves, int assoc_any,
                           int64_t slice, uint64_t budget, uint32_t* stats, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const size_t lds = lds_bytes;
  const dim3 grid((unsigned)waves), block(kWave);
  auto go = [&](auto kernel) -> hipError_t {
    hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    kernel<<<grid, block, lds, s>>>(items, nitems, queue, sw, slice, budget, stats);
    return hipGetLastError();
  };
  if (assoc_any) {
    return use_cache ? go(seq_walk_kernel<true, true>) : go(seq_walk_kernel<true, false>);
  }
  return use_cache ? go(seq_walk_kernel<false, true>) : go(seq_walk_kernel<false, false>);
}

}  // namespace sa
