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
// channels.  Lane 0 is then complete; after the visit every row shifts left by
// one lane (DPP row_shl:1) and lane 15, which has no source, keeps the DPP's
// fill value: x_{t+16}, read from an LDS ring of the next inputs.
//
// Inputs.  A gather kernel stages each visit's source pixel in visit order at
// the source's own width (4 bytes for 8-bit RGBA); at the start of each
// 16-visit chunk the walk associates the next chunk's pixels in registers
// (loaded one chunk earlier) and writes them to the ring.
//
// The serial chain per visit is: tap (mul, add), the channel byte, the memo
// key (four lane reads, scalar ORs and hash), the memo read, and the residual
// against the palette colour -- which is read from LDS only when the index
// differs from the previous visit's (55-70% of visits repeat it).  The clamped
// target is formed in the shadow of the memo read.
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
// once per image (a search image above 96 KiB -- palettes of more than about
// 1400 colours -- is read where it lies in HBM instead); the first three
// descent levels are one table read, the rest are uniform reads, and the
// candidate range is scored across lanes
// with a DPP min-reduction (the last index at the minimum wins, which is the
// reference's `<=` rule).
#include "sa_internal.h"
#include "sa_math.h"

#include <type_traits>

// v_writelane_b32: `v` with lane `lane` (uniform) replaced by the uniform `s`.
__device__ int llvm_writelane(int s, int lane, int v) __asm("llvm.amdgcn.writelane.i32");

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
typedef const uint32_t __attribute__((address_space(1))) gcu32;
typedef uint16_t __attribute__((address_space(1))) gu16;

__device__ __forceinline__ int64_t rfl64(int64_t v) {
  const uint32_t lo = (uint32_t)rfl((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)rfl((int)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// DPP move in which every lane has a source (no fill value needed).
template <int CTRL>
__device__ __forceinline__ double dpp_all(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

constexpr int kRowShl1 = 0x101;  // lane j <- lane j+1; lane 15 has no source

// LDS cache of the HBM memo table's 64-entry lines.  Entries are index+1
// (0 = untouched), bytes when the palette has at most 255 colours (twice the
// lines in the same LDS), 16-bit up to 65535 colours, 32-bit for 65536.
struct Memo {
  uint32_t* tag;    // per slot: bit 31 empty, bit 30 dirty, bits 0..29 line id
  void* line;       // 64 entries per slot
  void* table;      // HBM, 1 << K entries, zeroed before the walk
  uint32_t slots;   // any count: the hash scales into it
  // This lane's channel part of the key (memo_key_part for its row):
  // bfe(c8, k_hi_off, k_hi_w) << k_hi_shl | bfe(c8, 2, k_lo_w) << k_lo_shl.
  uint32_t k_hi_off, k_hi_w, k_hi_shl, k_lo_w, k_lo_shl;
};

// Rows 0..2 (r, g, b) give their top 4 bits to the line and bits 2..3 to the
// offset; row 3 (alpha) gives its top 6 bits to the line when associated.
template <bool ASSOC>
__device__ __forceinline__ void memo_key_consts(Memo& m, int ch) {
  m.k_hi_off = ch == 3 ? 2u : 4u;
  m.k_hi_w = ch == 3 ? (ASSOC ? 6u : 0u) : 4u;
  m.k_hi_shl = ch == 3 ? 18u : (uint32_t)(6 + 4 * ch);
  m.k_lo_w = ch == 3 ? 0u : 2u;
  m.k_lo_shl = ch == 3 ? 0u : (uint32_t)(2 * ch);
}

constexpr uint32_t kEmptyTag = 0x80000000u;
constexpr uint32_t kDirty = 0x40000000u;

// Multiplicative hash of the line id, scaled into [0, slots).
__device__ __forceinline__ uint32_t memo_slot(const Memo& m, uint32_t line) {
  return __umulhi(line * 0x9E3779B1u, m.slots);
}

template <int EB>
using MemoEntry = typename std::conditional<
    EB == 1, uint8_t, typename std::conditional<EB == 2, uint16_t, uint32_t>::type>::type;

// Makes `line` resident in `slot` (uniform).  Returns 1 when it read HBM.
template <int EB>
__device__ __forceinline__ uint32_t memo_fill(const Memo& m, uint32_t slot, uint32_t line,
                                              uint32_t old) {
  typedef MemoEntry<EB> E;
  typedef E __attribute__((address_space(1))) gE;
  const int lane = threadIdx.x;
  E* dst = (E*)m.line + 64 * slot;
  uint32_t loaded = 0;
  if (old & kEmptyTag) {
    // Never held a line: this line was never written, so it is all zero.
    dst[lane] = 0;
  } else {
    if (old & kDirty) ((gE*)m.table)[64 * (old & 0x3FFFFFFFu) + lane] = dst[lane];
    dst[lane] = ((const gE*)m.table)[64 * line + lane];
    loaded = 1;
  }
  if (lane == 0) m.tag[slot] = line;
  __builtin_amdgcn_wave_barrier();
  return loaded;
}

struct Counters {
  uint32_t misses = 0, fills = 0, loads = 0;
};

// The staged source pixel of visit v (clamped into the image, so the loads
// never branch): 16 bytes from its start, which the stage's padding covers.
__device__ __forceinline__ uint4 load_stage(const unsigned char* stage, int shift, int64_t v,
                                            int64_t n) {
  const gcu32* p = (const gcu32*)(stage + ((v < n ? v : n - 1) << shift));
  return make_uint4(p[0], p[1], p[2], p[3]);
}

// This lane's channel of a staged pixel as the walk's input: Q16 floats as
// load_pixel reads them, associated as AssociateAlphaPixel does
// (include/rd_types.h:179-192), +0.0 (the zero taps a visit near the start
// adds turn -0.0 into +0.0, exactly as the reference's first adds do).
template <bool ASSOC>
__device__ __forceinline__ double stage_input(uint4 w, int shift, int ch) {
  float c, a;
  if (shift == 2) {
    c = (float)((w.x >> (8 * ch)) & 255u) * 257.0f;
    a = (float)(w.x >> 24) * 257.0f;
  } else if (shift == 3) {
    const uint32_t h = ch < 2 ? w.x : w.y;
    c = (float)((h >> (16 * (ch & 1))) & 0xFFFFu);
    a = (float)(w.y >> 16);
  } else {
    c = __uint_as_float(ch == 0 ? w.x : ch == 1 ? w.y : ch == 2 ? w.z : w.w);
    a = __uint_as_float(w.w);
  }
  double x = (double)c;
  if (ASSOC && ch < 3 && a != 65535.0f) x = (kScale * (double)a) * x;
  return x + 0.0;
}

// The next inputs wait in an LDS ring of kRing visits, 4 doubles each; a lane
// reads and writes its own channel (`ring` points at it).
constexpr int kRing = kSeqRingBytes / 32;

// The channel byte scale_char_clamped(clamp_pixel(v)) from the sum itself:
// for v in [0, 65535) the float is clamp_pixel's; below 0 the sum s + 0.5 is
// at most 0.5 (and a NaN gives 0 through the max), and at or above 65535 (or
// inf) the final min gives 255, as the clamp would.  Two dependent operations
// fewer than clamping first.
__device__ __forceinline__ unsigned key_char(double v) {
  const float s = (float)((double)(float)v * kInv257);
  return (unsigned)fminf(fmaxf(s + 0.5f, 0.0f), 255.0f);
}

struct Visit {
  double acc;      // running sum of this lane's in-flight visit
  double eb;       // residual of the last visit, in every lane of its row
  double col;      // this row's channel of palette colour `last`
  uint32_t obuf;   // palette indices of the current 64-visit window
  int last;        // index of the last visit (-1: none yet)
};

// One visit (visit t); on return lane j holds visit t+1+j.
template <bool ASSOC, bool CACHE, int EB>
__device__ __forceinline__ void visit(Visit& w, double W, int64_t t, double* ring, const Memo& m,
                                      const Search& s, Counters& cnt) {
  const int lane = threadIdx.x;
  const int ch = lane >> 4;
  // x_{t+16}, which lane 15 restarts with at the end of the visit; read
  // first, so its wait is long over when the memo reads complete (LDS reads
  // return in order).
  const double xin = ring[4 * ((t + 16) & (kRing - 1))];
  __builtin_amdgcn_sched_barrier(0);
  // Every in-flight visit adds its tap of e_{t-1}; lane 0 is then complete.
  w.acc = w.acc + W * w.eb;
  const unsigned c8 = key_char(w.acc);
  // The clamped target itself (the residual's minuend and the search's
  // target) is formed while the memo read is in flight.
  double pd, pdb;
  auto target = [&]() {
    pd = (double)clamp_pixel(w.acc);
    pdb = dpp_all<kRowBcast0>(pd);
  };
  // The channel bytes of visit t sit in lanes 0, 16, 32, 48.
  auto search = [&]() {
    ++cnt.misses;
    return lookup<ASSOC>(s, (unsigned)rl((int)c8, 0), (unsigned)rl((int)c8, 16),
                         (unsigned)rl((int)c8, 32), ASSOC ? (unsigned)rl((int)c8, 48) : 0u,
                         rl(pd, 0), rl(pd, 16), rl(pd, 32), ASSOC ? rl(pd, 48) : 65535.0);
  };
  int idx;
  if (CACHE) {
    // Each row forms its channel's part of the memo key (memo_key_part with
    // this lane's constants: no branch); four lane reads and three ORs join
    // them (the channel bytes themselves are read only on a miss).
    const uint32_t kc =
        (__builtin_amdgcn_ubfe(c8, m.k_hi_off, m.k_hi_w) << m.k_hi_shl) |
        (__builtin_amdgcn_ubfe(c8, 2u, m.k_lo_w) << m.k_lo_shl);
    uint32_t key = (uint32_t)rl((int)kc, 0) | (uint32_t)rl((int)kc, 16) | (uint32_t)rl((int)kc, 32);
    if (ASSOC) key |= (uint32_t)rl((int)kc, 48);
    const uint32_t line = key >> 6;
    const uint32_t slot = memo_slot(m, line);
    typedef MemoEntry<EB> E;
    E* const entp = (E*)m.line + 64 * slot + (key & 63);
    const uint32_t tgv = m.tag[slot];
    const uint32_t entv = *entp;
    // Nothing else is scheduled ahead of these reads (in-order issue: work
    // placed before them would delay the chain), and the target is formed
    // behind them.
    __builtin_amdgcn_sched_barrier(0);
    target();
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t tg = (uint32_t)rfl((int)tgv);
    uint32_t ent = (uint32_t)rfl((int)entv);
    idx = (int)ent - 1;
    // The hit (line resident, entry touched) falls through: (ent - 1) >> 16
    // is nonzero only for ent == 0 (entries are at most 65536).
    if (__builtin_expect((((tg & ~kDirty) ^ line) | ((ent - 1u) >> 16)) != 0, 0)) {
      if ((tg & ~kDirty) != line) {
        ++cnt.fills;
        cnt.loads += memo_fill<EB>(m, slot, line, tg);
        ent = (uint32_t)rfl((int)*entp);
        idx = (int)ent - 1;
      }
      if (ent == 0) {
        idx = search();
        if (lane == 0) {
          *entp = (E)(idx + 1);
          m.tag[slot] = line | kDirty;
        }
      }
    }
  } else {
    target();
    idx = search();
  }
  // Residual against the associated palette colour, per row's channel, formed
  // in every lane from lane 0's target.  Consecutive visits often take the
  // same entry: its colour is then already in a register.
  if (__builtin_expect(idx != w.last, 0)) {
    w.col = s.col[4 * idx + ch];
    w.last = idx;
  }
  w.eb = pdb - w.col;
  w.obuf = (uint32_t)llvm_writelane(idx, (int)(t & 63), (int)w.obuf);
  // Shift: lane j takes visit t+1+j; lane 15 restarts with x_{t+16}.
  w.acc = dpp<kRowShl1>(w.acc, xin);
}

// Writes the indices of visits (t & ~63) .. t (the window's filled part).
__device__ __forceinline__ void flush_window(uint16_t* idx_out, uint32_t obuf, int64_t t) {
  const int64_t v = (t & ~(int64_t)63) + threadIdx.x;
  if (v <= t) ((gu16*)idx_out)[v] = (uint16_t)obuf;
}

// Visits [t0, stop).  t0 and stop - t0 are multiples of 48 unless stop is the
// end.  A walk that stops early leaves its queue (the in-flight sums and
// residuals) in it.state and flushes the indices of its last, partial 64-visit
// window; resuming loads them back, so a sliced walk computes exactly what an
// uninterrupted one does.  `deadline` (a s_memrealtime tick, 0: none) ends the
// walk at the first 48-visit boundary past it.  Returns where the walk stopped.
//
// Inputs: at the start of each 16-visit chunk the next chunk's staged pixels
// (loaded one chunk earlier) are associated and written to the ring, and the
// chunk after that is loaded; each visit reads its x_{t+16} from the ring.
template <bool ASSOC, bool CACHE, int EB, bool IMG_LDS>
__device__ int64_t walk_image(const SeqItem& it, double W, const Memo& m, const Search& s,
                              double* ring_base, Counters& cnt, int64_t t0, int64_t stop,
                              uint64_t deadline) {
  const int lane = threadIdx.x;
  const int j = lane & 15;
  const int ch = lane >> 4;
  const int64_t n = rfl64(it.n);
  t0 = rfl64(t0);
  stop = rfl64(stop);
  const int shift = rfl(seq_stage_shift(it.src_fmt));
  const unsigned char* stage = (const unsigned char*)it.stage;
  double* ring = ring_base + ch;
  Visit w;
  w.last = -1;
  w.col = 0.0;
  if (t0 == 0) {
    w.eb = 0.0;
    w.obuf = 0;
    w.acc = stage_input<ASSOC>(load_stage(stage, shift, j, n), shift, ch);
  } else {
    w.acc = ((gdouble*)it.state->acc)[lane];
    w.eb = ((gdouble*)it.state->e)[lane];
    const int64_t v = (t0 & ~(int64_t)63) + lane;
    w.obuf = v < t0 ? (uint32_t)((const gu16*)it.idx)[v] : 0u;
  }
  ring[4 * ((t0 + 16 + j) & (kRing - 1))] =
      stage_input<ASSOC>(load_stage(stage, shift, t0 + 16 + j, n), shift, ch);
  // Settle these loads here, before the prefetch below: left pending, the
  // compiler waits for them inside the first chunk instead.
  __builtin_amdgcn_s_waitcnt(0);
  uint4 raw = load_stage(stage, shift, t0 + 32 + j, n);
  int third = 0;
  for (int64_t base = t0;; base += 16) {
    if (base != t0) {
      if (++third == 3) {
        third = 0;
        if (deadline && __builtin_amdgcn_s_memrealtime() >= deadline) {
          stop = base;
          break;
        }
      }
      ring[4 * ((base + 16 + j) & (kRing - 1))] = stage_input<ASSOC>(raw, shift, ch);
      raw = load_stage(stage, shift, base + 32 + j, n);
    }
    // A 32-bit count keeps the loop test on the scalar unit (it has no
    // 64-bit ordered compare).
    const int len = rfl((int)(stop - base < 16 ? stop - base : 16));
    if (len == 16) {
      // Whole chunks run four visits per loop trip: the loop's own scalar
      // count, compare and taken branch were a measurable part of a hit visit
      // (240 against 248 ns per visit on the largest still; a full unroll
      // outgrows the instruction cache and is slower again).
#pragma unroll 4
      for (int k = 0; k < 16; ++k) visit<ASSOC, CACHE, EB>(w, W, base + k, ring, m, s, cnt);
    } else {
      for (int k = 0; k < len; ++k) visit<ASSOC, CACHE, EB>(w, W, base + k, ring, m, s, cnt);
    }
    const int64_t last = base + len - 1;
    if ((last & 63) == 63 || last + 1 == n) flush_window(it.idx, w.obuf, last);
    if (len < 16 || base + 16 == stop) break;
  }
  if (stop < n) {
    // The last window is partial unless stop - 1 ended it (then it was
    // flushed in the loop); the queue stays for the next launch.
    const int64_t last = stop - 1;
    if ((last & 63) != 63) flush_window(it.idx, w.obuf, last);
    it.state->acc[lane] = w.acc;
    it.state->e[lane] = w.eb;
  }
  return stop;
}

// The walk of an image whose search image stays in HBM, out of line so the
// common case's code stays as compact as before.
template <bool ASSOC, bool CACHE>
__device__ __attribute__((noinline)) int64_t walk_image_hbm(const SeqItem& it, double W,
                                                            const Memo& m, const Search& s,
                                                            double* ring, Counters& cnt,
                                                            int64_t t0, int64_t stop,
                                                            uint64_t deadline, int eb) {
  if (CACHE && eb == 4)
    return walk_image<ASSOC, CACHE, 4, false>(it, W, m, s, ring, cnt, t0, stop, deadline);
  return walk_image<ASSOC, CACHE, 2, false>(it, W, m, s, ring, cnt, t0, stop, deadline);
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
    const uint32_t slots = (uint32_t)rfl(it.memo_slots);
    const int eb = rfl(it.memo_eb);
    m.slots = slots;
    memo_key_consts<ASSOC>(m, lane >> 4);
    const uint32_t memo_bytes = CACHE ? seq_memo_bytes((int)slots, eb) : 0u;
    m.tag = (uint32_t*)lds;
    m.line = lds + 4 * slots;
    const PaletteDev pd = *it.pal;
    // The search image in LDS, or read where it lies when it is too large
    // (the memo cache then has the LDS).  Each case is its own instantiation,
    // so the LDS one keeps LDS reads (a pointer that may be either would make
    // every search read a flat access).
    const bool img_lds = pd.bytes <= kSeqImageLdsMax;
    {
      if (img_lds) copy16((uint4*)(lds + memo_bytes), (const uint4*)pd.image, (uint32_t)pd.bytes / 16);
      if (CACHE) {
        if (t0 == 0) {
          for (uint32_t k = lane; k < slots; k += kWave) m.tag[k] = kEmptyTag;
        } else {  // resume: the memo cache exactly as the last launch left it
          copy16((uint4*)lds, (const uint4*)it.snap, memo_bytes / 16);
        }
      }
      __syncthreads();
    }
    auto search_at = [&](const unsigned char* img) {
      Search s;
      s.col = (const double*)img;
      s.cent = (const float4*)(img + pd.off_cent);
      s.sub = (const int2*)(img + pd.off_sub);
      s.desc = (const uint32_t*)(img + pd.off_desc);
      s.child = (const int16_t*)(img + pd.off_child);
      s.cand = (const uint16_t*)(img + pd.off_cand);
      s.kids = pd.kids;
      return s;
    };
    m.table = it.memo;
    const uint64_t deadline =
        (budget && it.state) ? __builtin_amdgcn_s_memrealtime() + budget : 0;
    if (img_lds) {
      const Search s = search_at(lds + memo_bytes);
      double* ring = (double*)(lds + memo_bytes + pd.bytes);
      if (CACHE && eb == 1)
        stop = walk_image<ASSOC, CACHE, 1, true>(it, W, m, s, ring, cnt, t0, stop, deadline);
      else
        stop = walk_image<ASSOC, CACHE, 2, true>(it, W, m, s, ring, cnt, t0, stop, deadline);
    } else {  // more than about 1400 colours: 2- or 4-byte entries
      const Search s = search_at((const unsigned char*)pd.image);
      stop = walk_image_hbm<ASSOC, CACHE>(it, W, m, s, (double*)(lds + memo_bytes), cnt, t0,
                                          stop, deadline, eb);
    }
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
                                                         int key_bits) {
  const SeqItem it = items[blockIdx.y];
  uint4* t = (uint4*)it.memo;
  const int64_t w = ((int64_t)it.memo_eb << key_bits) / 16;
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < w; k += (int64_t)gridDim.x * 256)
    t[k] = make_uint4(0, 0, 0, 0);
}

// Gather: each visit's source pixel in visit order, in the source's width
// (seq_stage_shift); the walk associates it.
__global__ __launch_bounds__(256) void seq_gather_kernel(const SeqItem* __restrict__ items) {
  const SeqItem it = items[blockIdx.y];
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= it.n) return;
  const int64_t p = it.curve[v];
  switch (it.src_fmt) {
    case kRgb8: {
      const uint8_t* s = (const uint8_t*)it.src + 3 * p;
      ((uint32_t*)it.stage)[v] = (uint32_t)s[0] | ((uint32_t)s[1] << 8) | ((uint32_t)s[2] << 16) |
                                 0xFF000000u;
      break;
    }
    case kRgba8: ((uint32_t*)it.stage)[v] = ((const uint32_t*)it.src)[p]; break;
    case kRgb16: {
      const uint16_t* s = (const uint16_t*)it.src + 3 * p;
      ((uint2*)it.stage)[v] = make_uint2((uint32_t)s[0] | ((uint32_t)s[1] << 16),
                                         (uint32_t)s[2] | 0xFFFF0000u);
      break;
    }
    case kRgba16: ((uint2*)it.stage)[v] = ((const uint2*)it.src)[p]; break;
    default: ((float4*)it.stage)[v] = ((const float4*)it.src)[p]; break;
  }
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
  memo_clear_kernel<<<dim3(64, (unsigned)nitems), 256, 0, s>>>(items, key_bits);
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
