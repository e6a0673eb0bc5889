// sa_walk.hip -- the exact sequential Riemersma walk, batched.
//
// Restates RiemersmaWalkCpu (src/rd_riemersma_cpu.cpp:39-113, :359-417), which
// is the walk quantize.c:RiemersmaDither performs.  Bit-exact: every double
// operation is the reference's, in the reference's order, one rounding each
// (built with -ffp-contract=off).
//
// Mapping to CDNA4.  The walk is a serial chain, so one wavefront walks one
// image and many images walk at once (a persistent grid pulls images from a
// device queue, largest first).  Inside the wavefront the 64 lanes are four
// DPP rows of 16: row c holds channel c (r, g, b, a) and lane j of a row holds
// the running sum of one of the next 16 visits.  Visit T's diffusion is
//
//     p_T = x_T + w0*e_{T-16} + w1*e_{T-15} + ... + w15*e_{T-1}
//
// summed left to right.  Only the last term depends on the visit just
// finished, so when e_{T-1} is known every in-flight visit T..T+15 adds its
// own tap of it in ONE v_mul_f64 + ONE v_add_f64 for all four channels (lane j
// holds the weight for its visit's tap).  The lane holding visit T is then
// complete; after the visit it restarts with x_{T+16}.  Each visit therefore
// issues the reference's 32 multiply-adds per channel spread over 16 lanes,
// and the serial chain per visit is one mul, one add, the clamp, the memo
// probe and the residual.
//
// Memo (quantize.c's cache, src/rd_riemersma_cpu.cpp:75-94).  First-touch
// semantics are observable, so it is reproduced exactly: an LDS direct-mapped
// table of 2^S entries {valid, spilled, tag, index} in front of a per-image
// global table that holds every entry ever evicted.  A key whose slot was never
// evicted is known untouched without leaving LDS; a key whose slot has spilled
// is looked up in the global table (whose group of keys is marked untouched on
// the slot's first spill, so no per-image clear is needed).
//
// Palette search: the search image (sa_palette.h) is copied into LDS once per
// image; a memo miss descends the octree with uniform LDS reads and scores the
// candidate range across lanes (last index at the minimum wins, which is the
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

template <int L>
__device__ __forceinline__ double row_bcast(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x150 + L, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x150 + L, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, kWave));
  return v;
}

struct Search {
  const double* col;     // 4 per entry (LDS)
  const float4* ent;     // (LDS)
  const int2* sub;       // (LDS)
  const int16_t* child;  // (LDS)
  const uint16_t* cand;  // (LDS)
  int kids;
};

// Descent + candidate scan for one clamped target; uniform result.
// t0..t3: the target (double) already uniform; c0..c3 the scaled bytes.
template <bool ASSOC>
__device__ __attribute__((noinline)) int lookup(const Search& s, unsigned c0, unsigned c1, unsigned c2, unsigned c3,
                      double t0, double t1, double t2, double t3) {
  const int lane = threadIdx.x;
  int row = 0, prev = 0;
  for (int lev = kMaxDepth - 1; lev > 0; --lev) {
    const int id = node_id(ASSOC, c0, c1, c2, c3, lev);
    const int ch = rfl((int)s.child[row * s.kids + id]);
    if (ch == kNoChild) break;
    prev = row;
    if (ch == kLeaf) break;
    row = ch;
  }
  const int2 r = s.sub[prev];
  const int off = rfl(r.x), len = rfl(r.y);
  const double t[4] = {t0, t1, t2, t3};
  double best = kFarDistance;
  int bidx = 0;
  for (int b = 0; b < len; b += kWave) {
    const int k = b + lane;
    double d = __builtin_inf();
    int ci = 0;
    if (k < len) {
      ci = (int)s.cand[off + k];
      const float4 e4 = s.ent[ci];
      const float e[4] = {e4.x, e4.y, e4.z, e4.w};
      d = candidate_distance(ASSOC, e, t);
    }
    const double m = wave_min(d);
    if (m <= best) {
      const unsigned long long mask = __ballot(k < len && d == m);
      const int last = 63 - __builtin_clzll(mask);
      best = m;
      bidx = rl(ci, last);
    }
  }
  return bidx;
}

struct Memo {
  uint32_t* tab;  // LDS, 1<<S entries
  int32_t* spill; // global, 1<<K entries
  int S;
  uint32_t M;
  int K;
};

__device__ __forceinline__ int memo_slot_key(const Memo& m, uint32_t slot, uint32_t tag) {
  return (int)((tag << m.S) | (slot ^ (tag & m.M)));
}

// Memo miss for one visit (uniform): the key is not in its LDS slot.  `ent`
// is the slot's current content.
template <bool ASSOC>
__device__ __attribute__((noinline)) int memo_miss(const Memo& m, const Search& s, uint32_t key,
                                                   uint32_t ent, unsigned c0, unsigned c1,
                                                   unsigned c2, unsigned c3, double t0, double t1,
                                                   double t2, double t3) {
  const uint32_t tag = key >> m.S;
  const uint32_t slot = (key ^ tag) & m.M;
  const bool valid = (ent >> 31) != 0;
  const bool spilled = ((ent >> 30) & 1u) != 0;
  const int lane = threadIdx.x;
  int found = -1;
  if (spilled) found = rfl(m.spill[key]);
  const int idx = found >= 0 ? found : lookup<ASSOC>(s, c0, c1, c2, c3, t0, t1, t2, t3);
  bool now_spilled = spilled;
  if (valid) {
    // Evict the occupant to the global table.  On this slot's first spill,
    // first mark every key that maps here as untouched.
    if (!spilled) {
      const int group = 1 << (m.K - m.S);
      for (int t = lane; t < group; t += kWave) m.spill[memo_slot_key(m, slot, (uint32_t)t)] = -1;
      __threadfence_block();
    }
    if (lane == 0) m.spill[memo_slot_key(m, slot, (ent >> 16) & 0x3FFFu)] = (int)(ent & 0xFFFFu);
    __threadfence_block();
    now_spilled = true;
  }
  if (lane == 0)
    m.tab[slot] = 0x80000000u | (now_spilled ? 0x40000000u : 0u) | (tag << 16) | (uint32_t)idx;
  return idx;
}

struct WalkState {
  double acc;   // running sum of this lane's in-flight visit
  double e;     // residual of the visit this lane completed last
  double xn1, xn2, xn3;  // restart values for the next three chunks
  uint32_t obuf;         // palette indices of the current 64-visit window
};

__device__ __forceinline__ double load_x(const double* xs, int64_t v, int64_t n, int ch) {
  return v < n ? xs[4 * v + ch] : 0.0;
}

// One visit at phase I (visit T = base + I).  Returns false when T >= n.
template <int I, bool ASSOC, bool CACHE>
__device__ __forceinline__ void visit(WalkState& w, const double (&W)[16], int64_t base,
                                      const Memo& m, const Search& s) {
  const int lane = threadIdx.x;
  const int j = lane & 15;
  const int ch = lane >> 4;
  // e_{T-1} lives in lane (I-1)&15 of every row.
  const double eb = row_bcast<(I + 15) & 15>(w.e);
  const double prod = W[I] * eb;
  w.acc = w.acc + prod;
  const float pc = clamp_pixel(w.acc);
  const double pd = (double)pc;
  const unsigned c8 = scale_char(pc);
  const unsigned c0 = (unsigned)rl((int)c8, I);
  const unsigned c1 = (unsigned)rl((int)c8, 16 + I);
  const unsigned c2 = (unsigned)rl((int)c8, 32 + I);
  const unsigned c3 = ASSOC ? (unsigned)rl((int)c8, 48 + I) : 0u;
  int idx;
  if (CACHE) {
    const uint32_t key = (uint32_t)memo_key(ASSOC, c0, c1, c2, c3);
    const uint32_t tag = key >> m.S;
    const uint32_t slot = (key ^ tag) & m.M;
    const uint32_t ent = (uint32_t)rfl((int)m.tab[slot]);
    if ((ent >> 31) != 0 && ((ent >> 16) & 0x3FFFu) == tag) {
      idx = (int)(ent & 0xFFFFu);
    } else {
      idx = memo_miss<ASSOC>(m, s, key, ent, c0, c1, c2, c3, rl(pd, I), rl(pd, 16 + I),
                             rl(pd, 32 + I), ASSOC ? rl(pd, 48 + I) : 65535.0);
    }
  } else {
    idx = lookup<ASSOC>(s, c0, c1, c2, c3, rl(pd, I), rl(pd, 16 + I), rl(pd, 32 + I),
                        ASSOC ? rl(pd, 48 + I) : 65535.0);
  }
  // Residual against the associated palette colour, per row's channel.
  w.e = pd - s.col[4 * idx + ch];
  w.obuf = lane == (int)((base + I) & 63) ? (uint32_t)idx : w.obuf;
  w.acc = (j == I) ? w.xn1 : w.acc;
}

template <bool ASSOC, bool CACHE, int... Is>
__device__ __forceinline__ bool chunk(WalkState& w, const double (&W)[16], int64_t base,
                                      int64_t n, const Memo& m, const Search& s,
                                      std::integer_sequence<int, Is...>) {
  if (base + 16 <= n) {
    (visit<Is, ASSOC, CACHE>(w, W, base, m, s), ...);
    return true;
  }
  // Tail chunk: stop at n.
  bool go = true;
  ((go = go && (base + Is < n), go ? (visit<Is, ASSOC, CACHE>(w, W, base, m, s), 0) : 0), ...);
  return false;
}

template <bool ASSOC, bool CACHE>
__device__ void walk_image(const SeqItem& it, const double (&W)[16], const Memo& m,
                           const Search& s) {
  const int lane = threadIdx.x;
  const int j = lane & 15;
  const int ch = lane >> 4;
  const int64_t n = it.n;
  const double* xs = it.xs;
  WalkState w;
  w.e = 0.0;
  w.acc = load_x(xs, j, n, ch) + 0.0;
  w.xn1 = load_x(xs, 16 + j, n, ch) + 0.0;
  w.xn2 = load_x(xs, 32 + j, n, ch) + 0.0;
  w.xn3 = load_x(xs, 48 + j, n, ch) + 0.0;
  w.obuf = 0;
  for (int64_t base = 0; base < n; base += 16) {
    const bool full = chunk<ASSOC, CACHE>(w, W, base, n, m, s,
                                          std::make_integer_sequence<int, 16>{});
    const int64_t next = base + 16;
    if ((next & 63) == 0 || !full || next >= n) {
      const int64_t wbase = (next - 1) & ~(int64_t)63;
      const int64_t v = wbase + lane;
      if (v < n && v < next) it.idx[v] = (uint16_t)w.obuf;
    }
    w.xn1 = w.xn2;
    w.xn2 = w.xn3;
    w.xn3 = load_x(xs, next + 48 + j, n, ch) + 0.0;
  }
}

template <bool ASSOC, bool CACHE>
__global__ __launch_bounds__(kWave) void seq_walk_kernel(const SeqItem* __restrict__ items,
                                                         int nitems, int32_t* queue,
                                                         const double* __restrict__ sw,
                                                         int memo_bits) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int lane = threadIdx.x;
  const int j = lane & 15;
  // Per-lane weight of each phase: lane j at phase I adds tap 15-((j-I)&15).
  double W[16];
#pragma unroll
  for (int I = 0; I < 16; ++I) W[I] = sw[15 - ((j - I) & 15)];
  Memo m;
  m.S = memo_bits;
  m.M = (1u << memo_bits) - 1u;
  m.K = ASSOC ? 24 : 18;
  m.tab = (uint32_t*)lds;
  unsigned char* img = lds + (CACHE ? (4u << memo_bits) : 0u);
  for (;;) {
    int q = 0;
    if (lane == 0) q = atomicAdd(queue, 1);
    q = __shfl(q, 0, kWave);
    if (q >= nitems) break;
    const SeqItem it = items[q];
    const PaletteDev pd = *it.pal;
    // Stage the search image and clear the memo.
    {
      const uint4* src = (const uint4*)pd.image;
      uint4* dst = (uint4*)img;
      for (int k = lane; k < pd.bytes / 16; k += kWave) dst[k] = src[k];
      if (CACHE)
        for (uint32_t k = lane; k < (1u << memo_bits); k += kWave) m.tab[k] = 0u;
      __syncthreads();
    }
    Search s;
    s.col = (const double*)img;
    s.ent = (const float4*)(img + pd.off_ent);
    s.sub = (const int2*)(img + pd.off_sub);
    s.child = (const int16_t*)(img + pd.off_child);
    s.cand = (const uint16_t*)(img + pd.off_cand);
    s.kids = pd.kids;
    m.spill = it.spill;
    walk_image<ASSOC, CACHE>(it, W, m, s);
    __syncthreads();
  }
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

hipError_t launch_seq_walk(const SeqItem* items, int nitems, int32_t* queue, const double* sw,
                           int memo_bits, int use_cache, int lds_image_bytes, int waves,
                           int assoc_any, hipStream_t s) {
  if (nitems <= 0) return hipSuccess;
  const size_t lds = (use_cache ? (4u << memo_bits) : 0u) + (size_t)lds_image_bytes;
  const dim3 grid((unsigned)waves), block(kWave);
  auto go = [&](auto kernel) -> hipError_t {
    hipError_t e = hipFuncSetAttribute((const void*)kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    kernel<<<grid, block, lds, s>>>(items, nitems, queue, sw, memo_bits);
    return hipGetLastError();
  };
  if (assoc_any) {
    return use_cache ? go(seq_walk_kernel<true, true>) : go(seq_walk_kernel<true, false>);
  }
  return use_cache ? go(seq_walk_kernel<false, true>) : go(seq_walk_kernel<false, false>);
}

}  // namespace sa
