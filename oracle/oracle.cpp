// oracle.cpp -- CPU restatement of the reference Riemersma hot path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  The product never links this.
//
// Built with g++ -O2 -ffp-contract=off (oracle/Makefile): every double and
// float operation below rounds exactly once, in the order written, which is
// the order the reference writes them in.
//
// How each part is pinned (tests/test_oracle.py, tests/test_numeric.py):
//   * weights        -- exact IEEE hex of all 16 values (closed form 16^(-i/15))
//                       and the reference's table; against oracle/_ref.
//   * curve          -- closed form == explicit recursion for levels 0..9
//                       (brute force); endpoint (2^L-1,0); permutation of the
//                       grid; block-curve length pixels+1; unvisited-pixel rule;
//                       against oracle/_ref BuildCurveIndex.
//   * scale_char     -- exhaustive equality with a long-double reformulation
//                       on every float in [0, 65536) (tests/test_oracle.py).
//   * octree         -- reference KAT (tests/in_gray_alpha.png at 2 colours:
//                       52 nodes, 1 colour 32612/32768); golden fixtures from
//                       oracle/_ref ColorTree::Build on seeded images.
//   * sequential     -- golden fixtures from oracle/_ref RiemersmaWalkCpu; a
//                       hand-computed 1x1 walk; memo first-touch property.
//   * blocks         -- golden fixtures from oracle/_ref RiemersmaBlocksCpu;
//                       block >= n equals the cache-free sequential walk on
//                       the block curve.
//   * planar YUV     -- 4:4:4 in (swscale's matrix) and out (BT.601), and
//                       rgba64 out of YUV in: against oracle/_ref's fused
//                       gather/scatter of RiemersmaBlocksCpu.  4:2:0 in (a
//                       CUDA-only reader in the reference): even coordinates
//                       copy the chroma sample, constant chroma stays
//                       constant, and odd ones are the rounded bilinear mean,
//                       checked in numpy; the result then feeds the pinned
//                       4:4:4 path.
//   * screen dithers -- bayer-ordered, void-and-cluster (64, 32), clustered-
//                       dot: against oracle/_ref's own plugins run through
//                       FindDither (tests/test_screens.py).
//   * inverse filter -- triangular Toeplitz solve of (I-H) g = delta (numpy).
//   * Jacobi         -- the reference has no CPU Jacobi and its CUDA engine
//                       cannot be built here, so no golden vector exists; it
//                       is pinned by what the mathematics fixes.  One sweep
//                       from a given q (orc_jacobi_sweep) equals the exact
//                       closed form over piecewise-constant residuals --
//                       prefix sums of g over each segment's overlap with the
//                       tap range, in rational arithmetic -- within the
//                       standard rounding bound of the sweep's float type,
//                       and picks the nearest entry of the exact target
//                       wherever the margin clears that bound; an in-place
//                       update, a dropped tap, a flipped sign and a late
//                       residual planted here each fail it.  iterations = k
//                       equals k composed sweeps from InitIndex (the
//                       ping-pong), and the scatter writes pal[q] at each
//                       pixel's owning position.  Zero diffusion equals the
//                       block walk at diffusion 0.  What stays fixed only by
//                       this restatement is the rounding order inside the
//                       bound (d ascending, one rounding per multiply and per
//                       add, src/rd_approx_cuda.cu:209-216).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace {

constexpr double kRange = 65535.0;       // QuantumRange
constexpr double kScale = 1.0 / 65535.0; // QuantumScale
constexpr double kEps = 1.0e-12;         // MagickEpsilon
constexpr int kQueue = 16;               // ErrorQueueLength
constexpr int kMaxDepth = 8;             // MaxTreeDepth
constexpr int kMaxColormap = 65536;      // MaxColormapSize
constexpr size_t kMaxQNodes = 266817;    // MaxQNodes

struct D4 {
  double c[4] = {0.0, 0.0, 0.0, 0.0};
};

// rd_types.h:30-33 MagickSafeReciprocal.
double safe_recip(double x) {
  const double sign = x < 0.0 ? -1.0 : 1.0;
  return (sign * x) >= kEps ? 1.0 / x : sign * (1.0 / kEps);
}

// rd_types.h:39-43 ClampPixel -> float.
float clamp_px(double v) {
  if (v < 0.0) return 0.0f;
  if (v >= kRange) return (float)kRange;
  return (float)v;
}

// rd_types.h:133-138 ScaleQuantumToChar: float divide, float add, truncate.
unsigned scale_char(float q) {
  if (std::isnan(q) || q <= 0.0f) return 0;
  const float s = q / 257.0f;
  if (s >= 255.0f) return 255;
  return (unsigned)(unsigned char)(s + 0.5f);
}

// rd_types.h:141-153 CacheOffset: 6 bits per channel.
int memo_key(bool assoc, const D4& p) {
  int k = (int)(scale_char(clamp_px(p.c[0])) >> 2) |
          ((int)(scale_char(clamp_px(p.c[1])) >> 2) << 6) |
          ((int)(scale_char(clamp_px(p.c[2])) >> 2) << 12);
  if (assoc) k |= (int)(scale_char(clamp_px(p.c[3])) >> 2) << 18;
  return k;
}

// rd_types.h:169-176: weight *= exp(log(16)/15), reciprocal stored.
void raw_weights(double w[kQueue]) {
  double weight = 1.0;
  for (int i = 0; i < kQueue; ++i) {
    w[i] = safe_recip(weight);
    weight *= std::exp(std::log(1.0 / (1.0 / 16.0)) / (kQueue - 1.0));
  }
}

// rd_riemersma_cpu.cpp:370-373: (1/16 * diffusion) * weights[k].
void scaled_weights(double diffusion, double sw[kQueue]) {
  double w[kQueue];
  raw_weights(w);
  for (int k = 0; k < kQueue; ++k) sw[k] = (1.0 / 16.0) * diffusion * w[k];
}

// rd_types.h:179-192 AssociateAlphaPixel (float pixel -> double).
D4 assoc_px(bool assoc, const float* px) {
  D4 o;
  o.c[3] = (double)px[3];
  if (!assoc || px[3] == (float)kRange) {
    o.c[0] = (double)px[0];
    o.c[1] = (double)px[1];
    o.c[2] = (double)px[2];
    return o;
  }
  const double a = kScale * (double)px[3];
  o.c[0] = a * (double)px[0];
  o.c[1] = a * (double)px[1];
  o.c[2] = a * (double)px[2];
  return o;
}

// rd_types.h:195-208 AssociateAlphaPixelInfo (palette entry).
D4 assoc_info(bool assoc, const double* e) {
  D4 o;
  o.c[3] = e[3];
  if (!assoc || e[3] == kRange) {
    o.c[0] = e[0];
    o.c[1] = e[1];
    o.c[2] = e[2];
    return o;
  }
  const double a = kScale * e[3];
  o.c[0] = a * e[0];
  o.c[1] = a * e[1];
  o.c[2] = a * e[2];
  return o;
}

}  // namespace

// ===========================================================================
// Octree.  Restates src/rd_octree.cpp:39-409 (itself quantize.c's
// Classify/Prune/Reduce/DefineImageColormap/ClosestColor).
// ===========================================================================
struct ONode {
  int32_t child[16];
  int32_t parent;
  double total[4];
  double qe;
  uint64_t nu;
  uint32_t color_number;
  uint32_t id;
  uint32_t level;
};

struct OrcTree {
  std::vector<ONode> nodes;
  std::vector<double> cmap;  // 4 doubles per entry
  bool assoc = false;
  int maxc = 0;
  int depth = 3;
  int count = 0;  // colour count (classification counter, then colormap size)
  size_t live = 0;
  double prune_thr = 0.0;
  double next_thr = 0.0;

  int kids() const { return assoc ? 16 : 8; }

  int32_t add_node(uint32_t id, uint32_t level, int32_t parent) {
    ONode n;
    for (int i = 0; i < 16; ++i) n.child[i] = -1;
    n.parent = parent;
    for (int c = 0; c < 4; ++c) n.total[c] = 0.0;
    n.qe = 0.0;
    n.nu = 0;
    n.color_number = 0;
    n.id = id;
    n.level = level;
    nodes.push_back(n);
    return (int32_t)nodes.size() - 1;
  }

  // rd_octree.cpp:50-67 PackChannels / NodeId.
  void pack(const D4& p, unsigned c[4]) const {
    for (int k = 0; k < 4; ++k) c[k] = scale_char(clamp_px(p.c[k]));
  }
  int node_id(const unsigned c[4], int index) const {
    int id = (int)((c[0] >> index) & 1u);
    id |= (int)((c[1] >> index) & 1u) << 1;
    id |= (int)((c[2] >> index) & 1u) << 2;
    if (assoc) id |= (int)((c[3] >> index) & 1u) << 3;
    return id;
  }

  // rd_octree.cpp:69-93 PruneChild: gate on the live counter.
  void prune_child(int32_t n) {
    for (int i = 0; i < kids(); ++i) {
      const int32_t ch = nodes[n].child[i];
      if (ch >= 0) prune_child(ch);
    }
    if (live > (size_t)maxc) {
      const int32_t p = nodes[n].parent;
      if (p >= 0) {
        nodes[p].nu += nodes[n].nu;
        for (int c = 0; c < 4; ++c) nodes[p].total[c] += nodes[n].total[c];
        nodes[p].child[nodes[n].id] = -1;
      }
      --live;
    }
  }
  // rd_octree.cpp:95-102.
  void prune_level(int32_t n) {
    for (int i = 0; i < kids(); ++i) {
      const int32_t ch = nodes[n].child[i];
      if (ch >= 0) prune_level(ch);
    }
    if ((int)nodes[n].level == depth) prune_child(n);
  }
  // rd_octree.cpp:104-111.
  void prune_to_cube(int32_t n) {
    for (int i = 0; i < kids(); ++i) {
      const int32_t ch = nodes[n].child[i];
      if (ch >= 0) prune_to_cube(ch);
    }
    if ((int)nodes[n].level > depth) prune_child(n);
  }
  // rd_octree.cpp:113-127.
  void reduce(int32_t n) {
    for (int i = 0; i < kids(); ++i) {
      const int32_t ch = nodes[n].child[i];
      if (ch >= 0) reduce(ch);
    }
    if (nodes[n].qe <= prune_thr) {
      prune_child(n);
    } else {
      if (nodes[n].nu > 0) ++count;
      if (nodes[n].qe < next_thr) next_thr = nodes[n].qe;
    }
  }
  // rd_octree.cpp:129-140: pre-order, cut off at the live counter.
  size_t flatten(int32_t n, size_t off, std::vector<double>& out) const {
    if (off >= live) return 0;
    out[off] = nodes[n].qe;
    size_t k = 1;
    for (int i = 0; i < kids(); ++i) {
      const int32_t ch = nodes[n].child[i];
      if (ch >= 0) k += flatten(ch, off + k, out);
    }
    return k;
  }
  // rd_octree.cpp:142-176.
  void define_colormap(int32_t n) {
    for (int i = 0; i < kids(); ++i) {
      const int32_t ch = nodes[n].child[i];
      if (ch >= 0) define_colormap(ch);
    }
    ONode& q = nodes[n];
    if (q.nu == 0) return;
    if ((size_t)count >= cmap.size() / 4) cmap.resize((cmap.size() / 4 * 2 + 1) * 4, 0.0);
    double* e = &cmap[4 * (size_t)count];
    const double alpha = safe_recip((double)q.nu);
    if (!assoc) {
      e[0] = (double)(float)(alpha * kRange * q.total[0]);
      e[1] = (double)(float)(alpha * kRange * q.total[1]);
      e[2] = (double)(float)(alpha * kRange * q.total[2]);
      e[3] = kRange;
    } else {
      const double opacity = alpha * kRange * q.total[3];
      e[3] = (double)(float)opacity;
      if (e[3] == kRange) {
        e[0] = (double)(float)(alpha * kRange * q.total[0]);
        e[1] = (double)(float)(alpha * kRange * q.total[1]);
        e[2] = (double)(float)(alpha * kRange * q.total[2]);
      } else {
        const double gamma = safe_recip(kScale * e[3]);
        e[0] = (double)(float)(alpha * gamma * kRange * q.total[0]);
        e[1] = (double)(float)(alpha * gamma * kRange * q.total[1]);
        e[2] = (double)(float)(alpha * gamma * kRange * q.total[2]);
      }
    }
    q.color_number = (uint32_t)count;
    ++count;
  }
  // rd_octree.cpp:178-217: post-order scan with per-channel early exits.
  void closest(int32_t n, const D4& t, double* dist, int* cn) const {
    for (int i = 0; i < kids(); ++i) {
      const int32_t ch = nodes[n].child[i];
      if (ch >= 0) closest(ch, t, dist, cn);
    }
    const ONode& q = nodes[n];
    if (q.nu == 0) return;
    const double* p = &cmap[4 * (size_t)q.color_number];
    double alpha = 1.0, beta = 1.0;
    if (assoc) {
      alpha = kScale * p[3];
      beta = kScale * t.c[3];
    }
    double v = alpha * p[0] - beta * t.c[0];
    double d = v * v;
    if (d <= *dist) {
      v = alpha * p[1] - beta * t.c[1];
      d += v * v;
      if (d <= *dist) {
        v = alpha * p[2] - beta * t.c[2];
        d += v * v;
        if (d <= *dist) {
          if (assoc) {
            v = p[3] - t.c[3];
            d += v * v;
          }
          if (d <= *dist) {
            *dist = d;
            *cn = (int)q.color_number;
          }
        }
      }
    }
  }
  // rd_riemersma_cpu.cpp:82-92: descend over bit 7..1, search parent's subtree.
  int lookup(const D4& t) const {
    unsigned c[4];
    pack(t, c);
    int32_t n = 0;
    for (int i = kMaxDepth - 1; i > 0; --i) {
      const int32_t ch = nodes[n].child[node_id(c, i)];
      if (ch < 0) break;
      n = ch;
    }
    double dist = 4.0 * ((kRange + 1.0) * (kRange + 1.0)) + 1.0;
    int idx = 0;
    closest(nodes[n].parent, t, &dist, &idx);
    return idx;
  }

  // rd_octree.cpp:305-362: one scanline, run-length encoded.
  void classify_row(const float* row, int64_t width, int level_limit, int colors_level) {
    const double midpoint = kRange / 2.0;
    int64_t x = 0;
    while (x < width) {
      int64_t run = 1;
      while (x + run < width && row[4 * x] == row[4 * (x + run)] &&
             row[4 * x + 1] == row[4 * (x + run) + 1] &&
             row[4 * x + 2] == row[4 * (x + run) + 2] &&
             row[4 * x + 3] == row[4 * (x + run) + 3])
        ++run;
      const D4 px = assoc_px(assoc, row + 4 * x);
      unsigned c[4];
      pack(px, c);
      int index = kMaxDepth - 1;
      double bisect = (kRange + 1.0) / 2.0;
      double mid[4] = {midpoint, midpoint, midpoint, midpoint};
      int32_t n = 0;
      for (int level = 1; level <= level_limit; ++level) {
        bisect *= 0.5;
        const int id = node_id(c, index);
        for (int k = 0; k < 4; ++k) mid[k] += (id & (1 << k)) ? bisect : -bisect;
        if (nodes[n].child[id] < 0) {
          const int32_t made = add_node((uint32_t)id, (uint32_t)level, n);
          nodes[n].child[id] = made;
          ++live;
          if (level == colors_level) ++count;
        }
        n = nodes[n].child[id];
        const double e0 = kScale * (px.c[0] - mid[0]);
        const double e1 = kScale * (px.c[1] - mid[1]);
        const double e2 = kScale * (px.c[2] - mid[2]);
        const double e3 = assoc ? kScale * (px.c[3] - mid[3]) : 0.0;
        double dist = e0 * e0 + e1 * e1 + e2 * e2 + e3 * e3;
        if (std::isnan(dist)) dist = 0.0;
        nodes[n].qe += (double)run * std::sqrt(dist);
        nodes[0].qe += nodes[n].qe;  // IM's double count into the root
        --index;
      }
      ONode& leaf = nodes[n];
      leaf.nu += (uint64_t)run;
      leaf.total[0] += (double)run * kScale * (double)clamp_px(px.c[0]);
      leaf.total[1] += (double)run * kScale * (double)clamp_px(px.c[1]);
      leaf.total[2] += (double)run * kScale * (double)clamp_px(px.c[2]);
      leaf.total[3] += (double)run * kScale * (double)clamp_px(assoc ? px.c[3] : kRange);
      x += run;
    }
  }
};

namespace {

int quantize_error_cmp(const void* a, const void* b) {  // rd_octree.cpp:23-29
  const double p = *(const double*)a;
  const double q = *(const double*)b;
  if (p > q) return 1;
  if (std::fabs(q - p) <= kEps) return 0;
  return -1;
}

}  // namespace

extern "C" OrcTree* orc_tree_build(const float* rgba, int64_t w, int64_t h,
                                   int max_colors, int assoc, int gray) {
  OrcTree* t = new OrcTree();
  t->cmap.assign((size_t)kMaxColormap * 4, 0.0);
  for (size_t i = 0; i < (size_t)kMaxColormap; ++i) t->cmap[4 * i + 3] = kRange;
  t->assoc = assoc != 0;
  t->maxc = max_colors <= 0 ? kMaxColormap : std::min(max_colors, kMaxColormap);
  // rd_octree.cpp:255-275: depth = Log4(colours)+2, adjusted.
  int depth = 1;
  for (size_t c = (size_t)t->maxc; c != 0; ++depth) c >>= 2;
  if (depth > 2) --depth;
  if (t->assoc && depth > 5) --depth;
  if (gray) depth = kMaxDepth;
  if (depth > kMaxDepth) depth = kMaxDepth;
  if (depth < 2) depth = 2;
  t->depth = depth;
  t->add_node(0, 0, -1);
  t->live = 1;
  t->nodes[0].parent = 0;  // root is its own parent (rd_octree.cpp:285)

  int64_t y = 0;
  for (; y < h; ++y) {  // loop 1: full depth, colours counted at level 8
    if (t->live > kMaxQNodes) {
      t->prune_level(0);
      --t->depth;
    }
    t->classify_row(rgba + 4 * y * w, w, kMaxDepth, kMaxDepth);
    if (t->count > t->maxc) {
      t->prune_to_cube(0);
      break;
    }
  }
  ++y;
  for (; y < h; ++y) {  // loop 2: cube depth
    if (t->live > kMaxQNodes) {
      t->prune_level(0);
      --t->depth;
    }
    t->classify_row(rgba + 4 * y * w, w, t->depth, t->depth);
  }

  // rd_octree.cpp:388-404 ReduceImageColors.
  if (t->count > t->maxc) {
    t->next_thr = 0.0;
    if (t->live > 0) {
      std::vector<double> err(t->live, 0.0);
      (void)t->flatten(0, 0, err);
      const size_t n = t->live;
      std::qsort(err.data(), n, sizeof(double), quantize_error_cmp);
      const size_t aggressive = (110u * ((size_t)t->maxc + 1u)) / 100u;
      if (n > aggressive) t->next_thr = err[n - aggressive];
    }
  }
  while (t->count > t->maxc) {
    t->prune_thr = t->next_thr;
    t->next_thr = t->nodes[0].qe - 1.0;
    t->count = 0;
    t->reduce(0);
  }
  t->count = 0;
  t->define_colormap(0);
  return t;
}

extern "C" void orc_tree_free(OrcTree* t) { delete t; }
extern "C" int orc_tree_color_count(const OrcTree* t) { return t->count; }
extern "C" int64_t orc_tree_node_count(const OrcTree* t) { return (int64_t)t->nodes.size(); }
extern "C" int orc_tree_depth(const OrcTree* t) { return t->depth; }
extern "C" int orc_tree_assoc(const OrcTree* t) { return t->assoc ? 1 : 0; }

extern "C" void orc_tree_colormap(const OrcTree* t, double* out) {
  std::memcpy(out, t->cmap.data(), sizeof(double) * 4 * (size_t)t->count);
}

extern "C" void orc_tree_nodes(const OrcTree* t, int32_t* child16, int32_t* parent,
                               uint32_t* color_number, uint64_t* number_unique,
                               double* total4, double* qe, uint32_t* id,
                               uint32_t* level) {
  for (size_t i = 0; i < t->nodes.size(); ++i) {
    const ONode& n = t->nodes[i];
    if (child16) for (int c = 0; c < 16; ++c) child16[16 * i + c] = n.child[c];
    if (parent) parent[i] = n.parent;
    if (color_number) color_number[i] = n.color_number;
    if (number_unique) number_unique[i] = n.nu;
    if (total4) for (int c = 0; c < 4; ++c) total4[4 * i + c] = n.total[c];
    if (qe) qe[i] = n.qe;
    if (id) id[i] = n.id;
    if (level) level[i] = n.level;
  }
}

extern "C" int orc_tree_lookup(const OrcTree* t, const double target[4]) {
  D4 d;
  for (int c = 0; c < 4; ++c) d.c[c] = target[c];
  return t->lookup(d);
}

// ===========================================================================
// Numeric contract exports.
// ===========================================================================
extern "C" void orc_error_weights(double out[16]) { raw_weights(out); }
extern "C" void orc_scaled_weights(double diffusion, double out[16]) {
  scaled_weights(diffusion, out);
}
extern "C" unsigned orc_scale_char(float q) { return scale_char(q); }
extern "C" float orc_clamp_pixel(double v) { return clamp_px(v); }

// ===========================================================================
// Curve.
// ===========================================================================
// rd_riemersma.h:88-94: level = ceil(log2(max(w,h))), via log2 on the host.
extern "C" int orc_curve_level(int64_t w, int64_t h) {
  const int64_t extent = w > h ? w : h;
  if (extent <= 0) return 0;
  int level = (int)std::log2((double)extent);
  if (((int64_t)1 << level) < extent) ++level;
  return level;
}

// rd_riemersma_cpu.cpp:167-189: textbook d2xy (the reference's orientation).
extern "C" void orc_hilbert_point(int64_t n, int level, int* ox, int* oy) {
  int64_t x = 0, y = 0;
  const int64_t span = (int64_t)1 << level;
  for (int64_t s = 1; s < span; s <<= 1) {
    const int64_t rx = 1 & (n / 2);
    const int64_t ry = 1 & (n ^ rx);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      std::swap(x, y);
    }
    x += s * rx;
    y += s * ry;
    n /= 4;
  }
  *ox = (int)x;
  *oy = (int)y;
}

namespace {

// IM's Riemersma() recursion, as the sequence of moves per (level, heading).
// Directions: 0=W 1=E 2=N 3=S.  At level 1 a heading expands into three
// visits (each followed by a move); above level 1 into
// R(a) D(m1) R(b) D(m2) R(c) D(m3) R(d).  Same schedule as the reference's
// tables (rd_riemersma.h:103-114), written as strings here.
const char* kLeafMoves[4] = {"ESW", "WNE", "SEN", "NWS"};
const char* kNodeSteps[4] = {"NeWsWwS", "SwEnEeN", "WsNeNnE", "EnSwSsW"};

int dir_of(char c) {
  switch (c) {
    case 'W': case 'w': return 0;
    case 'E': case 'e': return 1;
    case 'N': case 'n': return 2;
    default: return 3;
  }
}

template <typename Fn>
void recurse(int level, int heading, int* x, int* y, Fn&& visit) {
  auto move = [&](int d) {
    if (d == 0) --*x;
    else if (d == 1) ++*x;
    else if (d == 2) --*y;
    else ++*y;
  };
  if (level == 1) {
    for (int k = 0; k < 3; ++k) {
      visit(*x, *y);
      move(dir_of(kLeafMoves[heading][k]));
    }
    return;
  }
  const char* s = kNodeSteps[heading];
  for (int k = 0; k < 7; ++k) {
    const char c = s[k];
    if (c >= 'A' && c <= 'Z') {
      recurse(level - 1, dir_of(c), x, y, visit);
    } else {
      visit(*x, *y);
      move(dir_of(c));
    }
  }
}

// Visit sequence of the sequential walk: the recursion started heading North,
// then the trailing ForgetGravity visit at the resting cursor
// (rd_riemersma_cpu.cpp:403-416).
template <typename Fn>
void walk_order(int level, Fn&& visit) {
  int x = 0, y = 0;
  if (level > 0) recurse(level, 2, &x, &y, visit);
  visit(x, y);
}

}  // namespace

extern "C" int64_t orc_recursive_walk(int level, int* xs, int* ys) {
  int64_t k = 0;
  walk_order(level, [&](int x, int y) {
    xs[k] = x;
    ys[k] = y;
    ++k;
  });
  return k;
}

extern "C" int64_t orc_seq_order(int64_t w, int64_t h, int32_t* order) {
  const int level = orc_curve_level(w, h);
  int64_t k = 0;
  walk_order(level, [&](int x, int y) {
    if (x >= 0 && y >= 0 && x < w && y < h) order[k++] = (int32_t)(y * w + x);
  });
  return k;
}

// rd_riemersma_cpu.cpp:269-356: in-bounds of [0, 4^L-1) in ascending n, then
// the origin, owner = last writer.
extern "C" int64_t orc_block_order(int64_t w, int64_t h, int32_t* order, int32_t* owner) {
  const int level = orc_curve_level(w, h);
  for (int64_t p = 0; p < w * h; ++p) owner[p] = -1;
  if (level <= 0) return 0;
  const int64_t leaves = ((int64_t)1 << (2 * level)) - 1;
  int64_t k = 0;
  for (int64_t n = 0; n < leaves; ++n) {
    int x, y;
    orc_hilbert_point(n, level, &x, &y);
    if (x < w && y < h) order[k++] = (int32_t)(y * w + x);
  }
  order[k++] = 0;
  for (int64_t i = 0; i < k; ++i) owner[order[i]] = (int32_t)i;
  return k;
}

// ===========================================================================
// Sequential walk.  rd_riemersma_cpu.cpp:39-113 (Visit) and :359-417.
// ===========================================================================
namespace {

struct SeqState {
  const OrcTree* t;
  double sw[kQueue];
  D4 err[kQueue];
  std::vector<int> memo;
  bool use_cache;
  int64_t misses = 0;
};

int visit(SeqState& s, float* px) {
  const OrcTree& t = *s.t;
  const bool assoc = t.assoc;
  D4 p = assoc_px(assoc, px);
  for (int i = 0; i < kQueue; ++i) {
    const double w = s.sw[i];
    p.c[0] += w * s.err[i].c[0];
    p.c[1] += w * s.err[i].c[1];
    p.c[2] += w * s.err[i].c[2];
    if (assoc) p.c[3] += w * s.err[i].c[3];
  }
  p.c[0] = (double)clamp_px(p.c[0]);
  p.c[1] = (double)clamp_px(p.c[1]);
  p.c[2] = (double)clamp_px(p.c[2]);
  if (assoc) p.c[3] = (double)clamp_px(p.c[3]);
  const int key = memo_key(assoc, p);
  int idx = s.use_cache ? s.memo[(size_t)key] : -1;
  if (idx < 0) {
    idx = t.lookup(p);
    ++s.misses;
    if (s.use_cache) s.memo[(size_t)key] = idx;
  }
  const double* e = &t.cmap[4 * (size_t)idx];
  px[0] = (float)e[0];
  px[1] = (float)e[1];
  px[2] = (float)e[2];
  if (assoc) px[3] = (float)e[3];
  for (int i = 0; i < kQueue - 1; ++i) s.err[i] = s.err[i + 1];
  const D4 col = assoc_info(assoc, e);
  s.err[kQueue - 1].c[0] = p.c[0] - col.c[0];
  s.err[kQueue - 1].c[1] = p.c[1] - col.c[1];
  s.err[kQueue - 1].c[2] = p.c[2] - col.c[2];
  if (assoc) s.err[kQueue - 1].c[3] = p.c[3] - col.c[3];
  return idx;
}

int64_t walk_seq_impl(const OrcTree* t, float* rgba, int64_t w, int64_t h,
                      double diffusion, int use_cache, int32_t* index_out) {
  SeqState s;
  s.t = t;
  scaled_weights(diffusion, s.sw);
  s.use_cache = use_cache != 0;
  if (s.use_cache) s.memo.assign(t->assoc ? ((size_t)1 << 24) : ((size_t)1 << 18), -1);
  const int level = orc_curve_level(w, h);
  int64_t k = 0;
  walk_order(level, [&](int x, int y) {
    if (x < 0 || y < 0 || x >= w || y >= h) return;
    const int idx = visit(s, rgba + 4 * ((int64_t)y * w + x));
    if (index_out) index_out[k] = idx;
    ++k;
  });
  return s.misses;
}

}  // namespace

extern "C" void orc_walk_seq(const OrcTree* t, float* rgba, int64_t w, int64_t h,
                             double diffusion, int use_cache) {
  walk_seq_impl(t, rgba, w, h, diffusion, use_cache, nullptr);
}

extern "C" void orc_walk_seq_index(const OrcTree* t, float* rgba, int64_t w, int64_t h,
                                   double diffusion, int use_cache, int32_t* index_out) {
  walk_seq_impl(t, rgba, w, h, diffusion, use_cache, index_out);
}

// Number of memo misses (full octree searches) of one walk.
extern "C" int64_t orc_walk_seq_misses(const OrcTree* t, float* rgba, int64_t w, int64_t h,
                                       double diffusion) {
  return walk_seq_impl(t, rgba, w, h, diffusion, 1, nullptr);
}

extern "C" int orc_walk_seq_many(OrcTree* const* trees, float* const* images,
                                 const int64_t* ws, const int64_t* hs, int n,
                                 double diffusion, int threads) {
  std::vector<int> ord((size_t)n);
  for (int i = 0; i < n; ++i) ord[(size_t)i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    return ws[a] * hs[a] > ws[b] * hs[b];
  });
  std::vector<std::thread> pool;
  std::vector<int> next(1, 0);
  std::mutex* mu = new std::mutex();
  if (threads < 1) threads = 1;
  for (int k = 0; k < threads; ++k) {
    pool.emplace_back([&]() {
      for (;;) {
        int j;
        {
          std::lock_guard<std::mutex> g(*mu);
          if (next[0] >= n) return;
          j = ord[(size_t)next[0]++];
        }
        walk_seq_impl(trees[j], images[j], ws[j], hs[j], diffusion, 1, nullptr);
      }
    });
  }
  for (auto& th : pool) th.join();
  delete mu;
  return 0;
}

// ===========================================================================
// Block engine.  rd_riemersma_cpu.cpp:477-805 (unfused path and the rgba64
// fused scatter).  CPU semantics for associated alpha (one association).
// ===========================================================================
namespace {

int blocks_indices(const OrcTree* t, const float* batch, int64_t w, int64_t h,
                   int frames, int block, double diffusion,
                   std::vector<int32_t>& curve, std::vector<int32_t>& owner,
                   std::vector<uint8_t>& chosen) {
  if (t->count > 256) return -1;
  if (frames < 1 || block < kQueue) return -1;
  double sw[kQueue];
  scaled_weights(diffusion, sw);
  curve.assign((size_t)(w * h + 1), 0);
  owner.assign((size_t)(w * h), -1);
  const int64_t n = orc_block_order(w, h, curve.data(), owner.data());
  curve.resize((size_t)n);
  chosen.assign((size_t)n * (size_t)frames, 0);
  const bool assoc = t->assoc;
  const int64_t pixels = w * h;
  const int64_t nblocks = (n + block - 1) / block;
  for (int f = 0; f < frames; ++f) {
    const float* frame = batch + 4 * (size_t)f * (size_t)pixels;
    for (int64_t b = 0; b < nblocks; ++b) {
      D4 err[kQueue];
      const int64_t begin = b * block;
      const int64_t end = std::min(begin + (int64_t)block, n);
      for (int64_t i = begin; i < end; ++i) {
        D4 p = assoc_px(assoc, frame + 4 * (size_t)curve[(size_t)i]);
        for (int k = 0; k < kQueue; ++k) {
          const double wk = sw[k];
          p.c[0] += wk * err[k].c[0];
          p.c[1] += wk * err[k].c[1];
          p.c[2] += wk * err[k].c[2];
          if (assoc) p.c[3] += wk * err[k].c[3];
        }
        p.c[0] = (double)clamp_px(p.c[0]);
        p.c[1] = (double)clamp_px(p.c[1]);
        p.c[2] = (double)clamp_px(p.c[2]);
        if (assoc) p.c[3] = (double)clamp_px(p.c[3]);
        const int idx = t->lookup(p);
        chosen[(size_t)f * (size_t)n + (size_t)i] = (uint8_t)idx;
        for (int k = 0; k < kQueue - 1; ++k) err[k] = err[k + 1];
        const D4 col = assoc_info(assoc, &t->cmap[4 * (size_t)idx]);
        err[kQueue - 1].c[0] = p.c[0] - col.c[0];
        err[kQueue - 1].c[1] = p.c[1] - col.c[1];
        err[kQueue - 1].c[2] = p.c[2] - col.c[2];
        if (assoc) err[kQueue - 1].c[3] = p.c[3] - col.c[3];
      }
    }
  }
  return 0;
}

}  // namespace

extern "C" int orc_walk_blocks(const OrcTree* t, float* batch, int64_t w, int64_t h,
                               int frames, int block, double diffusion,
                               uint8_t* index_out) {
  std::vector<int32_t> curve, owner;
  std::vector<uint8_t> chosen;
  if (blocks_indices(t, batch, w, h, frames, block, diffusion, curve, owner, chosen) != 0)
    return -1;
  const int64_t n = (int64_t)curve.size();
  const int64_t pixels = w * h;
  for (int f = 0; f < frames; ++f) {
    float* frame = batch + 4 * (size_t)f * (size_t)pixels;
    for (int64_t i = 0; i < n; ++i) {
      const int32_t pix = curve[(size_t)i];
      if (owner[(size_t)pix] != i) continue;
      const double* e = &t->cmap[4 * (size_t)chosen[(size_t)f * (size_t)n + (size_t)i]];
      frame[4 * pix + 0] = (float)e[0];
      frame[4 * pix + 1] = (float)e[1];
      frame[4 * pix + 2] = (float)e[2];
      if (t->assoc) frame[4 * pix + 3] = (float)e[3];
    }
  }
  if (index_out) std::memcpy(index_out, chosen.data(), chosen.size());
  return 0;
}

extern "C" int orc_walk_blocks_u16(const OrcTree* t, const float* batch, int64_t w,
                                   int64_t h, int frames, int block, double diffusion,
                                   uint16_t* out) {
  std::vector<int32_t> curve, owner;
  std::vector<uint8_t> chosen;
  if (blocks_indices(t, batch, w, h, frames, block, diffusion, curve, owner, chosen) != 0)
    return -1;
  const int64_t n = (int64_t)curve.size();
  const int64_t pixels = w * h;
  for (int f = 0; f < frames; ++f) {
    const float* src = batch + 4 * (size_t)f * (size_t)pixels;
    uint16_t* u = out + 4 * (size_t)f * (size_t)pixels;
    for (int64_t i = 0; i < n; ++i) {
      const int32_t pix = curve[(size_t)i];
      if (owner[(size_t)pix] != i) continue;
      const double* e = &t->cmap[4 * (size_t)chosen[(size_t)f * (size_t)n + (size_t)i]];
      u[4 * pix + 0] = (uint16_t)(float)e[0];
      u[4 * pix + 1] = (uint16_t)(float)e[1];
      u[4 * pix + 2] = (uint16_t)(float)e[2];
      u[4 * pix + 3] = t->assoc ? (uint16_t)(float)e[3] : (uint16_t)src[4 * pix + 3];
    }
    for (int64_t pix = 0; pix < pixels; ++pix) {  // unvisited: source value
      if (owner[(size_t)pix] >= 0) continue;
      for (int c = 0; c < 4; ++c) u[4 * pix + c] = (uint16_t)src[4 * pix + c];
    }
  }
  return 0;
}

// ===========================================================================
// Video wire formats of the block engine: planar 8-bit YUV in and out.
// ===========================================================================
namespace {

// SwsYuvToRgb16, include/rd_types.h:75-97, in the reference's statement order.
uint16_t sws_clip16(int x) {
  const int v = x + (1 << 15);
  return (uint16_t)(v < 0 ? 0 : (v > 65535 ? 65535 : v));
}
void sws_yuv_to_rgb16(unsigned y, unsigned u, unsigned v, uint16_t* out) {
  int yy = (int)((y - 128u) * 512u);
  const int uu = (int)((u - 128u) * 512u);
  const int vv = (int)((v - 128u) * 512u);
  yy += 0x10000;
  yy -= 8192;
  yy *= 9539;
  yy += (1 << 13) - (1 << 29);
  out[0] = sws_clip16((vv * 13075 + yy) >> 14);
  out[1] = sws_clip16((vv * -6660 + uu * -3209 + yy) >> 14);
  out[2] = sws_clip16((uu * 16525 + yy) >> 14);
}

// d_chroma_bilinear, src/rd_blocks_cuda.cu:651-675.
unsigned chroma_bilinear(const uint8_t* plane, int64_t stride, int64_t cx, int64_t cy, int ox,
                         int oy, int64_t w2, int64_t h2) {
  const int64_t x1 = cx + 1 < w2 ? cx + 1 : cx;
  const int64_t y1 = cy + 1 < h2 ? cy + 1 : cy;
  const int a = plane[cy * stride + cx], b = plane[cy * stride + x1];
  const int c = plane[y1 * stride + cx], d = plane[y1 * stride + x1];
  const int top = a * (2 - ox) + b * ox;
  const int bot = c * (2 - ox) + d * ox;
  return (unsigned)((top * (2 - oy) + bot * oy + 2) >> 2);
}

// The fused scatter's RGB -> YUV 4:4:4 (src/rd_riemersma_cpu.cpp:718-730).
void rgb_to_yuv444(float fr, float fg, float fb, uint8_t* y, uint8_t* u, uint8_t* v) {
  const int r8 = (int)fr >> 8, g8 = (int)fg >> 8, b8 = (int)fb >> 8;
  const int yv = ((66 * r8 + 129 * g8 + 25 * b8 + 128) >> 8) + 16;
  const int uv = ((-38 * r8 - 74 * g8 + 112 * b8 + 128) >> 8) + 128;
  const int vv = ((112 * r8 - 94 * g8 - 18 * b8 + 128) >> 8) + 128;
  *y = (uint8_t)(yv < 0 ? 0 : (yv > 255 ? 255 : yv));
  *u = (uint8_t)(uv < 0 ? 0 : (uv > 255 ? 255 : uv));
  *v = (uint8_t)(vv < 0 ? 0 : (vv > 255 ? 255 : vv));
}

}  // namespace

// Planar frames (fmt 7: 4:4:4, 8: 4:2:0 with even w, h) widened to RgbaF, as
// the reader and the gathers do (RawYuv444ToFloats / gather_pixel,
// src/rd_riemersma_cpu.cpp:496-507; BlkGatherYuv420Kernel,
// src/rd_blocks_cuda.cu:699-740): alpha 65535.
extern "C" int orc_yuv_to_rgbaf(const uint8_t* planes, int64_t w, int64_t h, int frames, int fmt,
                                float* out) {
  if (fmt != 7 && fmt != 8) return -1;
  if (fmt == 8 && ((w & 1) || (h & 1))) return -1;
  const int64_t pp = w * h, w2 = w / 2, h2 = h / 2;
  const int64_t fbytes = fmt == 7 ? 3 * pp : pp + 2 * w2 * h2;
  for (int f = 0; f < frames; ++f) {
    const uint8_t* s = planes + (size_t)f * (size_t)fbytes;
    float* o = out + 4 * (size_t)f * (size_t)pp;
    for (int64_t p = 0; p < pp; ++p) {
      unsigned u, v;
      if (fmt == 7) {
        u = s[pp + p];
        v = s[2 * pp + p];
      } else {
        const int64_t y = p / w, x = p % w;
        u = chroma_bilinear(s + pp, w2, x >> 1, y >> 1, (int)(x & 1), (int)(y & 1), w2, h2);
        v = chroma_bilinear(s + pp + w2 * h2, w2, x >> 1, y >> 1, (int)(x & 1), (int)(y & 1), w2,
                            h2);
      }
      uint16_t rgb[3];
      sws_yuv_to_rgb16(s[p], u, v, rgb);
      o[4 * p + 0] = (float)rgb[0];
      o[4 * p + 1] = (float)rgb[1];
      o[4 * p + 2] = (float)rgb[2];
      o[4 * p + 3] = (float)65535.0;
    }
  }
  return 0;
}

// The block walk with planar 4:4:4 output (the CPU engine's fused scatter,
// src/rd_riemersma_cpu.cpp:711-730, and its unvisited pixels, :776-792):
// Y, Cb, Cr planes of w*h bytes per frame.
extern "C" int orc_walk_blocks_yuv444(const OrcTree* t, const float* batch, int64_t w, int64_t h,
                                      int frames, int block, double diffusion, uint8_t* out) {
  std::vector<int32_t> curve, owner;
  std::vector<uint8_t> chosen;
  if (blocks_indices(t, batch, w, h, frames, block, diffusion, curve, owner, chosen) != 0)
    return -1;
  const int64_t n = (int64_t)curve.size();
  const int64_t pp = w * h;
  for (int f = 0; f < frames; ++f) {
    const float* src = batch + 4 * (size_t)f * (size_t)pp;
    uint8_t* o = out + 3 * (size_t)f * (size_t)pp;
    for (int64_t i = 0; i < n; ++i) {
      const int32_t pix = curve[(size_t)i];
      if (owner[(size_t)pix] != i) continue;
      const double* e = &t->cmap[4 * (size_t)chosen[(size_t)f * (size_t)n + (size_t)i]];
      rgb_to_yuv444((float)e[0], (float)e[1], (float)e[2], o + pix, o + pp + pix,
                    o + 2 * pp + pix);
    }
    for (int64_t pix = 0; pix < pp; ++pix) {  // unvisited: the source, converted
      if (owner[(size_t)pix] >= 0) continue;
      rgb_to_yuv444(src[4 * pix + 0], src[4 * pix + 1], src[4 * pix + 2], o + pix, o + pp + pix,
                    o + 2 * pp + pix);
    }
  }
  return 0;
}

// ===========================================================================
// Threshold-screen plugin dithers (src/rd_dithers_ordered.cc, src/rd_dithers.cc).
// ===========================================================================
namespace {

// BuildVoidCluster, src/rd_dithers_ordered.cc:220-330.
std::vector<uint16_t> vc_ranks(int n) {
  const size_t np = (size_t)n * n;
  const double k5[5] = {0.12012, 0.23387, 0.29202, 0.23387, 0.12012};
  std::vector<double> on(np), tmp(np), filt(np);
  uint64_t s = 0x9E3779B97F4A7C15ull;
  for (size_t i = 0; i < np; ++i) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    on[i] = ((s >> 33) % 10 == 0) ? 1.0 : 0.0;
  }
  auto wrap = [n](int a) { return (a % n + n) % n; };
  auto blur = [&]() {
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        double acc = 0.0;
        for (int k = 0; k < 5; ++k) acc += k5[k] * on[(size_t)y * n + wrap(x + k - 2)];
        tmp[(size_t)y * n + x] = acc;
      }
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        double acc = 0.0;
        for (int k = 0; k < 5; ++k) acc += k5[k] * tmp[(size_t)wrap(y + k - 2) * n + x];
        filt[(size_t)y * n + x] = acc;
      }
  };
  auto amax_on = [&]() {
    size_t b = np;
    double bv = 0.0;
    for (size_t i = 0; i < np; ++i)
      if (on[i] > 0.5 && (b == np || filt[i] > bv)) { bv = filt[i]; b = i; }
    return b;
  };
  auto amin_off = [&]() {
    size_t b = np;
    double bv = 0.0;
    for (size_t i = 0; i < np; ++i)
      if (on[i] <= 0.5 && (b == np || filt[i] < bv)) { bv = filt[i]; b = i; }
    return b;
  };
  for (size_t it = 0; it < 4 * np; ++it) {
    blur();
    const size_t t = amax_on();
    if (t == np) break;
    on[t] = 0.0;
    blur();
    const size_t v = amin_off();
    if (v == np) break;
    on[v] = 1.0;
    if (t == v) break;
  }
  const std::vector<double> eq = on;
  size_t dots = 0;
  for (size_t i = 0; i < np; ++i) if (on[i] > 0.5) ++dots;
  std::vector<uint16_t> rank(np, 0);
  size_t r = dots;
  while (r > 0) {
    --r;
    blur();
    const size_t t = amax_on();
    if (t == np) break;
    on[t] = 0.0;
    rank[t] = (uint16_t)(r + 1);
  }
  on = eq;
  for (size_t q = dots; q < np; ++q) {
    blur();
    const size_t v = amin_off();
    if (v == np) break;
    on[v] = 1.0;
    rank[v] = (uint16_t)(q + 1);
  }
  return rank;
}

// OrderedPick, src/rd_dithers_ordered.cc:79-145.
int ordered_pick(const double* ent, int count, float r, float g, float b, float t) {
  int first = 0, second = 0;
  float d1 = -1.0f, d2 = -1.0f;
  for (int i = 0; i < count; ++i) {
    const float dr = r - (float)ent[4 * i], dg = g - (float)ent[4 * i + 1];
    const float db = b - (float)ent[4 * i + 2];
    const float d = dr * dr + dg * dg + db * db;
    if (d1 < 0.0f || d < d1) {
      d2 = d1; second = first; d1 = d; first = i;
    } else if (d2 < 0.0f || d < d2) {
      d2 = d; second = i;
    }
  }
  if (d2 < 0.0f) return first;
  if (d2 <= 0.0f) return second;
  const float s1 = std::sqrt(d1), s2 = std::sqrt(d2);
  return t < s1 / (s1 + s2) ? second : first;
}

const unsigned char kBayer[64] = {0,  32, 8,  40, 2,  34, 10, 42, 48, 16, 56, 24, 50, 18, 58, 26,
                                  12, 44, 4,  36, 14, 46, 6,  38, 60, 28, 52, 20, 62, 30, 54, 22,
                                  3,  35, 11, 43, 1,  33, 9,  41, 51, 19, 59, 27, 49, 17, 57, 25,
                                  15, 47, 7,  39, 13, 45, 5,  37, 63, 31, 55, 23, 61, 29, 53, 21};
const unsigned char kCell[64] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0,
                                 0, 0, 1, 1, 1, 1, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0,
                                 0, 1, 1, 1, 1, 1, 1, 0, 0, 0, 1, 1, 1, 1, 0, 0,
                                 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

}  // namespace

// The plugin's run over `frames` RgbaF frames in place: screen 0 bayer-ordered,
// 1 void-and-cluster (64), 2 void-and-cluster-fast (32), 3 clustered-dot.
// Returns 0, or -1 for no palette / an empty frame (Prepare).
extern "C" int orc_dither_screen(const double* ent, int count, float* rgba, int64_t w, int64_t h,
                                 int frames, int screen, double diffusion) {
  if (count <= 0 || w <= 0 || h <= 0 || screen < 0 || screen > 3) return -1;
  static std::mutex mu;
  static std::vector<uint16_t> v64, v32;
  {
    std::lock_guard<std::mutex> g(mu);
    if (screen == 1 && v64.empty()) v64 = vc_ranks(64);
    if (screen == 2 && v32.empty()) v32 = vc_ranks(32);
  }
  int lo_i = 0, hi_i = 0;
  double lo = 0.0, hi = 0.0;
  for (int i = 0; i < count; ++i) {  // rd_dithers.cc:317-328
    const double l = 0.299 * ent[4 * i] + 0.587 * ent[4 * i + 1] + 0.114 * ent[4 * i + 2];
    if (i == 0 || l < lo) { lo = l; lo_i = i; }
    if (i == 0 || l > hi) { hi = l; hi_i = i; }
  }
  const double span = hi > lo ? hi - lo : 1.0;
  const int64_t pp = w * h;
  for (int f = 0; f < frames; ++f)
    for (int64_t y = 0; y < h; ++y)
      for (int64_t x = 0; x < w; ++x) {
        float* px = rgba + 4 * ((size_t)f * pp + (size_t)(y * w + x));
        int best;
        if (screen == 3) {
          const double lum = 0.299 * px[0] + 0.587 * px[1] + 0.114 * px[2];
          const double t = (lum - lo) / span;
          const double cell = (double)kCell[(y & 7) * 8 + (x & 7)];
          double dark = cell / 32.0;
          if (cell == 0) dark = 0.0;
          dark = 0.5 + (dark - 0.5) * diffusion;
          best = (1.0 - t) >= dark ? lo_i : hi_i;
        } else {
          float t;
          if (screen == 0) {
            t = ((float)kBayer[(y & 7) * 8 + (x & 7)] + 0.5f) / 64.0f;
          } else {
            const int n = screen == 1 ? 64 : 32;
            const std::vector<uint16_t>& tab = screen == 1 ? v64 : v32;
            t = (float)tab[(size_t)(y % n) * n + (size_t)(x % n)] / (float)((double)n * n);
          }
          const float sc = 0.5f + (t - 0.5f) * (float)diffusion;
          best = ordered_pick(ent, count, px[0], px[1], px[2], sc);
        }
        px[0] = (float)ent[4 * best];
        px[1] = (float)ent[4 * best + 1];
        px[2] = (float)ent[4 * best + 2];
      }
  return 0;
}

// ===========================================================================
// Jacobi approximation.  src/rd_approx_cuda.cu:63-384 restated on the CPU.
// ===========================================================================
extern "C" void orc_inverse_filter(int taps, double diffusion, double* g) {
  double w[kQueue];
  raw_weights(w);
  g[0] = 1.0;
  for (int d = 1; d <= taps; ++d) {
    double acc = 0.0;
    const int kmax = d < kQueue ? d : kQueue;
    for (int k = 1; k <= kmax; ++k) acc += (diffusion / 16.0) * w[k - 1] * g[d - k];
    g[d] = acc;
  }
}

namespace {

// rd_cuda_common.cuh:57-72: the clamped double target, associated from double.
D4 assoc_from_double(bool assoc, double r, double g, double b, double a) {
  D4 o;
  const double ca = (double)clamp_px(a);
  o.c[3] = ca;
  if (!assoc || a == kRange) {
    o.c[0] = (double)clamp_px(r);
    o.c[1] = (double)clamp_px(g);
    o.c[2] = (double)clamp_px(b);
    return o;
  }
  const double alpha = kScale * ca;
  o.c[0] = alpha * (double)clamp_px(r) + 0.0;
  o.c[1] = alpha * (double)clamp_px(g) + 0.0;
  o.c[2] = alpha * (double)clamp_px(b) + 0.0;
  return o;
}

// rd_cuda_common.cuh:451-476: plain ascending scan, last exact tie wins.
int nearest_linear(const OrcTree* t, const D4& tg) {
  double dist = 4.0 * ((kRange + 1.0) * (kRange + 1.0)) + 1.0;
  int best = 0;
  const double beta = kScale * tg.c[3];
  for (int i = 0; i < t->count; ++i) {
    const double* e = &t->cmap[4 * (size_t)i];
    const double alpha = t->assoc ? kScale * e[3] : 1.0;
    const double dr = alpha * e[0] + -beta * tg.c[0];
    const double dg = alpha * e[1] + -beta * tg.c[1];
    const double db = alpha * e[2] + -beta * tg.c[2];
    double d = dr * dr + (dg * dg + db * db);
    if (t->assoc) {
      const double da = e[3] + -tg.c[3];
      d = d + da * da;
    }
    if (d <= dist) {
      dist = d;
      best = i;
    }
  }
  return best;
}

// One sweep of JacobiStepKernel (src/rd_approx_cuda.cu:156-228) over every
// position: t = pal[q_in[i]] + sum_{d=0..min(taps,i)} g[d] (x[i-d] - pal[q_in[i-d]]),
// accumulated in T per channel with d ascending, widened to double, clamped
// and associated (rd_cuda_common.cuh:57-72), then looked up.  Positions are
// independent given q_in, so they are split over threads; the result does
// not depend on the split.  t_out (n x 4 doubles, or null) receives the
// unclamped target before association.
//
// `mutant` plants a known mistake for the tests that show the closed forms
// catch it (0 = the restatement): 1 updates q in place (Gauss-Seidel),
// 2 drops the last tap, 3 flips the residual's sign, 4 takes the residual
// one position late.
template <typename T>
void sweep_range(const OrcTree* t, const T* cx, const T* g, int taps, int use_tree,
                 const uint8_t* q_in, uint8_t* q_out, double* t_out, int64_t lo, int64_t hi,
                 int mutant) {
  const bool assoc = t->assoc;
  const uint8_t* qr = mutant == 1 ? q_out : q_in;
  for (int64_t i = lo; i < hi; ++i) {
    T acc[4] = {T(0), T(0), T(0), T(0)};
    int64_t dmax = taps < i ? taps : i;
    if (mutant == 2 && dmax == taps) dmax = taps - 1;
    for (int64_t d = 0; d <= dmax; ++d) {
      int64_t j = i - d;
      if (mutant == 4 && j > 0) j -= 1;
      const T* xs = &cx[4 * (size_t)j];
      const uint8_t qj = d == 0 ? q_in[(size_t)i] : qr[(size_t)j];
      const double* p = &t->cmap[4 * (size_t)qj];
      const T gk = g[(size_t)d];
      for (int c = 0; c < 4; ++c) {
        const T r = mutant == 3 ? ((T)p[c] - xs[c]) : (xs[c] - (T)p[c]);
        acc[c] += gk * r;
      }
    }
    const double* p0 = &t->cmap[4 * (size_t)q_in[(size_t)i]];
    const double vr = p0[0] + (double)acc[0];
    const double vg = p0[1] + (double)acc[1];
    const double vb = p0[2] + (double)acc[2];
    const double va = assoc ? p0[3] + (double)acc[3] : kRange;
    if (t_out) {
      t_out[4 * (size_t)i + 0] = vr;
      t_out[4 * (size_t)i + 1] = vg;
      t_out[4 * (size_t)i + 2] = vb;
      t_out[4 * (size_t)i + 3] = va;
    }
    const D4 tg = assoc_from_double(assoc, vr, vg, vb, va);
    q_out[(size_t)i] = (uint8_t)(use_tree ? t->lookup(tg) : nearest_linear(t, tg));
  }
}

template <typename T>
void sweep(const OrcTree* t, const T* cx, const T* g, int taps, int use_tree, const uint8_t* q_in,
           uint8_t* q_out, double* t_out, int64_t n, int threads, int mutant) {
  if (mutant == 1 || threads <= 1 || n < 65536) {
    sweep_range<T>(t, cx, g, taps, use_tree, q_in, q_out, t_out, 0, n, mutant);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t per = (n + threads - 1) / threads;
  for (int k = 0; k < threads; ++k) {
    const int64_t lo = k * per, hi = std::min(n, lo + per);
    if (lo >= hi) break;
    pool.emplace_back(sweep_range<T>, t, cx, g, taps, use_tree, q_in, q_out, t_out, lo, hi,
                      mutant);
  }
  for (auto& th : pool) th.join();
}

// InitIndexKernel (src/rd_approx_cuda.cu:127-148): the nearest entry of the
// raw pixel, narrowed to float and associated as the pixel path does.
template <typename T>
void init_index(const OrcTree* t, const T* cx, int use_tree, uint8_t* q, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    float v[4];
    for (int c = 0; c < 4; ++c) v[c] = (float)cx[4 * (size_t)i + c];
    const D4 tg = assoc_px(t->assoc, v);
    q[(size_t)i] = (uint8_t)(use_tree ? t->lookup(tg) : nearest_linear(t, tg));
  }
}

template <typename T>
std::vector<T> filter_as(int taps, double diffusion) {
  std::vector<double> gd((size_t)taps + 1);
  orc_inverse_filter(taps, diffusion, gd.data());
  std::vector<T> g((size_t)taps + 1);
  for (size_t i = 0; i < g.size(); ++i) g[i] = (T)gd[i];
  return g;
}

int jacobi_threads() {
  const unsigned hc = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(hc, 16u));
}

template <typename T>
int jacobi_impl(const OrcTree* t, float* rgba, int64_t w, int64_t h, int iterations,
                int taps, int use_tree, double diffusion, uint8_t* q_out) {
  if (t->count < 1 || t->count > 256) return -1;
  if (taps < 1 || taps > 4096) return -1;
  const bool assoc = t->assoc;
  std::vector<int32_t> curve((size_t)(w * h + 1)), owner((size_t)(w * h));
  const int64_t n = orc_block_order(w, h, curve.data(), owner.data());
  const std::vector<T> g = filter_as<T>(taps, diffusion);
  std::vector<T> cx((size_t)n * 4);
  for (int64_t i = 0; i < n; ++i)  // GatherKernelT
    for (int c = 0; c < 4; ++c) cx[4 * (size_t)i + c] = (T)rgba[4 * (size_t)curve[(size_t)i] + c];
  std::vector<uint8_t> q0((size_t)n), q1((size_t)n);
  init_index<T>(t, cx.data(), use_tree, q0.data(), n);
  const int threads = jacobi_threads();
  for (int it = 0; it < iterations; ++it) {  // the ping-pong of src/rd_approx_cuda.cu:356-368
    sweep<T>(t, cx.data(), g.data(), taps, use_tree, q0.data(), q1.data(), nullptr, n, threads, 0);
    std::swap(q0, q1);
  }
  for (int64_t i = 0; i < n; ++i) {  // ScatterKernel (owner-filtered)
    const int32_t pix = curve[(size_t)i];
    if (owner[(size_t)pix] != i) continue;
    const double* p = &t->cmap[4 * (size_t)q0[(size_t)i]];
    rgba[4 * pix + 0] = (float)p[0];
    rgba[4 * pix + 1] = (float)p[1];
    rgba[4 * pix + 2] = (float)p[2];
    rgba[4 * pix + 3] = assoc ? (float)p[3] : (float)65535.0;
  }
  if (q_out) std::memcpy(q_out, q0.data(), (size_t)n);
  (void)assoc;
  return 0;
}

}  // namespace

extern "C" int orc_jacobi(const OrcTree* t, float* rgba, int64_t w, int64_t h,
                          int iterations, int taps, int fp64, int use_tree,
                          double diffusion, uint8_t* q_out) {
  if (fp64) return jacobi_impl<double>(t, rgba, w, h, iterations, taps, use_tree, diffusion, q_out);
  return jacobi_impl<float>(t, rgba, w, h, iterations, taps, use_tree, diffusion, q_out);
}

extern "C" int orc_jacobi_init(const OrcTree* t, const void* cx, int64_t n, int fp64,
                               int use_tree, uint8_t* q) {
  if (t->count < 1 || t->count > 256 || n < 0) return -1;
  if (fp64)
    init_index<double>(t, (const double*)cx, use_tree, q, n);
  else
    init_index<float>(t, (const float*)cx, use_tree, q, n);
  return 0;
}

extern "C" int orc_jacobi_sweep(const OrcTree* t, const void* cx, int64_t n, const uint8_t* q_in,
                                uint8_t* q_out, double* t_out, int taps, int fp64, int use_tree,
                                double diffusion, int mutant) {
  if (t->count < 1 || t->count > 256 || n < 0) return -1;
  if (taps < 1 || taps > 4096) return -1;
  if (mutant == 1) std::memcpy(q_out, q_in, (size_t)n);
  const int threads = jacobi_threads();
  if (fp64) {
    const std::vector<double> g = filter_as<double>(taps, diffusion);
    sweep<double>(t, (const double*)cx, g.data(), taps, use_tree, q_in, q_out, t_out, n, threads,
                  mutant);
  } else {
    const std::vector<float> g = filter_as<float>(taps, diffusion);
    sweep<float>(t, (const float*)cx, g.data(), taps, use_tree, q_in, q_out, t_out, n, threads,
                 mutant);
  }
  return 0;
}

// ===========================================================================
// Video palette montage.  src/rd_video.cpp:2373-2380 (grid), :2404-2418
// (widen), :2441-2470 (place_raw), :2636-2660 (cells past the last sample
// repeat it), restated on the CPU.  Frames are the decoder's bytes: fmt 1
// rgb24, 2 rgba, 3 rgb48, 4 rgba64le.  out: mw*mh RgbaF.
// ===========================================================================
namespace {
void widen_px(int fmt, const void* frame, int64_t o, float* d) {
  if (fmt == 1 || fmt == 2) {
    const int ch = fmt == 1 ? 3 : 4;
    const uint8_t* p = (const uint8_t*)frame + ch * o;
    d[0] = (float)p[0] * 257.0f;
    d[1] = (float)p[1] * 257.0f;
    d[2] = (float)p[2] * 257.0f;
    d[3] = ch == 4 ? (float)p[3] * 257.0f : 65535.0f;
  } else {
    const int ch = fmt == 3 ? 3 : 4;
    const uint16_t* p = (const uint16_t*)frame + ch * o;
    d[0] = (float)p[0];
    d[1] = (float)p[1];
    d[2] = (float)p[2];
    d[3] = ch == 4 ? (float)p[3] : 65535.0f;
  }
}
}  // namespace

extern "C" int orc_video_montage(const void* const* frames, int sampled, int fmt, int64_t w,
                                 int64_t h, int want, int tile, int cols, float* out,
                                 int64_t* mw_out, int64_t* mh_out) {
  if (sampled < 1 || want < sampled || w < 1 || h < 1) return -1;
  const int64_t cell_w = tile > 0 ? tile : w, cell_h = tile > 0 ? tile : h;
  const int c = cols > 0 ? cols : (int)std::ceil(std::sqrt((double)want));
  const int rows = (want + c - 1) / c;
  const int64_t mw = (int64_t)c * cell_w, mh = (int64_t)rows * cell_h;
  if (mw_out) *mw_out = mw;
  if (mh_out) *mh_out = mh;
  if (!out) return 0;
  for (int s = 0; s < sampled; ++s) {  // place_raw
    const int64_t cx = s % c, cy = s / c;
    for (int64_t ty = 0; ty < cell_h; ++ty) {
      const int64_t sy = tile > 0 ? ty * h / cell_h : ty;
      float* row = out + 4 * ((cy * cell_h + ty) * mw + cx * cell_w);
      for (int64_t tx = 0; tx < cell_w; ++tx) {
        const int64_t sx = tile > 0 ? tx * w / cell_w : tx;
        widen_px(fmt, frames[s], sy * w + sx, row + 4 * tx);
      }
    }
  }
  const int64_t lx = (sampled - 1) % c, ly = (sampled - 1) / c;
  for (int s = sampled; s < c * rows; ++s) {  // repeat the last sample
    const int64_t cx = s % c, cy = s / c;
    for (int64_t ty = 0; ty < cell_h; ++ty)
      std::memcpy(out + 4 * ((cy * cell_h + ty) * mw + cx * cell_w),
                  out + 4 * ((ly * cell_h + ty) * mw + lx * cell_w), 16 * (size_t)cell_w);
  }
  return 0;
}

// ===========================================================================
// Error-diffusion plugin dithers, restated from src/rd_dithers.cc:74-220 and
// src/rd_dither_common.h (NearestEntry, ErrorPlane, Accumulate, WriteChosen):
// kind 0 floyd-steinberg, 1 atkinson, 2 jarvis, in place on RgbaF frames.
// ===========================================================================
namespace {
int nearest_entry_f(const double* ent, int count, float r, float g, float b) {
  int best = 0;
  float best_d = -1.0f;
  for (int i = 0; i < count; ++i) {
    const float dr = r - (float)ent[4 * i];
    const float dg = g - (float)ent[4 * i + 1];
    const float db = b - (float)ent[4 * i + 2];
    const float d = dr * dr + dg * dg + db * db;
    if (best_d < 0.0f || d < best_d) {
      best_d = d;
      best = i;
    }
  }
  return best;
}

struct Plane {
  std::vector<float> r, g, b;
  size_t n = 0;
  void add(size_t i, float w, float er, float eg, float eb) {
    if (i >= n) return;
    r[i] += w * er;
    g[i] += w * eg;
    b[i] += w * eb;
  }
};
}  // namespace

extern "C" int orc_dither_diffuse(const double* ent, int count, float* rgba, int64_t w64,
                                  int64_t h64, int frames, int kind, double diffusion) {
  if (count <= 0 || w64 <= 0 || h64 <= 0 || kind < 0 || kind > 2) return -1;
  const size_t w = (size_t)w64, h = (size_t)h64, n = w * h;
  const float d = (float)diffusion;
  Plane e;
  e.n = n;
  for (int f = 0; f < frames; ++f) {
    e.r.assign(n, 0.f);
    e.g.assign(n, 0.f);
    e.b.assign(n, 0.f);
    float* fr = rgba + 4 * n * (size_t)f;
    for (size_t y = 0; y < h; ++y)
      for (size_t x = 0; x < w; ++x) {
        const size_t i = y * w + x;
        float* p = fr + 4 * i;
        const float r = p[0] + e.r[i] * d;
        const float g = p[1] + e.g[i] * d;
        const float b = p[2] + e.b[i] * d;
        const int k = nearest_entry_f(ent, count, r, g, b);
        const double* c = ent + 4 * k;
        const float er = r - (float)c[0], eg = g - (float)c[1], eb = b - (float)c[2];
        if (kind == 0) {  // :98-105
          if (x + 1 < w) e.add(i + 1, 7.0f / 16.0f, er, eg, eb);
          e.add(i + w - (x > 0 ? 1 : 0), 3.0f / 16.0f, er, eg, eb);
          e.add(i + w, 5.0f / 16.0f, er, eg, eb);
          if (x + 1 < w) e.add(i + w + 1, 1.0f / 16.0f, er, eg, eb);
        } else if (kind == 1) {  // :134-141
          const float q = 1.0f / 8.0f;
          if (x + 1 < w) e.add(i + 1, q, er, eg, eb);
          if (x + 2 < w) e.add(i + 2, q, er, eg, eb);
          if (x > 0) e.add(i + w - 1, q, er, eg, eb);
          e.add(i + w, q, er, eg, eb);
          if (x + 1 < w) e.add(i + w + 1, q, er, eg, eb);
          if (y + 1 < h) e.add(i + 2 * w, q, er, eg, eb);
        } else {  // :198-212: columns in size_t arithmetic, clamped only at the borders
          static const int tdy[12] = {2, 2, 2, 2, 1, 1, 1, 1, 0, 0, 0, 0};
          static const int tdx[12] = {-2, -1, 0, 1, -2, -1, 0, 1, -2, -1, 1, 2};
          static const float tw[12] = {1.0f / 48.0f, 3.0f / 48.0f, 5.0f / 48.0f, 3.0f / 48.0f,
                                       7.0f / 48.0f, 5.0f / 48.0f, 7.0f / 48.0f, 5.0f / 48.0f,
                                       3.0f / 48.0f, 1.0f / 48.0f, 3.0f / 48.0f, 1.0f / 48.0f};
          for (int t = 0; t < 12; ++t) {
            size_t tx = x + (size_t)(ptrdiff_t)tdx[t];
            size_t ty = y + (size_t)(ptrdiff_t)tdy[t];
            if (tdx[t] < 0 && x == 0) tx = 0;
            if (tdx[t] > 0 && x + 1 >= w) tx = w - 1;
            if (tdy[t] > 0 && y + 1 >= h) ty = h - 1;
            e.add(ty * w + tx, tw[t], er, eg, eb);
          }
        }
        p[0] = (float)c[0];  // WriteChosen: r, g, b; alpha untouched
        p[1] = (float)c[1];
        p[2] = (float)c[2];
      }
  }
  return 0;
}
