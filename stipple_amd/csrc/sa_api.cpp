// sa_api.cpp -- the C-ABI (include/stipple_amd.h) over the HIP kernels.
//
// Host-side responsibilities only: argument checks with the reference's
// refusals, the palette search image, the per-geometry curve cache, the error
// weights (computed with the host libm, exactly as the reference computes
// them, never on the device), stream-ordered scratch, and launches.
#include "../../include/stipple_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sa_internal.h"
#include "sa_math.h"
#include "sa_octree.h"
#include "sa_palette.h"

struct sa_palette_s {
  sa::PaletteDev hdr;        // host copy of the header
  sa::PaletteDev* dev_hdr = nullptr;
  void* dev_image = nullptr;
  int count = 0;
  int assoc = 0;
  // Set when the palette was built on the device.
  double* dev_entries = nullptr;
  int depth = 0;
  int64_t nodes = 0;
  int exact_root = 0;
  bool has_tree = false;
  sa::TreeDev tree{};
  void* dev_block = nullptr;   // built palettes: one allocation for image, hdr, entries
  void* tree_block = nullptr;  // built with keep_tree: the node arrays
};

namespace {

int bind_noalloc_scratch(int device, char* err, size_t errlen);  // sa_walk_blocks_noalloc

// ---- tuning knobs (set by an autotuner; none of them changes a result) ------
struct Knobs {
  std::atomic<int> blocks_threads{256};  // block walk: blocks per workgroup
  std::atomic<int> approx_per{4};        // Jacobi sweep: outputs per thread
  std::atomic<int> seq_memo_bits{0};     // sequential walk: memo slots (log2); 0 = largest fit
};
Knobs g_knobs;


using sa::PaletteDev;

std::mutex g_mu;
int g_device = -1;
uint32_t* g_stats = nullptr;  // walk counters: misses, line fills, line loads from HBM
int g_cus = 256;
size_t g_lds_per_cu = 160 * 1024;
char g_name[256] = "";

int fail(char* err, size_t n, int code, const char* fmt, ...) {
  if (err && n) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, n, fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIP_TRY(expr)                                                                \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(err, errlen, SA_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

int ensure_init(char* err, size_t errlen) {
  std::lock_guard<std::mutex> g(g_mu);
  if (g_device >= 0) return SA_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(err, errlen, SA_EDEVICE, "no HIP device available");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, dev));
  g_cus = prop.multiProcessorCount;
  if (!g_stats) {
    HIP_TRY(hipMalloc((void**)&g_stats, 16));
    HIP_TRY(hipMemset(g_stats, 0, 16));
  }
  snprintf(g_name, sizeof(g_name), "%s (%s, %d CUs)", prop.name, prop.gcnArchName, g_cus);
  // Every build and walk allocates its scratch stream-ordered; keep the freed
  // memory in the pool instead of returning it to the HIP runtime at each
  // synchronisation (the default threshold, 0, makes every palette build map
  // its ~100 MB of pair arrays afresh).
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool) {
    uint64_t keep = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
  }
  g_device = dev;
  return SA_OK;
}

// Scaled weights (1/16 * diffusion) * 16^(-k/15), via the host libm: the
// reference computes them on the host and uploads them (include/rd_types.h:
// 169-176, src/rd_riemersma_cpu.cpp:370-373, include/rd_cuda_common.cuh:4-10).
void scaled_weights(double diffusion, double sw[16]) {
  double weight = 1.0;
  for (int i = 0; i < 16; ++i) {
    const double w = sa::safe_recip(weight);
    sw[i] = (1.0 / 16.0) * diffusion * w;
    weight *= std::exp(std::log(1.0 / (1.0 / 16.0)) / (16 - 1.0));
  }
}

// ---- pinned staging for small host->device argument arrays ----------------
struct Staging {
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;
};
std::mutex g_stage_mu;
std::vector<Staging> g_stage;

// Copies `bytes` from `src` to device `dst` on `s` through a pinned buffer
// that is not reused until the copy has executed.
hipError_t staged_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
  std::lock_guard<std::mutex> g(g_stage_mu);
  Staging* pick = nullptr;
  for (Staging& st : g_stage) {
    if (st.cap >= bytes && hipEventQuery(st.done) == hipSuccess) {
      pick = &st;
      break;
    }
  }
  if (!pick) {
    Staging st;
    st.cap = std::max<size_t>(bytes, 64 * 1024);
    hipError_t e = hipHostMalloc(&st.host, st.cap, hipHostMallocDefault);
    if (e != hipSuccess) return e;
    e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming);
    if (e != hipSuccess) return e;
    g_stage.push_back(st);
    pick = &g_stage.back();
  }
  std::memcpy(pick->host, src, bytes);
  hipError_t e = hipMemcpyAsync(dst, pick->host, bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipEventRecord(pick->done, s);
}

// ---- curve cache ---------------------------------------------------------
struct CurveEntry {
  int64_t w, h;
  int32_t* order;  // w*h
  size_t bytes;
  hipEvent_t ready;  // recorded after the build on the building stream
};
std::mutex g_curve_mu;
std::list<CurveEntry> g_curves;  // most recently used first
size_t g_curve_bytes = 0;
const size_t kCurveBudget = (size_t)8 << 30;

int curve_level(int64_t w, int64_t h) {  // include/rd_riemersma.h:88-94, on the host
  const int64_t extent = w > h ? w : h;
  if (extent <= 0) return 0;
  int level = (int)std::log2((double)extent);
  if (((int64_t)1 << level) < extent) ++level;
  return level;
}

int build_curve(int64_t w, int64_t h, int32_t* order, hipStream_t s, char* err, size_t errlen) {
  const int level = curve_level(w, h);
  if (level > 15) return fail(err, errlen, SA_EUNSUPPORTED, "image side above 32768");
  const int64_t chunks = sa::curve_chunks(level);
  int32_t* counts = nullptr;
  HIP_TRY(sa::scratch_alloc((void**)&counts, sizeof(int32_t) * chunks, s));
  HIP_TRY(sa::launch_curve_count(level, w, h, counts, chunks, s));
  HIP_TRY(sa::launch_curve_scan(counts, chunks, s));
  HIP_TRY(sa::launch_curve_emit(level, w, h, counts, chunks, order, s));
  HIP_TRY(sa::scratch_free(counts, s));
  return SA_OK;
}

// Returns a cached device curve for (w, h), building it on `s` if needed.
int get_curve(int64_t w, int64_t h, hipStream_t s, const int32_t** out, char* err,
              size_t errlen) {
  std::lock_guard<std::mutex> g(g_curve_mu);
  for (auto it = g_curves.begin(); it != g_curves.end(); ++it) {
    if (it->w == w && it->h == h) {
      g_curves.splice(g_curves.begin(), g_curves, it);
      // Built on another stream, possibly still in flight: order after it.
      HIP_TRY(hipStreamWaitEvent(s, g_curves.front().ready, 0));
      *out = g_curves.front().order;
      return SA_OK;
    }
  }
  CurveEntry e;
  e.w = w;
  e.h = h;
  e.bytes = sizeof(int32_t) * (size_t)(w * h);
  // Evict least recently used entries beyond the budget.  The device must be
  // done with them: synchronise before freeing.
  if (g_curve_bytes + e.bytes > kCurveBudget && !g_curves.empty()) {
    HIP_TRY(hipDeviceSynchronize());
    while (g_curve_bytes + e.bytes > kCurveBudget && !g_curves.empty()) {
      HIP_TRY(hipFree(g_curves.back().order));
      HIP_TRY(hipEventDestroy(g_curves.back().ready));
      g_curve_bytes -= g_curves.back().bytes;
      g_curves.pop_back();
    }
  }
  HIP_TRY(hipMalloc((void**)&e.order, e.bytes));
  const int rc = build_curve(w, h, e.order, s, err, errlen);
  if (rc != SA_OK) {
    (void)hipFree(e.order);
    return rc;
  }
  HIP_TRY(hipEventCreateWithFlags(&e.ready, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(e.ready, s));
  g_curves.push_front(e);
  g_curve_bytes += e.bytes;
  *out = e.order;
  return SA_OK;
}

// ---- weights cache (per diffusion value) ---------------------------------
std::mutex g_w_mu;
std::map<double, double*> g_weights;

int get_weights(double diffusion, hipStream_t s, const double** out, char* err, size_t errlen) {
  std::lock_guard<std::mutex> g(g_w_mu);
  auto it = g_weights.find(diffusion);
  if (it != g_weights.end()) {
    *out = it->second;
    return SA_OK;
  }
  double sw[16];
  scaled_weights(diffusion, sw);
  double* d = nullptr;
  HIP_TRY(hipMalloc((void**)&d, sizeof(sw)));
  HIP_TRY(hipMemcpy(d, sw, sizeof(sw), hipMemcpyHostToDevice));
  (void)s;
  g_weights[diffusion] = d;
  *out = d;
  return SA_OK;
}

// ---- kernel timing (the RD_KERNEL_TIMING of src/rd_blocks_cuda.cu:820-837) --
struct Timed {
  const char* name;
  hipEvent_t a, b;
};
std::mutex g_t_mu;
bool g_timing = false;
std::vector<Timed> g_timed;

// Brackets one launch with events on its stream when timing is enabled.
template <typename Fn>
hipError_t timed(const char* name, hipStream_t s, Fn&& launch) {
  if (!g_timing) return launch();
  Timed t{name, nullptr, nullptr};
  hipError_t e = hipEventCreate(&t.a);
  if (e == hipSuccess) e = hipEventCreate(&t.b);
  if (e == hipSuccess) e = hipEventRecord(t.a, s);
  if (e != hipSuccess) return e;
  e = launch();
  if (e != hipSuccess) return e;
  e = hipEventRecord(t.b, s);
  std::lock_guard<std::mutex> g(g_t_mu);
  g_timed.push_back(t);
  return e;
}

// Staged inputs of one image: its pixels in visit order at the source's
// width, and 16 bytes of padding (the walk loads 16 bytes from each start).
size_t stage_bytes(int fmt, int64_t px) {
  return (((size_t)px << sa::seq_stage_shift(fmt)) + 16 + 255) & ~(size_t)255;
}

// The walk's LDS memo cache for one image: entries of the fewest bytes that
// hold index+1 (1 up to 255 colours, 2 up to 65535, 4 for 65536), and as many
// line slots as the LDS left
// beside the search image and the input ring holds (a multiple of 4, at most
// 4096); `memo_bits` > 0 (or the seq.memo_bits knob, as a cap) asks for
// 1 << memo_bits slots instead.
struct MemoShape {
  int slots = 0;
  int eb = 2;         // bytes per entry
  size_t bytes = 0;   // LDS (and snapshot) bytes
  int key_bits = 0;   // the HBM table holds 1 << key_bits entries
  size_t table = 0;   // its bytes
};

MemoShape memo_shape(const sa_seq_options& o, int assoc, int img_bytes, int count);

bool valid_in_fmt(int f) { return f >= sa::kRgbaF32 && f <= sa::kRgba16; }
bool valid_out_fmt(int f) { return f >= sa::kRgbaF32 && f <= sa::kIndex16; }
// The block engine and the palette build also read planar YUV 4:4:4 / 4:2:0;
// the block engine writes 4:4:4.
bool valid_blocks_in_fmt(int f) { return valid_in_fmt(f) || f == sa::kYuv444 || f == sa::kYuv420; }
bool valid_blocks_out_fmt(int f) { return valid_out_fmt(f) || f == sa::kYuv444; }

}  // namespace

// ===========================================================================
extern "C" int sa_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0 ? 1 : 0;
}

extern "C" int sa_init(int device, char* err, size_t errlen) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(err, errlen, SA_EDEVICE, "no HIP device available");
  if (device < 0 || device >= n) return fail(err, errlen, SA_EINVAL, "device %d of %d", device, n);
  HIP_TRY(hipSetDevice(device));
  {
    std::lock_guard<std::mutex> g(g_mu);
    g_device = -1;
  }
  const int rc = ensure_init(err, errlen);
  if (rc) return rc;
  return bind_noalloc_scratch(device, err, errlen);
}

extern "C" int sa_device_name(char* out, size_t len) {
  char err[256];
  if (ensure_init(err, sizeof(err)) != SA_OK) return SA_EDEVICE;
  snprintf(out, len, "%s", g_name);
  return SA_OK;
}

// ---- palette ----------------------------------------------------------------
extern "C" int sa_palette_create(const double* entries, int count, int assoc,
                                 const int32_t* child16, const int32_t* parent,
                                 const uint32_t* color_number, const uint64_t* number_unique,
                                 int64_t nodes, sa_palette* out, char* err, size_t errlen) {
  (void)parent;
  if (!out) return fail(err, errlen, SA_EINVAL, "null output handle");
  *out = nullptr;
  if (count < 1) return fail(err, errlen, SA_EREFUSED, "empty palette");
  if (count > 65536) return fail(err, errlen, SA_EINVAL, "palette above 65536 entries");
  if (nodes < 1 || !child16 || !color_number || !number_unique || !entries)
    return fail(err, errlen, SA_EINVAL, "missing tree arrays");
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  const int kids = assoc ? 16 : 8;
  // Reachable nodes from the root, pre-order; internal ones get rows.
  std::vector<int32_t> row_of((size_t)nodes, -1);
  std::vector<int32_t> rows;  // node id per row
  std::vector<int32_t> stack{0};
  std::vector<char> seen((size_t)nodes, 0);
  while (!stack.empty()) {
    const int32_t n = stack.back();
    stack.pop_back();
    if (n < 0 || n >= nodes || seen[(size_t)n]) return fail(err, errlen, SA_EINVAL, "malformed tree");
    seen[(size_t)n] = 1;
    bool internal = false;
    for (int c = 0; c < kids; ++c) internal |= child16[16 * n + c] >= 0;
    if (internal || n == 0) {
      row_of[(size_t)n] = (int32_t)rows.size();
      rows.push_back(n);
    }
    for (int c = kids - 1; c >= 0; --c)
      if (child16[16 * n + c] >= 0) stack.push_back(child16[16 * n + c]);
  }
  if (rows.size() > 32767) return fail(err, errlen, SA_EUNSUPPORTED, "tree above 32767 internal nodes");
  // Post-order candidate list and the range of every internal node's subtree.
  std::vector<uint16_t> cand;
  std::vector<int32_t> sub_off(rows.size(), 0), sub_len(rows.size(), 0);
  struct Fr {
    int32_t n;
    int next;
    int32_t first;
  };
  std::vector<Fr> st{{0, 0, 0}};
  while (!st.empty()) {
    Fr& f = st.back();
    if (f.next < kids) {
      const int32_t c = child16[16 * f.n + f.next];
      ++f.next;
      if (c >= 0) st.push_back({c, 0, (int32_t)cand.size()});
      continue;
    }
    if (number_unique[f.n] != 0) {
      const uint32_t cn = color_number[f.n];
      if ((int)cn >= count) return fail(err, errlen, SA_EINVAL, "colour number out of range");
      cand.push_back((uint16_t)cn);
    }
    const int32_t r = row_of[(size_t)f.n];
    if (r >= 0) {
      sub_off[(size_t)r] = f.first;
      sub_len[(size_t)r] = (int32_t)cand.size() - f.first;
    }
    st.pop_back();
  }
  // Search image.
  const sa::PaletteLayout L = sa::palette_layout(count, (int)rows.size(), kids);
  std::vector<unsigned char> img((size_t)L.bytes, 0);
  double* col = (double*)(img.data() + L.off_col);
  float* ent = (float*)(img.data() + L.off_ent);
  float* cent = (float*)(img.data() + L.off_cent);
  int32_t* sub = (int32_t*)(img.data() + L.off_sub);
  uint32_t* desc = (uint32_t*)(img.data() + L.off_desc);
  int16_t* child = (int16_t*)(img.data() + L.off_child);
  uint16_t* cnd = (uint16_t*)(img.data() + L.off_cand);
  for (int i = 0; i < count; ++i) {
    const double* e = entries + 4 * i;
    for (int c = 0; c < 4; ++c) {
      if ((double)(float)e[c] != e[c])
        return fail(err, errlen, SA_EUNSUPPORTED, "palette entry %d is not a Quantum (float) value", i);
      ent[4 * i + c] = (float)e[c];
    }
    sa::assoc_entry(assoc != 0, e, col + 4 * i);
  }
  for (size_t r = 0; r < rows.size(); ++r) {
    sub[2 * r] = sub_off[r];
    sub[2 * r + 1] = sub_len[r];
    const int32_t n = rows[r];
    for (int c = 0; c < kids; ++c) {
      const int32_t ch = child16[16 * n + c];
      int16_t v = sa::kNoChild;
      if (ch >= 0) v = row_of[(size_t)ch] >= 0 ? (int16_t)row_of[(size_t)ch] : sa::kLeaf;
      child[r * kids + c] = v;
    }
  }
  for (size_t i = 0; i < cand.size() && i < (size_t)count; ++i) {
    cnd[i] = cand[i];
    for (int c = 0; c < 4; ++c) cent[4 * i + c] = ent[4 * cand[i] + c];
  }
  for (int e = 0; e < sa::desc_entries(kids); ++e) desc[e] = sa::descent3(child, kids, (uint32_t)e);
  sa_palette p = new sa_palette_s();
  p->count = count;
  p->assoc = assoc ? 1 : 0;
  PaletteDev& h = p->hdr;
  h.count = count;
  h.assoc = p->assoc;
  h.kids = kids;
  h.rows = (int32_t)rows.size();
  h.bytes = L.bytes;
  h.off_ent = L.off_ent;
  h.off_cent = L.off_cent;
  h.off_sub = L.off_sub;
  h.off_desc = L.off_desc;
  h.off_child = L.off_child;
  h.off_cand = L.off_cand;
  h.pad_ = 0;
  hipError_t e = hipMalloc(&p->dev_image, (size_t)L.bytes);
  if (e == hipSuccess) e = hipMemcpy(p->dev_image, img.data(), (size_t)L.bytes, hipMemcpyHostToDevice);
  h.image = p->dev_image;
  if (e == hipSuccess) e = hipMalloc((void**)&p->dev_hdr, sizeof(PaletteDev));
  if (e == hipSuccess) e = hipMemcpy(p->dev_hdr, &h, sizeof(PaletteDev), hipMemcpyHostToDevice);
  // The entries as given (doubles), for the screen dithers and sa_palette_entries.
  if (e == hipSuccess) e = hipMalloc((void**)&p->dev_entries, sizeof(double) * 4 * (size_t)count);
  if (e == hipSuccess)
    e = hipMemcpy(p->dev_entries, entries, sizeof(double) * 4 * (size_t)count,
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    sa_palette_destroy(p);
    return fail(err, errlen, SA_EDEVICE, "palette upload: %s", hipGetErrorString(e));
  }
  *out = p;
  return SA_OK;
}

namespace {
// Destroyed palettes' device blocks.  Work queued before a destroy may still
// read the palette, so its blocks cannot go back at once; a device
// synchronisation per destroy, though, stalls whichever thread drops a
// palette (a serving loop's palette builder waited a whole walk step per
// palette).  So the blocks wait here and go back after one synchronisation
// once they hold more than kDeferBytes; no work queued after the destroy may
// use the palette (the contract), so that synchronisation covers every use.
std::mutex g_defer_mu;
std::vector<void*> g_defer;
size_t g_defer_bytes = 0;
constexpr size_t kDeferBytes = (size_t)256 << 20;

void defer_free(void* q, size_t bytes) {
  if (!q) return;
  std::vector<void*> drop;
  {
    std::lock_guard<std::mutex> g(g_defer_mu);
    g_defer.push_back(q);
    g_defer_bytes += bytes;
    if (g_defer_bytes <= kDeferBytes) return;
    drop.swap(g_defer);
    g_defer_bytes = 0;
  }
  (void)hipDeviceSynchronize();
  for (void* d : drop) (void)hipFree(d);
}
}  // namespace

extern "C" int sa_palette_destroy(sa_palette p) {
  if (!p) return SA_OK;
  // Sizes for the deferred list's budget (estimates: the blocks' exact sizes
  // are not kept).
  const size_t img = (size_t)p->hdr.bytes + 256 + sizeof(double) * 4 * (size_t)p->count;
  if (p->dev_block) {
    defer_free(p->dev_block, img);
  } else {
    defer_free(p->dev_image, (size_t)p->hdr.bytes);
    defer_free(p->dev_hdr, 256);
    defer_free(p->dev_entries, sizeof(double) * 4 * (size_t)p->count);
  }
  if (p->tree_block) defer_free(p->tree_block, (size_t)std::max<int64_t>(p->nodes, 1) * 128);
  delete p;
  return SA_OK;
}

namespace {
// A palette around a device build's outputs.  The header is the one k_reduce
// wrote (the same layout function gives the same offsets), so it needs no
// read back.
sa_palette wrap_built(const sa::OctreeOut& o, const sa::OctreeResult& r, int assoc) {
  sa_palette p = new sa_palette_s();
  p->dev_block = o.block;
  p->dev_image = o.image;
  p->dev_hdr = o.hdr;
  p->dev_entries = o.entries;
  p->count = r.count;
  p->assoc = assoc ? 1 : 0;
  p->depth = r.depth;
  p->nodes = r.nodes;
  p->exact_root = r.exact_root;
  p->has_tree = o.keep_tree != 0;
  p->tree = o.tree;
  p->tree_block = o.tree_block;
  const int kids = assoc ? 16 : 8;
  const sa::PaletteLayout L = sa::palette_layout(r.count, r.rows, kids);
  PaletteDev& h = p->hdr;
  h.count = r.count;
  h.assoc = p->assoc;
  h.kids = kids;
  h.rows = r.rows;
  h.bytes = L.bytes;
  h.off_ent = L.off_ent;
  h.off_cent = L.off_cent;
  h.off_sub = L.off_sub;
  h.off_desc = L.off_desc;
  h.off_child = L.off_child;
  h.off_cand = L.off_cand;
  h.pad_ = 0;
  h.image = o.image;
  return p;
}
}  // namespace

extern "C" int sa_palette_build(const void* src, int src_fmt, int64_t w, int64_t h, int colors,
                                int assoc, int gray, int keep_tree, void* stream, sa_palette* out,
                                char* err, size_t errlen) {
  if (!out) return fail(err, errlen, SA_EINVAL, "null output handle");
  *out = nullptr;
  if (!src || w < 1 || h < 1) return fail(err, errlen, SA_EINVAL, "bad image");
  if (!valid_blocks_in_fmt(src_fmt)) return fail(err, errlen, SA_EINVAL, "bad pixel format");
  if (src_fmt == sa::kYuv420 && ((w & 1) || (h & 1)))
    return fail(err, errlen, SA_EUNSUPPORTED, "4:2:0 input needs even width and height");
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  sa::OctreeOut o;
  o.keep_tree = keep_tree ? 1 : 0;
  const sa::OctreeResult r = sa::build_octree(src, src_fmt, w, h, colors, assoc, gray, s, &o);
  if (r.status != 0) return fail(err, errlen, r.status, "%s", r.msg);
  *out = wrap_built(o, r, assoc);
  return SA_OK;
}

// ---- video palette montage ----------------------------------------------------------
namespace {
// The montage geometry of src/rd_video.cpp:2373-2380 (cols: ceil(sqrt(want))
// unless given).
int montage_shape(int64_t w, int64_t h, int want, int tile, int cols, int64_t* cell_w,
                  int64_t* cell_h, int* cols_out, int64_t* mw, int64_t* mh) {
  if (w < 1 || h < 1 || want < 1) return SA_EINVAL;
  *cell_w = tile > 0 ? tile : w;
  *cell_h = tile > 0 ? tile : h;
  int c = cols > 0 ? cols : (int)std::ceil(std::sqrt((double)want));
  if (c < 1) c = 1;
  const int rows = (want + c - 1) / c;
  *cols_out = c;
  *mw = (int64_t)c * *cell_w;
  *mh = (int64_t)rows * *cell_h;
  return SA_OK;
}
}  // namespace

extern "C" int sa_video_montage(const void* const* frames, int n, int src_fmt, int64_t width,
                                int64_t height, int want, int tile, int cols, void* montage,
                                int64_t* mw, int64_t* mh, int* gray, void* stream, char* err,
                                size_t errlen) {
  int64_t cw = 0, ch = 0, W = 0, H = 0;
  int c = 0;
  if (montage_shape(width, height, want, tile, cols, &cw, &ch, &c, &W, &H) != SA_OK)
    return fail(err, errlen, SA_EINVAL, "bad montage geometry (%lldx%lld, want %d)",
                (long long)width, (long long)height, want);
  if (mw) *mw = W;
  if (mh) *mh = H;
  if (!montage) return SA_OK;  // a size query
  if (n < 1 || n > want || !frames)
    return fail(err, errlen, SA_EINVAL, "montage: %d sampled frames for %d cells", n, want);
  if (src_fmt != sa::kRgb8 && src_fmt != sa::kRgba8 && src_fmt != sa::kRgb16 &&
      src_fmt != sa::kRgba16)
    return fail(err, errlen, SA_EINVAL, "montage: frames must be RGB8, RGBA8, RGB16 or RGBA16");
  if (W * H > ((int64_t)1 << 28))
    return fail(err, errlen, SA_EUNSUPPORTED, "montage of %lld pixels: above 2^28",
                (long long)(W * H));
  for (int i = 0; i < n; ++i)
    if (!frames[i]) return fail(err, errlen, SA_EINVAL, "montage: frame %d is null", i);
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  unsigned char* d = nullptr;
  const size_t ptr_bytes = (sizeof(void*) * (size_t)n + 255) & ~(size_t)255;
  HIP_TRY(sa::scratch_alloc((void**)&d, ptr_bytes + 256, s));
  HIP_TRY(staged_copy(d, frames, sizeof(void*) * (size_t)n, s));
  int* d_gray = (int*)(d + ptr_bytes);
  HIP_TRY(sa::launch_video_montage((const void* const*)d, n, src_fmt, width, height, cw, ch, c, W,
                                   H, montage, d_gray, s));
  if (gray) {
    int* h_gray = nullptr;
    HIP_TRY(hipHostMalloc((void**)&h_gray, sizeof(int), hipHostMallocDefault));
    hipError_t e = hipMemcpyAsync(h_gray, d_gray, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    *gray = *h_gray;
    (void)hipHostFree(h_gray);
    if (e != hipSuccess) {
      (void)sa::scratch_free(d, s);
      return fail(err, errlen, SA_EDEVICE, "montage: %s", hipGetErrorString(e));
    }
  }
  HIP_TRY(sa::scratch_free(d, s));
  return SA_OK;
}

extern "C" int sa_video_palette(const void* const* frames, int n, int src_fmt, int64_t width,
                                int64_t height, int want, int tile, int cols, int colors,
                                int keep_tree, void* stream, sa_palette* out, char* err,
                                size_t errlen) {
  if (!out) return fail(err, errlen, SA_EINVAL, "null output handle");
  *out = nullptr;
  int64_t W = 0, H = 0;
  int rc = sa_video_montage(frames, n, src_fmt, width, height, want, tile, cols, nullptr, &W, &H,
                            nullptr, stream, err, errlen);
  if (rc) return rc;
  rc = ensure_init(err, errlen);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  void* m = nullptr;
  HIP_TRY(sa::scratch_alloc(&m, (size_t)(8 * W * H), s));
  int gray = 0;
  rc = sa_video_montage(frames, n, src_fmt, width, height, want, tile, cols, m, &W, &H, &gray,
                        stream, err, errlen);
  if (rc == SA_OK)
    rc = sa_palette_build(m, sa::kRgba16, W, H, colors, 0, gray, keep_tree, stream, out, err,
                          errlen);
  HIP_TRY(sa::scratch_free(m, s));
  return rc;
}

extern "C" int sa_palette_build_batch(const sa_build* items, int n, void* stream, sa_palette* out,
                                      int* status, char* err, size_t errlen) {
  if (n < 0 || (n > 0 && (!items || !out))) return fail(err, errlen, SA_EINVAL, "bad build list");
  for (int i = 0; i < n; ++i) out[i] = nullptr;
  if (n == 0) return SA_OK;
  for (int i = 0; i < n; ++i) {
    const sa_build& b = items[i];
    if (!b.src || b.width < 1 || b.height < 1)
      return fail(err, errlen, SA_EINVAL, "item %d: bad image", i);
    if (!valid_blocks_in_fmt(b.src_fmt))
      return fail(err, errlen, SA_EINVAL, "item %d: bad pixel format", i);
    if (b.src_fmt == sa::kYuv420 && ((b.width & 1) || (b.height & 1)))
      return fail(err, errlen, SA_EUNSUPPORTED, "item %d: 4:2:0 input needs even width and height",
                  i);
  }
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  std::vector<sa::OctreeReq> req(n);
  std::vector<sa::OctreeOut> o(n);
  std::vector<sa::OctreeResult> r(n);
  for (int i = 0; i < n; ++i) {
    const sa_build& b = items[i];
    req[i] = sa::OctreeReq{b.src, b.src_fmt, b.width, b.height, b.colors, b.assoc, b.gray};
    o[i].keep_tree = b.keep_tree ? 1 : 0;
  }
  sa::build_octree_batch(req.data(), n, (hipStream_t)stream, o.data(), r.data());
  int first = -1;
  for (int i = 0; i < n; ++i) {
    if (status) status[i] = r[i].status;
    if (r[i].status == 0) {
      out[i] = wrap_built(o[i], r[i], items[i].assoc);
    } else if (first < 0) {
      first = i;
    }
  }
  if (first >= 0) return fail(err, errlen, r[first].status, "item %d: %s", first, r[first].msg);
  return SA_OK;
}

extern "C" int sa_palette_info(sa_palette p, int* count, int* assoc, int* depth, int64_t* nodes,
                               int* exact_root) {
  if (!p) return SA_EINVAL;
  if (count) *count = p->count;
  if (assoc) *assoc = p->assoc;
  if (depth) *depth = p->depth;
  if (nodes) *nodes = p->nodes;
  if (exact_root) *exact_root = p->exact_root;
  return SA_OK;
}

extern "C" int sa_palette_entries(sa_palette p, double* out, char* err, size_t errlen) {
  if (!p || !out) return fail(err, errlen, SA_EINVAL, "null argument");
  if (p->dev_entries) {
    HIP_TRY(hipMemcpy(out, p->dev_entries, sizeof(double) * 4 * (size_t)p->count,
                      hipMemcpyDeviceToHost));
    return SA_OK;
  }
  // Uploaded palettes: the search image holds the entries as floats.
  std::vector<float> ent(4 * (size_t)p->count);
  HIP_TRY(hipMemcpy(ent.data(), (const unsigned char*)p->dev_image + p->hdr.off_ent,
                    sizeof(float) * ent.size(), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < ent.size(); ++i) out[i] = (double)ent[i];
  return SA_OK;
}

extern "C" int sa_palette_tree(sa_palette p, int32_t* child16, int32_t* parent,
                               uint32_t* color_number, uint64_t* number_unique, double* total4,
                               double* qe, uint32_t* id, uint32_t* level, char* err,
                               size_t errlen) {
  if (!p) return fail(err, errlen, SA_EINVAL, "null palette");
  if (!p->has_tree) return fail(err, errlen, SA_EINVAL, "palette was built without keep_tree");
  const size_t n = (size_t)p->nodes;
  const sa::TreeDev& t = p->tree;
  HIP_TRY(hipDeviceSynchronize());
  if (child16) HIP_TRY(hipMemcpy(child16, t.child, 64 * n, hipMemcpyDeviceToHost));
  if (parent) HIP_TRY(hipMemcpy(parent, t.parent, 4 * n, hipMemcpyDeviceToHost));
  if (color_number) HIP_TRY(hipMemcpy(color_number, t.color_number, 4 * n, hipMemcpyDeviceToHost));
  if (number_unique) HIP_TRY(hipMemcpy(number_unique, t.nu, 8 * n, hipMemcpyDeviceToHost));
  if (total4) HIP_TRY(hipMemcpy(total4, t.total, 32 * n, hipMemcpyDeviceToHost));
  if (qe) HIP_TRY(hipMemcpy(qe, t.qe, 8 * n, hipMemcpyDeviceToHost));
  if (id) HIP_TRY(hipMemcpy(id, t.id, 4 * n, hipMemcpyDeviceToHost));
  if (level) HIP_TRY(hipMemcpy(level, t.level, 4 * n, hipMemcpyDeviceToHost));
  return SA_OK;
}

extern "C" int sa_palette_count(sa_palette p) { return p ? p->count : 0; }

// ---- curve -------------------------------------------------------------------
extern "C" int sa_curve_order(int64_t w, int64_t h, int32_t* order, void* stream, char* err,
                              size_t errlen) {
  if (w < 1 || h < 1 || !order) return fail(err, errlen, SA_EINVAL, "bad geometry");
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  return build_curve(w, h, order, (hipStream_t)stream, err, errlen);
}

// ---- sequential engine -----------------------------------------------------
extern "C" int sa_walk_seq_batch(const sa_image* images, int n, const sa_seq_options* opt,
                                 void* stream, char* err, size_t errlen) {
  if (n < 0 || (n > 0 && !images)) return fail(err, errlen, SA_EINVAL, "bad image list");
  if (n == 0) return SA_OK;
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  sa_seq_options o{1.0, 1, 0, 0, 0};
  if (opt) o = *opt;
  for (int i = 0; i < n; ++i) {
    const sa_image& im = images[i];
    if (!im.src || !im.dst || !im.palette || im.width < 1 || im.height < 1)
      return fail(err, errlen, SA_EINVAL, "image %d: null pointer or empty geometry", i);
    if (!valid_in_fmt(im.src_fmt) || !valid_out_fmt(im.dst_fmt))
      return fail(err, errlen, SA_EINVAL, "image %d: bad pixel format", i);
    if (im.dst == im.src && (im.src_fmt != im.dst_fmt))
      return fail(err, errlen, SA_EINVAL, "image %d: in-place needs the same format", i);
  }
  const double* sw = nullptr;
  rc = get_weights(o.diffusion, s, &sw, err, errlen);
  if (rc) return rc;
  // Two groups by alpha association (different tree arity and key width).
  for (int assoc = 0; assoc < 2; ++assoc) {
    std::vector<int> sel;
    for (int i = 0; i < n; ++i)
      if (images[i].palette->assoc == assoc) sel.push_back(i);
    if (sel.empty()) continue;
    // Largest first: the makespan is set by the biggest walk.
    std::stable_sort(sel.begin(), sel.end(), [&](int a, int b) {
      return images[a].width * images[a].height > images[b].width * images[b].height;
    });
    int max_img = 0;
    int64_t max_n = 0, total = 0;
    for (int i : sel) {
      max_img = std::max(max_img, images[i].palette->hdr.bytes);
      const int64_t px = images[i].width * images[i].height;
      max_n = std::max(max_n, px);
      total += px;
    }
    if (o.memo_bits != 0 && (o.memo_bits < 4 || o.memo_bits > 10))
      return fail(err, errlen, SA_EINVAL, "memo_bits %d outside 4..10", o.memo_bits);
    std::vector<MemoShape> shape(sel.size());
    size_t lds = 0;
    for (size_t k = 0; k < sel.size(); ++k) {
      const sa_image& im = images[sel[k]];
      const int img_lds = sa::seq_img_lds_bytes(im.palette->hdr.bytes);
      shape[k] = memo_shape(o, assoc, img_lds, im.palette->count);
      lds = std::max(lds, shape[k].bytes + (size_t)img_lds + sa::kSeqRingBytes);
    }
    if (lds > g_lds_per_cu) return fail(err, errlen, SA_EUNSUPPORTED, "LDS %zu too large", lds);
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, g_lds_per_cu / lds));
    int waves = o.waves > 0 ? o.waves : g_cus * per_cu;
    waves = std::min<int>(waves, (int)sel.size());
    // Scratch: staged inputs, indices, HBM memo tables.
    std::vector<sa::SeqItem> items(sel.size());
    size_t xs_bytes = 0, idx_bytes = 0, spill_bytes = 0;
    for (size_t k = 0; k < sel.size(); ++k) {
      const int64_t px = images[sel[k]].width * images[sel[k]].height;
      xs_bytes += stage_bytes(images[sel[k]].src_fmt, px);
      idx_bytes += (2 * (size_t)px + 255) & ~(size_t)255;
      spill_bytes += shape[k].table;
    }
    unsigned char* scratch = nullptr;
    const size_t items_bytes = (sizeof(sa::SeqItem) * sel.size() + 255) & ~(size_t)255;
    const size_t total_bytes = items_bytes + 256 + xs_bytes + idx_bytes + spill_bytes;
    HIP_TRY(sa::scratch_alloc((void**)&scratch, total_bytes, s));
    unsigned char* p = scratch + items_bytes + 256;
    for (size_t k = 0; k < sel.size(); ++k) {
      const sa_image& im = images[sel[k]];
      const int64_t px = im.width * im.height;
      sa::SeqItem& it = items[k];
      const int32_t* curve = nullptr;
      rc = get_curve(im.width, im.height, s, &curve, err, errlen);
      if (rc) {
        (void)sa::scratch_free(scratch, s);
        return rc;
      }
      it.src = im.src;
      it.dst = im.dst;
      it.curve = curve;
      it.stage = p;
      p += stage_bytes(im.src_fmt, px);
      it.idx = (uint16_t*)p;
      p += (2 * (size_t)px + 255) & ~(size_t)255;
      it.memo = nullptr;
      if (o.use_cache) {
        it.memo = p;
        p += shape[k].table;
      }
      it.pal = im.palette->dev_hdr;
      it.state = nullptr;
      it.snap = nullptr;
      it.n = px;
      it.src_fmt = im.src_fmt;
      it.dst_fmt = im.dst_fmt;
      it.memo_slots = shape[k].slots;
      it.memo_eb = shape[k].eb;
    }
    sa::SeqItem* d_items = (sa::SeqItem*)scratch;
    int32_t* d_queue = (int32_t*)(scratch + items_bytes);
    HIP_TRY(staged_copy(d_items, items.data(), sizeof(sa::SeqItem) * items.size(), s));
    HIP_TRY(hipMemsetAsync(d_queue, 0, sizeof(int32_t), s));
    const int ni = (int)items.size();
    HIP_TRY(timed("seq_gather", s, [&] { return sa::launch_seq_gather(d_items, ni, max_n, s); }));
    if (o.use_cache) HIP_TRY(sa::launch_memo_clear(d_items, ni, assoc ? 24 : 18, s));
    HIP_TRY(timed("seq_walk", s, [&] {
      return sa::launch_seq_walk(d_items, ni, d_queue, sw, o.use_cache, lds, waves, assoc, 0, 0,
                                 g_stats, s);
    }));
    HIP_TRY(timed("seq_pack", s, [&] { return sa::launch_seq_pack(d_items, ni, max_n, s); }));
    HIP_TRY(sa::scratch_free(scratch, s));
    (void)total;
  }
  return SA_OK;
}

// ---- resumable walks: continuous batching -----------------------------------
namespace {

MemoShape memo_shape(const sa_seq_options& o, int assoc, int img_bytes, int count) {
  MemoShape m;
  if (!o.use_cache) return m;
  m.eb = count <= 255 ? 1 : count <= 65535 ? 2 : 4;
  m.key_bits = assoc ? 24 : 18;
  m.table = (size_t)m.eb << m.key_bits;
  const size_t per = sa::seq_memo_bytes(1, m.eb);
  if (o.memo_bits > 0) {
    m.slots = 1 << o.memo_bits;
  } else {
    const size_t budget = std::min<size_t>(g_lds_per_cu, 160 * 1024);
    const size_t used = (size_t)img_bytes + sa::kSeqRingBytes;
    const size_t room = budget > used ? budget - used : 0;
    m.slots = (int)std::min<size_t>(4096, room / per) & ~3;
    const int tuned = g_knobs.seq_memo_bits.load();
    if (tuned > 0) m.slots = std::min(m.slots, 1 << tuned);
    m.slots = std::max(m.slots, 16);
  }
  m.bytes = sa::seq_memo_bytes(m.slots, m.eb);
  return m;
}


struct JobItem {
  int64_t id = 0;
  sa::SeqItem it{};
  int64_t t = 0;               // progress (every step advances it deterministically)
  unsigned char* mem = nullptr;  // xs, idx, memo, state, snapshot
  size_t mem_bytes = 0;
  hipEvent_t mem_ev = nullptr;   // recorded after the block's last use when it is parked
  int assoc = 0;
  int img_bytes = 0;
};

// Device copy of `items` in stream-ordered scratch (freed by the caller).
int upload_items(const std::vector<sa::SeqItem>& items, hipStream_t s, unsigned char** out,
                 char* err, size_t errlen) {
  const size_t bytes = (sizeof(sa::SeqItem) * items.size() + 255) & ~(size_t)255;
  HIP_TRY(sa::scratch_alloc((void**)out, bytes + 256, s));
  HIP_TRY(staged_copy(*out, items.data(), sizeof(sa::SeqItem) * items.size(), s));
  return SA_OK;
}

}  // namespace

struct sa_walk_job_s {
  sa_seq_options opt{1.0, 1, 0, 0, 0};
  const double* sw = nullptr;
  std::vector<JobItem> items;  // unfinished
  int64_t next_id = 0;
  std::mutex mu;
  // Scratch blocks of retired images, kept for images added later: a
  // stream-ordered free costs milliseconds per large block on this runtime,
  // so a serving loop that retires hundreds of images would stall on them.
  // Each block carries an event recorded after its last use; a reuse waits on
  // it.  Oldest first; above kSpareBytes the oldest go back to the runtime.
  struct Spare {
    unsigned char* mem;
    size_t bytes;
    hipEvent_t ev;
  };
  std::vector<Spare> spare;
  size_t spare_bytes = 0;
  // Overlapped steps: the progress of the step in flight, copied to pinned
  // memory, and the ids it covered (two slots, alternating).
  struct Flight {
    std::vector<int64_t> ids;
    int64_t* host = nullptr;  // pinned, `cap` entries
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool live = false;
  } flight[2];
  int cur = 0;
};

namespace {

constexpr size_t kSpareBytes = (size_t)16 << 30;

// Parks an image's scratch block once the work queued on `s` so far is done
// with it; the oldest blocks above kSpareBytes are freed.
int park_block(sa_walk_job job, JobItem& J, hipStream_t s, char* err, size_t errlen) {
  if (!J.mem) return SA_OK;
  if (!J.mem_ev) HIP_TRY(hipEventCreateWithFlags(&J.mem_ev, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(J.mem_ev, s));
  job->spare.push_back({J.mem, J.mem_bytes, J.mem_ev});
  job->spare_bytes += J.mem_bytes;
  J.mem = nullptr;
  J.mem_ev = nullptr;
  while (job->spare_bytes > kSpareBytes) {
    const sa_walk_job_s::Spare o = job->spare.front();
    job->spare.erase(job->spare.begin());
    job->spare_bytes -= o.bytes;
    HIP_TRY(hipStreamWaitEvent(s, o.ev, 0));
    HIP_TRY(sa::scratch_free(o.mem, s));
    HIP_TRY(hipEventDestroy(o.ev));
  }
  return SA_OK;
}

// Gives J the smallest parked block of `bytes` to twice that, with `s`
// ordered after its last use; J.mem stays null when none fits.
int reuse_block(sa_walk_job job, JobItem& J, size_t bytes, hipStream_t s, char* err,
                size_t errlen) {
  int best = -1;
  for (int k = 0; k < (int)job->spare.size(); ++k) {
    const size_t b = job->spare[k].bytes;
    if (b >= bytes && b / 2 <= bytes && (best < 0 || b < job->spare[best].bytes)) best = k;
  }
  if (best < 0) return SA_OK;
  const sa_walk_job_s::Spare o = job->spare[best];
  job->spare.erase(job->spare.begin() + best);
  job->spare_bytes -= o.bytes;
  J.mem = o.mem;
  J.mem_bytes = o.bytes;
  J.mem_ev = o.ev;
  HIP_TRY(hipStreamWaitEvent(s, o.ev, 0));
  return SA_OK;
}

}  // namespace

extern "C" int sa_walk_job_create(const sa_seq_options* opt, sa_walk_job* out, char* err,
                                  size_t errlen) {
  if (!out) return fail(err, errlen, SA_EINVAL, "null job");
  *out = nullptr;
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  sa_walk_job j = new sa_walk_job_s;
  if (opt) j->opt = *opt;
  if (j->opt.memo_bits != 0 && (j->opt.memo_bits < 4 || j->opt.memo_bits > 10)) {
    delete j;
    return fail(err, errlen, SA_EINVAL, "memo_bits %d outside 4..10", opt->memo_bits);
  }
  rc = get_weights(j->opt.diffusion, nullptr, &j->sw, err, errlen);
  if (rc) {
    delete j;
    return rc;
  }
  *out = j;
  return SA_OK;
}

extern "C" int sa_walk_job_add(sa_walk_job job, const sa_image* images, int n, int64_t* ids,
                               void* stream, char* err, size_t errlen) {
  if (!job) return fail(err, errlen, SA_EINVAL, "null job");
  if (n < 0 || (n > 0 && !images)) return fail(err, errlen, SA_EINVAL, "bad image list");
  if (n == 0) return SA_OK;
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> g(job->mu);
  const sa_seq_options& o = job->opt;
  for (int i = 0; i < n; ++i) {
    const sa_image& im = images[i];
    if (!im.src || !im.dst || !im.palette || im.width < 1 || im.height < 1)
      return fail(err, errlen, SA_EINVAL, "image %d: null pointer or empty geometry", i);
    if (!valid_in_fmt(im.src_fmt) || !valid_out_fmt(im.dst_fmt))
      return fail(err, errlen, SA_EINVAL, "image %d: bad pixel format", i);
    if (im.dst == im.src && (im.src_fmt != im.dst_fmt))
      return fail(err, errlen, SA_EINVAL, "image %d: in-place needs the same format", i);
    const int img_lds = sa::seq_img_lds_bytes(im.palette->hdr.bytes);
    const MemoShape ms = memo_shape(o, im.palette->assoc, img_lds, im.palette->count);
    const size_t lds = ms.bytes + (size_t)img_lds + sa::kSeqRingBytes;
    if (lds > g_lds_per_cu) return fail(err, errlen, SA_EUNSUPPORTED, "image %d: LDS %zu too large", i, lds);
  }
  std::vector<JobItem> fresh;
  std::vector<sa::SeqItem> dev[2];
  int64_t max_n = 0;
  for (int i = 0; i < n; ++i) {
    const sa_image& im = images[i];
    JobItem J;
    J.id = job->next_id + i;
    J.assoc = im.palette->assoc;
    J.img_bytes = sa::seq_img_lds_bytes(im.palette->hdr.bytes);  // its LDS share
    const MemoShape ms = memo_shape(o, J.assoc, J.img_bytes, im.palette->count);
    const int64_t px = im.width * im.height;
    const size_t xs_b = stage_bytes(im.src_fmt, px), idx_b = (2 * (size_t)px + 255) & ~(size_t)255;
    const size_t memo_b = ms.table;
    const size_t st_b = (sizeof(sa::SeqState) + 255) & ~(size_t)255;
    const size_t snap_b = ms.bytes;
    int rc = SA_OK;
    const int32_t* curve = nullptr;
    rc = get_curve(im.width, im.height, s, &curve, err, errlen);
    const size_t need = xs_b + idx_b + memo_b + st_b + snap_b;
    if (rc == SA_OK) rc = reuse_block(job, J, need, s, err, errlen);
    if (rc == SA_OK && !J.mem) {
      hipError_t e = sa::scratch_alloc((void**)&J.mem, need, s);
      if (e != hipSuccess) rc = fail(err, errlen, SA_ENOMEM, "image %d: %s", i, hipGetErrorString(e));
      J.mem_bytes = need;
    }
    if (rc != SA_OK) {
      for (JobItem& f : fresh) (void)park_block(job, f, s, nullptr, 0);
      (void)park_block(job, J, s, nullptr, 0);
      return rc;
    }
    unsigned char* p = J.mem;
    sa::SeqItem& it = J.it;
    it.src = im.src;
    it.dst = im.dst;
    it.curve = curve;
    it.stage = p;
    p += xs_b;
    it.idx = (uint16_t*)p;
    p += idx_b;
    it.memo = o.use_cache ? p : nullptr;
    p += memo_b;
    it.state = (sa::SeqState*)p;
    p += st_b;
    it.snap = o.use_cache ? p : nullptr;
    it.pal = im.palette->dev_hdr;
    it.n = px;
    it.src_fmt = im.src_fmt;
    it.dst_fmt = im.dst_fmt;
    it.memo_slots = ms.slots;
    it.memo_eb = ms.eb;
    max_n = std::max(max_n, px);
    dev[J.assoc].push_back(it);
    fresh.push_back(J);
  }
  // Inputs in visit order, zeroed memo tables, progress 0 -- on `stream`.
  for (int assoc = 0; assoc < 2; ++assoc) {
    if (dev[assoc].empty()) continue;
    unsigned char* d = nullptr;
    int rc = upload_items(dev[assoc], s, &d, err, errlen);
    if (rc) {
      for (JobItem& f : fresh) (void)park_block(job, f, s, nullptr, 0);
      return rc;
    }
    const int ni = (int)dev[assoc].size();
    const sa::SeqItem* di = (const sa::SeqItem*)d;
    HIP_TRY(timed("seq_gather", s, [&] { return sa::launch_seq_gather(di, ni, max_n, s); }));
    if (o.use_cache) HIP_TRY(sa::launch_memo_clear(di, ni, assoc ? 24 : 18, s));
    HIP_TRY(sa::launch_seq_state_init(di, ni, s));
    HIP_TRY(sa::scratch_free(d, s));
  }
  for (int i = 0; i < n; ++i) {
    if (ids) ids[i] = fresh[i].id;
    job->items.push_back(fresh[i]);
  }
  job->next_id += n;
  return SA_OK;
}

namespace {

// One step of a job: a visit slice (slice > 0, progress known on the host),
// a time budget (budget > 0 ticks, progress read back) or to the end.
int job_step(sa_walk_job job, int64_t slice, uint64_t budget, hipStream_t s, int64_t* done,
             int cap, int* n_done, int* active, char* err, size_t errlen) {
  const sa_seq_options& o = job->opt;
  int nd = 0;
  // Progress of overlapped steps still in flight comes first.
  for (int k = 0; k < 2; ++k) {
    auto& F = job->flight[(job->cur + k) & 1];
    if (!F.live) continue;
    HIP_TRY(hipEventSynchronize(F.ev));
    for (size_t q = 0; q < F.ids.size(); ++q)
      for (JobItem& J : job->items)
        if (J.id == F.ids[q]) J.t = std::max(J.t, F.host[q]);
    F.live = false;
  }
  for (int assoc = 0; assoc < 2; ++assoc) {
    std::vector<size_t> sel;
    for (size_t k = 0; k < job->items.size(); ++k)
      if (job->items[k].assoc == assoc) sel.push_back(k);
    if (sel.empty()) continue;
    // Longest remaining walk first: it sets the step's length.
    std::stable_sort(sel.begin(), sel.end(), [&](size_t a, size_t b) {
      return job->items[a].it.n - job->items[a].t > job->items[b].it.n - job->items[b].t;
    });
    size_t lds = 0;
    for (size_t k : sel) {
      const JobItem& J = job->items[k];
      lds = std::max(lds, (o.use_cache ? sa::seq_memo_bytes(J.it.memo_slots, J.it.memo_eb) : 0u) +
                                 (size_t)J.img_bytes +
                                 sa::kSeqRingBytes);
    }
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, g_lds_per_cu / lds));
    int waves = o.waves > 0 ? o.waves : g_cus * per_cu;
    // A bounded step (slice or budget) over more images than wavefronts walks
    // the longest ones only; a step to the end walks every image (the
    // persistent grid pulls the rest as walks finish).
    if ((slice > 0 || budget > 0) && (int)sel.size() > waves) sel.resize((size_t)waves);
    std::vector<sa::SeqItem> its;
    for (size_t k : sel) its.push_back(job->items[k].it);
    waves = std::min<int>(waves, (int)its.size());
    const size_t items_bytes = (sizeof(sa::SeqItem) * its.size() + 255) & ~(size_t)255;
    unsigned char* d = nullptr;
    HIP_TRY(sa::scratch_alloc((void**)&d, items_bytes + 256 + 8 * its.size(), s));
    HIP_TRY(staged_copy(d, its.data(), sizeof(sa::SeqItem) * its.size(), s));
    int32_t* queue = (int32_t*)(d + items_bytes);
    int64_t* prog = (int64_t*)(d + items_bytes + 256);
    HIP_TRY(hipMemsetAsync(queue, 0, sizeof(int32_t), s));
    const int ni = (int)its.size();
    const sa::SeqItem* di = (const sa::SeqItem*)d;
    HIP_TRY(timed("seq_walk", s, [&] {
      return sa::launch_seq_walk(di, ni, queue, job->sw, o.use_cache, lds, waves, assoc, slice,
                                 budget, g_stats, s);
    }));
    std::vector<int64_t> t(its.size());
    if (budget > 0) {
      HIP_TRY(sa::launch_seq_progress(di, ni, prog, s));
      HIP_TRY(hipMemcpyAsync(t.data(), prog, 8 * its.size(), hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
    }
    HIP_TRY(sa::scratch_free(d, s));
    std::vector<sa::SeqItem> fin;
    int64_t max_n = 0;
    for (size_t q = 0; q < sel.size(); ++q) {
      JobItem& J = job->items[sel[q]];
      if (budget > 0)
        J.t = t[q];
      else
        J.t = slice > 0 ? std::min(J.it.n, J.t + slice) : J.it.n;
      if (J.t == J.it.n) {
        fin.push_back(J.it);
        max_n = std::max(max_n, J.it.n);
      }
    }
    if (!fin.empty()) {
      unsigned char* f = nullptr;
      int rc = upload_items(fin, s, &f, err, errlen);
      if (rc) return rc;
      const int nf = (int)fin.size();
      HIP_TRY(timed("seq_pack", s, [&] {
        return sa::launch_seq_pack((const sa::SeqItem*)f, nf, max_n, s);
      }));
      HIP_TRY(sa::scratch_free(f, s));
    }
  }
  // Retire the finished items (their scratch is freed after this step's work).
  std::vector<JobItem> keep;
  for (JobItem& J : job->items) {
    if (J.t == J.it.n) {
      if (done && nd < cap) done[nd] = J.id;
      ++nd;
      int rc = park_block(job, J, s, err, errlen);
      if (rc) return rc;
    } else {
      keep.push_back(J);
    }
  }
  job->items.swap(keep);
  if (n_done) *n_done = nd;
  if (active) *active = (int)job->items.size();
  return SA_OK;
}

}  // namespace

extern "C" int sa_walk_job_step(sa_walk_job job, int64_t visits, void* stream, int64_t* done,
                                int cap, int* n_done, int* active, char* err, size_t errlen) {
  if (!job) return fail(err, errlen, SA_EINVAL, "null job");
  std::lock_guard<std::mutex> g(job->mu);
  const int64_t slice = visits > 0 ? (visits + 47) / 48 * 48 : 0;
  return job_step(job, slice, 0, (hipStream_t)stream, done, cap, n_done, active, err, errlen);
}

extern "C" int sa_walk_job_step_timed(sa_walk_job job, double budget_us, void* stream,
                                      int64_t* done, int cap, int* n_done, int* active, char* err,
                                      size_t errlen) {
  if (!job) return fail(err, errlen, SA_EINVAL, "null job");
  if (!(budget_us > 0.0)) return fail(err, errlen, SA_EINVAL, "budget must be positive");
  std::lock_guard<std::mutex> g(job->mu);
  // s_memrealtime counts at 100 MHz.
  const uint64_t ticks = (uint64_t)std::max(1.0, budget_us * 100.0);
  return job_step(job, 0, ticks, (hipStream_t)stream, done, cap, n_done, active, err, errlen);
}

namespace {

// Packs and retires the items whose progress reached their end; `s` orders it.
int retire_finished(sa_walk_job job, hipStream_t s, int64_t* done, int cap, int* n_done,
                    char* err, size_t errlen) {
  int nd = 0;
  for (int assoc = 0; assoc < 2; ++assoc) {
    std::vector<sa::SeqItem> fin;
    int64_t max_n = 0;
    for (const JobItem& J : job->items)
      if (J.assoc == assoc && J.t == J.it.n) {
        fin.push_back(J.it);
        max_n = std::max(max_n, J.it.n);
      }
    if (fin.empty()) continue;
    unsigned char* f = nullptr;
    int rc = upload_items(fin, s, &f, err, errlen);
    if (rc) return rc;
    const int nf = (int)fin.size();
    HIP_TRY(timed("seq_pack", s, [&] {
      return sa::launch_seq_pack((const sa::SeqItem*)f, nf, max_n, s);
    }));
    HIP_TRY(sa::scratch_free(f, s));
  }
  std::vector<JobItem> keep;
  for (JobItem& J : job->items) {
    if (J.t == J.it.n) {
      if (done && nd < cap) done[nd] = J.id;
      ++nd;
      int rc = park_block(job, J, s, err, errlen);
      if (rc) return rc;
    } else {
      keep.push_back(J);
    }
  }
  job->items.swap(keep);
  if (n_done) *n_done = nd;
  return SA_OK;
}

}  // namespace

extern "C" int sa_walk_job_step_overlapped(sa_walk_job job, double budget_us, void* stream,
                                           int64_t* done, int cap, int* n_done, int* active,
                                           char* err, size_t errlen) {
  if (!job) return fail(err, errlen, SA_EINVAL, "null job");
  if (!(budget_us > 0.0)) return fail(err, errlen, SA_EINVAL, "budget must be positive");
  hipStream_t s = (hipStream_t)stream;
  std::lock_guard<std::mutex> g(job->mu);
  const sa_seq_options& o = job->opt;
  const uint64_t ticks = (uint64_t)std::max(1.0, budget_us * 100.0);
  // 1. Launch this step over every image not known to be finished (images
  //    that finished in the step still in flight are skipped by the kernel).
  auto& F = job->flight[job->cur];
  F.ids.clear();
  if (!job->items.empty()) {
    const size_t total = job->items.size();
    if (F.cap < total) {
      if (F.host) (void)hipHostFree(F.host);
      F.host = nullptr;
      F.cap = 0;
      HIP_TRY(hipHostMalloc((void**)&F.host, 8 * std::max<size_t>(total, 64), hipHostMallocDefault));
      F.cap = std::max<size_t>(total, 64);
    }
    if (!F.ev) HIP_TRY(hipEventCreateWithFlags(&F.ev, hipEventDisableTiming));
    size_t off = 0;
    for (int assoc = 0; assoc < 2; ++assoc) {
      std::vector<const JobItem*> cand;
      size_t lds = 0;
      for (const JobItem& J : job->items) {
        if (J.assoc != assoc) continue;
        cand.push_back(&J);
        lds = std::max(lds, (o.use_cache ? sa::seq_memo_bytes(J.it.memo_slots, J.it.memo_eb) : 0u) +
                                 (size_t)J.img_bytes +
                                 sa::kSeqRingBytes);
      }
      if (cand.empty()) continue;
      const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, g_lds_per_cu / lds));
      int waves = o.waves > 0 ? o.waves : g_cus * per_cu;
      // Most visits left first: the grid's first wavefronts (one per CU) take
      // the longest serial walks.  With more images than wavefronts the step
      // walks only the first `waves` (the rest wait), so the longest walks
      // advance a full budget every step instead of sharing rounds of the
      // launch with short ones.
      std::stable_sort(cand.begin(), cand.end(), [](const JobItem* a, const JobItem* b) {
        return a->it.n - a->t > b->it.n - b->t;
      });
      if ((int)cand.size() > waves) cand.resize((size_t)waves);
      std::vector<sa::SeqItem> its;
      for (const JobItem* J : cand) {
        its.push_back(J->it);
        F.ids.push_back(J->id);
      }
      waves = std::min<int>(waves, (int)its.size());
      const size_t items_bytes = (sizeof(sa::SeqItem) * its.size() + 255) & ~(size_t)255;
      unsigned char* d = nullptr;
      HIP_TRY(sa::scratch_alloc((void**)&d, items_bytes + 256 + 8 * its.size(), s));
      HIP_TRY(staged_copy(d, its.data(), sizeof(sa::SeqItem) * its.size(), s));
      int32_t* queue = (int32_t*)(d + items_bytes);
      int64_t* prog = (int64_t*)(d + items_bytes + 256);
      HIP_TRY(hipMemsetAsync(queue, 0, sizeof(int32_t), s));
      const int ni = (int)its.size();
      const sa::SeqItem* di = (const sa::SeqItem*)d;
      HIP_TRY(timed("seq_walk", s, [&] {
        return sa::launch_seq_walk(di, ni, queue, job->sw, o.use_cache, lds, waves, assoc, 0,
                                   ticks, g_stats, s);
      }));
      HIP_TRY(sa::launch_seq_progress(di, ni, prog, s));
      HIP_TRY(hipMemcpyAsync(F.host + off, prog, 8 * its.size(), hipMemcpyDeviceToHost, s));
      HIP_TRY(sa::scratch_free(d, s));
      off += its.size();
    }
    HIP_TRY(hipEventRecord(F.ev, s));
    F.live = true;
  }
  // 2. Collect the previous step's progress while this one runs.
  auto& P = job->flight[job->cur ^ 1];
  if (P.live) {
    HIP_TRY(hipEventSynchronize(P.ev));
    std::map<int64_t, int64_t> prog;
    for (size_t q = 0; q < P.ids.size(); ++q) prog[P.ids[q]] = P.host[q];
    for (JobItem& J : job->items) {
      auto it = prog.find(J.id);
      if (it != prog.end()) J.t = std::max(J.t, it->second);
    }
    P.live = false;
  }
  job->cur ^= 1;
  // 3. Pack and retire what has finished (ordered after this step).
  int rc = retire_finished(job, s, done, cap, n_done, err, errlen);
  if (rc) return rc;
  if (active) *active = (int)job->items.size();
  return SA_OK;
}

extern "C" int sa_walk_job_destroy(sa_walk_job job) {
  if (!job) return SA_OK;
  {
    std::lock_guard<std::mutex> g(job->mu);
    (void)hipDeviceSynchronize();
    // The blocks are the scratch cache's: they go back to it.
    for (JobItem& J : job->items) {
      (void)sa::scratch_free(J.mem, nullptr);
      if (J.mem_ev) (void)hipEventDestroy(J.mem_ev);
    }
    for (auto& o : job->spare) {
      (void)sa::scratch_free(o.mem, nullptr);
      (void)hipEventDestroy(o.ev);
    }
    for (auto& F : job->flight) {
      if (F.host) (void)hipHostFree(F.host);
      if (F.ev) (void)hipEventDestroy(F.ev);
    }
  }
  delete job;
  return SA_OK;
}

// ---- block engine ----------------------------------------------------------
namespace {
// A validated block-engine call: the reference's refusals, the weights, the
// block curve's length and gap cell, and the frame descriptors.
struct BlocksPlan {
  int64_t w = 0, h = 0, nb = 0, gap = -1;
  int n = 0, block = 0, assoc = -1, max_img = 0, max_count = 0, src_class = 0;
  int threads = 256;  // tuned launch shape
  const double* sw = nullptr;
  std::vector<sa::BlockFrame> fr;
};

int plan_blocks(const sa_image* frames, int n, int block, double diffusion, hipStream_t s,
                BlocksPlan& p, char* err, size_t errlen, bool no_alloc = false) {
  if (n < 1 || !frames)
    return fail(err, errlen, SA_EREFUSED,
                "invalid dither batch: frames must be >= 1 and block must be >= 16 (got frames=%d, block=%d)",
                n, block);
  if (block < 16)
    return fail(err, errlen, SA_EREFUSED,
                "invalid dither batch: frames must be >= 1 and block must be >= 16 (got frames=%d, block=%d)",
                n, block);
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  p.n = n;
  p.block = block;
  p.threads = g_knobs.blocks_threads.load();
  p.w = frames[0].width;
  p.h = frames[0].height;
  for (int i = 0; i < n; ++i) {
    const sa_image& f = frames[i];
    if (!f.src || !f.dst || !f.palette) return fail(err, errlen, SA_EINVAL, "frame %d: null pointer", i);
    if (f.width != p.w || f.height != p.h)
      return fail(err, errlen, SA_EINVAL, "frame %d: geometry differs", i);
    if (f.dst == f.src) return fail(err, errlen, SA_EINVAL, "frame %d: dst aliases src", i);
    if (!valid_blocks_in_fmt(f.src_fmt) || !valid_blocks_out_fmt(f.dst_fmt))
      return fail(err, errlen, SA_EINVAL, "frame %d: bad pixel format", i);
    if (f.src_fmt == sa::kYuv420 && ((f.width & 1) || (f.height & 1)))
      return fail(err, errlen, SA_EUNSUPPORTED, "frame %d: 4:2:0 input needs even width and height",
                  i);
    if (f.palette->count > 256)
      return fail(err, errlen, SA_EREFUSED,
                  "the blocks engine carries one palette index per byte, so the palette would alias "
                  "onto the first 256 (%d colours requested)", f.palette->count);
    if (p.assoc < 0) p.assoc = f.palette->assoc;
    if (p.assoc != f.palette->assoc)
      return fail(err, errlen, SA_EUNSUPPORTED, "frames mix associated and plain alpha");
    p.max_img = std::max(p.max_img, f.palette->hdr.bytes);
    p.max_count = std::max(p.max_count, f.palette->count);
    p.src_class = std::max(p.src_class, sa::blocks_src_class(f.src_fmt));
  }
  if (p.w < 1 || p.h < 1) return fail(err, errlen, SA_EINVAL, "bad geometry");
  if (sa::blocks_lds_bytes(p.max_img, p.max_count, p.src_class, 64) > sa::kBlocksMaxLds)
    return fail(err, errlen, SA_EUNSUPPORTED,
                "the blocks engine stages the palette's search image in LDS (%d bytes): too large",
                p.max_img);
  if (!no_alloc) {
    rc = get_weights(diffusion, s, &p.sw, err, errlen);
    if (rc) return rc;
  }
  const int level = curve_level(p.w, p.h);
  // Block curve length: in-bounds of [0, 4^L - 1) plus the origin; the gap
  // cell (2^L-1, 0) is in bounds exactly when w == 2^L.
  if (level > 0) {
    const bool gap_in = p.w == ((int64_t)1 << level);
    p.nb = p.w * p.h - (gap_in ? 1 : 0) + 1;
    if (gap_in) p.gap = p.w - 1;
  } else {
    p.gap = 0;  // 1x1: the block curve is empty (src/rd_riemersma_cpu.cpp:275)
  }
  if (no_alloc) return SA_OK;
  p.fr.resize((size_t)n);
  for (int i = 0; i < n; ++i) {
    p.fr[(size_t)i].src = frames[i].src;
    p.fr[(size_t)i].dst = frames[i].dst;
    p.fr[(size_t)i].pal = frames[i].palette->dev_hdr;
    p.fr[(size_t)i].src_fmt = frames[i].src_fmt;
    p.fr[(size_t)i].dst_fmt = frames[i].dst_fmt;
  }
  return SA_OK;
}

// The block walk and the gap fill, launched on s (kernel launches only, so a
// stream capture records exactly these).
hipError_t launch_blocks_plan(const BlocksPlan& p, const sa::BlockFrame* d_fr,
                              const int32_t* curve, const double* sw, hipStream_t s) {
  hipError_t e = sa::launch_blocks(d_fr, p.n, curve, p.w, p.h, p.nb, p.block, sw, p.assoc,
                                   p.max_img, p.max_count, p.src_class, p.threads, s);
  if (e == hipSuccess && p.gap >= 0) e = sa::launch_blocks_fill(d_fr, p.n, p.gap, p.w, p.h, s);
  return e;
}
}  // namespace

extern "C" int sa_walk_blocks(const sa_image* frames, int n, int block, double diffusion,
                              void* stream, char* err, size_t errlen) {
  hipStream_t s = (hipStream_t)stream;
  BlocksPlan p;
  int rc = plan_blocks(frames, n, block, diffusion, s, p, err, errlen);
  if (rc) return rc;
  const int32_t* curve = nullptr;
  rc = get_curve(p.w, p.h, s, &curve, err, errlen);
  if (rc) return rc;
  sa::BlockFrame* d_fr = nullptr;
  HIP_TRY(sa::scratch_alloc((void**)&d_fr, sizeof(sa::BlockFrame) * p.fr.size(), s));
  HIP_TRY(staged_copy(d_fr, p.fr.data(), sizeof(sa::BlockFrame) * p.fr.size(), s));
  HIP_TRY(timed("blocks_walk", s, [&] {
    return sa::launch_blocks(d_fr, p.n, curve, p.w, p.h, p.nb, p.block, p.sw, p.assoc, p.max_img,
                             p.max_count, p.src_class, p.threads, s);
  }));
  if (p.gap >= 0) HIP_TRY(sa::launch_blocks_fill(d_fr, p.n, p.gap, p.w, p.h, s));
  HIP_TRY(sa::scratch_free(d_fr, s));
  return SA_OK;
}

// ---- block walk without allocation ---------------------------------------------------
// sa_walk_blocks for callers that cannot allocate: the walk's scratch (frame
// table, weights, curve chunk counts, curve) lives in one fixed device array,
// so the call makes no hipMalloc/hipFree and no pinned-host allocation.  Its
// device address is taken once per device by sa_init (hipGetSymbolAddress);
// host code uses only that address, never the array's name.  Every call on a
// device shares its array: calls that overlap on one device race.
namespace {
constexpr size_t kNoAllocBytes = (size_t)40 << 20;        // the array: 40 MiB
constexpr size_t kNoAllocHead = (size_t)64 << 10;         // frames, weights, counts
constexpr int kNoAllocMaxFrames = 1024;
constexpr int kNoAllocMaxLevel = 12;                       // sides up to 4096
constexpr size_t kNoAllocWeightsOff = sizeof(sa::BlockFrame) * kNoAllocMaxFrames;
constexpr size_t kNoAllocCountsOff = kNoAllocWeightsOff + 256;
static_assert(kNoAllocCountsOff + sizeof(int32_t) * (((size_t)1 << (2 * kNoAllocMaxLevel)) / 4096) <=
                  kNoAllocHead,
              "the head holds the frame table, the weights and the chunk counts");
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_na_mu;
unsigned char* g_na_ptr[kMaxDevices] = {};  // scratch's address on each device

// Called by sa_init with `device` current.
int bind_noalloc_scratch(int device, char* err, size_t errlen) {
  if (device < 0 || device >= kMaxDevices) return SA_OK;  // the entry point then refuses
  std::lock_guard<std::mutex> g(g_na_mu);
  if (g_na_ptr[device]) return SA_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(err, errlen, SA_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_na_ptr[device] = (unsigned char*)ptr;
  return SA_OK;
}

// Up to 32 frame descriptors and the weights, carried by value in the kernel
// arguments (the runtime copies them at launch): no host staging is needed.
constexpr int kSetupFrames = 32;
struct NoAllocSetup {
  sa::BlockFrame fr[kSetupFrames];
  double sw[16];
  int first, count, with_weights;
};

__global__ void noalloc_setup_kernel(NoAllocSetup a, sa::BlockFrame* frames, double* sw) {
  const int t = threadIdx.x;
  if (t < a.count) frames[a.first + t] = a.fr[t];
  if (a.with_weights && t < 16) sw[t] = a.sw[t];
}
}  // namespace

extern "C" int sa_walk_blocks_noalloc(const sa_image* frames, int n, int block, double diffusion,
                                      void* stream, char* err, size_t errlen) {
  hipStream_t s = (hipStream_t)stream;
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> g(g_na_mu);
    if (dev >= 0 && dev < kMaxDevices) base = g_na_ptr[dev];
  }
  if (!base)
    return fail(err, errlen, SA_EINVAL, "sa_walk_blocks_noalloc: call sa_init on device %d first",
                dev);
  if (n > kNoAllocMaxFrames)
    return fail(err, errlen, SA_EUNSUPPORTED,
                "sa_walk_blocks_noalloc: %d frames, at most %d per call", n, kNoAllocMaxFrames);
  BlocksPlan p;
  int rc = plan_blocks(frames, n, block, diffusion, s, p, err, errlen, true);
  if (rc) return rc;
  const int level = curve_level(p.w, p.h);
  const size_t curve_bytes = sizeof(int32_t) * (size_t)(p.w * p.h);
  if (level > kNoAllocMaxLevel || kNoAllocHead + curve_bytes > kNoAllocBytes)
    return fail(err, errlen, SA_EUNSUPPORTED,
                "sa_walk_blocks_noalloc: a %lldx%lld frame needs %zu bytes of scratch; the fixed "
                "maximum is %zu (sides up to 4096, width*height up to %zu)",
                (long long)p.w, (long long)p.h, kNoAllocHead + curve_bytes, kNoAllocBytes,
                (kNoAllocBytes - kNoAllocHead) / sizeof(int32_t));
  sa::BlockFrame* d_fr = (sa::BlockFrame*)base;
  double* d_sw = (double*)(base + kNoAllocWeightsOff);
  int32_t* d_counts = (int32_t*)(base + kNoAllocCountsOff);
  int32_t* d_curve = (int32_t*)(base + kNoAllocHead);
  // Frame table and weights, 32 frames per launch.
  for (int first = 0; first < n; first += kSetupFrames) {
    NoAllocSetup a{};
    a.first = first;
    a.count = std::min(kSetupFrames, n - first);
    a.with_weights = first == 0;
    if (first == 0) scaled_weights(diffusion, a.sw);
    for (int k = 0; k < a.count; ++k) {
      const sa_image& f = frames[first + k];
      a.fr[k] = sa::BlockFrame{f.src, f.dst, f.palette->dev_hdr, f.src_fmt, f.dst_fmt};
    }
    noalloc_setup_kernel<<<dim3(1), dim3(64), 0, s>>>(a, d_fr, d_sw);
    HIP_TRY(hipGetLastError());
  }
  // The curve, rebuilt on every call (the array keeps only the last call's).
  const int64_t chunks = sa::curve_chunks(level);
  HIP_TRY(sa::launch_curve_count(level, p.w, p.h, d_counts, chunks, s));
  HIP_TRY(sa::launch_curve_scan(d_counts, chunks, s));
  HIP_TRY(sa::launch_curve_emit(level, p.w, p.h, d_counts, chunks, d_curve, s));
  p.sw = d_sw;
  HIP_TRY(launch_blocks_plan(p, d_fr, d_curve, d_sw, s));
  return SA_OK;
}

// ---- threshold-screen dithers ------------------------------------------------------
namespace {
// The four screens in one device buffer, built once: Bayer 8x8, void-and-
// cluster 64x64 and 32x32 ranks, the clustered-dot cell.
struct Screens {
  std::once_flag once;
  uint16_t* dev = nullptr;
  hipError_t status = hipSuccess;
};
Screens g_screens;
const int kScreenSide[4] = {8, 64, 32, 8};
const size_t kScreenOff[4] = {0, 64, 64 + 4096, 64 + 4096 + 1024};

hipError_t screens_ready() {
  std::call_once(g_screens.once, [] {
    std::vector<uint16_t> all(kScreenOff[3] + 64);
    std::copy(sa::kBayer8, sa::kBayer8 + 64, all.begin());
    const std::vector<uint16_t> v64 = sa::void_cluster_ranks(64), v32 = sa::void_cluster_ranks(32);
    std::copy(v64.begin(), v64.end(), all.begin() + kScreenOff[1]);
    std::copy(v32.begin(), v32.end(), all.begin() + kScreenOff[2]);
    std::copy(sa::kDot8, sa::kDot8 + 64, all.begin() + kScreenOff[3]);
    hipError_t e = hipMalloc((void**)&g_screens.dev, sizeof(uint16_t) * all.size());
    if (e == hipSuccess)
      e = hipMemcpy(g_screens.dev, all.data(), sizeof(uint16_t) * all.size(), hipMemcpyHostToDevice);
    g_screens.status = e;
  });
  return g_screens.status;
}
}  // namespace

extern "C" int sa_dither_screen(const sa_image* frames, int n, int screen, double diffusion,
                                void* stream, char* err, size_t errlen) {
  static const char* kWho[4] = {"bayer-ordered", "void-and-cluster", "void-and-cluster-fast",
                                "clustered-dot"};
  if (screen < 0 || screen > 3) return fail(err, errlen, SA_EINVAL, "unknown screen %d", screen);
  if (n < 1 || !frames) return fail(err, errlen, SA_EINVAL, "%s: no frames", kWho[screen]);
  const int64_t w = frames[0].width, h = frames[0].height;
  for (int i = 0; i < n; ++i) {
    const sa_image& f = frames[i];
    if (!f.palette || f.palette->count <= 0)
      return fail(err, errlen, SA_EREFUSED, "%s: no palette", kWho[screen]);
    if (f.width <= 0 || f.height <= 0)
      return fail(err, errlen, SA_EREFUSED, "%s: empty frame", kWho[screen]);
    if (!f.src || !f.dst) return fail(err, errlen, SA_EINVAL, "frame %d: null pointer", i);
    if (f.width != w || f.height != h)
      return fail(err, errlen, SA_EINVAL, "frame %d: geometry differs", i);
    if (!valid_blocks_in_fmt(f.src_fmt) || !valid_blocks_out_fmt(f.dst_fmt))
      return fail(err, errlen, SA_EINVAL, "frame %d: bad pixel format", i);
    if (f.dst == f.src && f.src_fmt != f.dst_fmt)
      return fail(err, errlen, SA_EINVAL, "frame %d: in-place needs the same format", i);
    if (f.src_fmt == sa::kYuv420 && ((w & 1) || (h & 1)))
      return fail(err, errlen, SA_EUNSUPPORTED, "frame %d: 4:2:0 input needs even width and height",
                  i);
    if (!f.palette->dev_entries)
      return fail(err, errlen, SA_EINVAL, "frame %d: palette without entries", i);
  }
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  HIP_TRY(screens_ready());
  hipStream_t s = (hipStream_t)stream;
  std::vector<sa::ScreenFrame> fr((size_t)n);
  for (int i = 0; i < n; ++i)
    fr[(size_t)i] = sa::ScreenFrame{frames[i].src, frames[i].dst, frames[i].palette->dev_hdr,
                                    frames[i].palette->dev_entries, frames[i].src_fmt,
                                    frames[i].dst_fmt};
  sa::ScreenFrame* d_fr = nullptr;
  const size_t fr_bytes = (sizeof(sa::ScreenFrame) * fr.size() + 255) & ~(size_t)255;
  HIP_TRY(sa::scratch_alloc((void**)&d_fr, fr_bytes + sizeof(sa::ScreenRange) * fr.size(), s));
  sa::ScreenRange* d_rng = (sa::ScreenRange*)((unsigned char*)d_fr + fr_bytes);
  HIP_TRY(staged_copy(d_fr, fr.data(), sizeof(sa::ScreenFrame) * fr.size(), s));
  HIP_TRY(timed("screen_dither", s, [&] {
    return sa::launch_screen(d_fr, n, w, h, g_screens.dev + kScreenOff[screen], kScreenSide[screen],
                             screen, diffusion, d_rng, s);
  }));
  HIP_TRY(sa::scratch_free(d_fr, s));
  return SA_OK;
}

// ---- error-diffusion dithers --------------------------------------------------------
extern "C" int sa_dither_diffuse(const sa_image* frames, int n, int kind, double diffusion,
                                 void* stream, char* err, size_t errlen) {
  static const char* kWho[3] = {"floyd-steinberg", "atkinson", "jarvis"};
  if (kind < 0 || kind > 2) return fail(err, errlen, SA_EINVAL, "unknown diffusion dither %d", kind);
  if (n < 1 || !frames) return fail(err, errlen, SA_EINVAL, "%s: no frames", kWho[kind]);
  const int64_t w = frames[0].width, h = frames[0].height;
  for (int i = 0; i < n; ++i) {
    const sa_image& f = frames[i];
    if (!f.palette || f.palette->count <= 0)
      return fail(err, errlen, SA_EREFUSED, "%s: no palette", kWho[kind]);
    if (f.width <= 0 || f.height <= 0)
      return fail(err, errlen, SA_EREFUSED, "%s: empty frame", kWho[kind]);
    if (!f.src || !f.dst) return fail(err, errlen, SA_EINVAL, "frame %d: null pointer", i);
    if (f.width != w || f.height != h)
      return fail(err, errlen, SA_EINVAL, "frame %d: geometry differs", i);
    if (!valid_blocks_in_fmt(f.src_fmt) || !valid_blocks_out_fmt(f.dst_fmt))
      return fail(err, errlen, SA_EINVAL, "frame %d: bad pixel format", i);
    if (f.dst == f.src && f.src_fmt != f.dst_fmt)
      return fail(err, errlen, SA_EINVAL, "frame %d: in-place needs the same format", i);
    if (f.src_fmt == sa::kYuv420 && ((w & 1) || (h & 1)))
      return fail(err, errlen, SA_EUNSUPPORTED, "frame %d: 4:2:0 input needs even width and height",
                  i);
  }
  int max_count = 0;
  for (int i = 0; i < n; ++i) max_count = std::max(max_count, frames[i].palette->count);
  if (sa::diffuse_lds_bytes(kind, w, max_count) > g_lds_per_cu)
    return fail(err, errlen, SA_EUNSUPPORTED,
                "%s: a frame %lld pixels wide with %d entries needs %zu bytes of LDS for its "
                "error rows and entries (at most %zu)", kWho[kind], (long long)w, max_count,
                sa::diffuse_lds_bytes(kind, w, max_count), g_lds_per_cu);
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  std::vector<sa::ScreenFrame> fr((size_t)n);
  for (int i = 0; i < n; ++i)
    fr[(size_t)i] = sa::ScreenFrame{frames[i].src, frames[i].dst, frames[i].palette->dev_hdr,
                                    frames[i].palette->dev_entries, frames[i].src_fmt,
                                    frames[i].dst_fmt};
  sa::ScreenFrame* d_fr = nullptr;
  HIP_TRY(sa::scratch_alloc((void**)&d_fr, sizeof(sa::ScreenFrame) * fr.size(), s));
  HIP_TRY(staged_copy(d_fr, fr.data(), sizeof(sa::ScreenFrame) * fr.size(), s));
  HIP_TRY(timed("diffuse_dither", s, [&] {
    return sa::launch_diffuse(d_fr, n, w, h, kind, max_count, (float)diffusion, s);
  }));
  HIP_TRY(sa::scratch_free(d_fr, s));
  return SA_OK;
}

// ---- Jacobi approximation ----------------------------------------------------------
// BuildInverseFilter (src/rd_approx_cuda.cu:73-88) with the host libm weights.
static void inverse_filter(int taps, double diffusion, std::vector<double>& g) {
  double w[16];
  double weight = 1.0;
  for (int i = 0; i < 16; ++i) {
    w[i] = sa::safe_recip(weight);
    weight *= std::exp(std::log(1.0 / (1.0 / 16.0)) / (16 - 1.0));
  }
  g.assign((size_t)taps + 1, 0.0);
  g[0] = 1.0;
  for (int d = 1; d <= taps; ++d) {
    double acc = 0.0;
    const int kmax = d < 16 ? d : 16;
    for (int k = 1; k <= kmax; ++k) acc += (diffusion / 16.0) * w[k - 1] * g[(size_t)(d - k)];
    g[(size_t)d] = acc;
  }
}

extern "C" int sa_inverse_filter(int taps, double diffusion, double* g, char* err,
                                 size_t errlen) {
  if (taps < 1 || taps > 4096) return fail(err, errlen, SA_EREFUSED, "--approx-taps out of range");
  if (!g) return fail(err, errlen, SA_EINVAL, "null output");
  std::vector<double> v;
  inverse_filter(taps, diffusion, v);
  std::copy(v.begin(), v.end(), g);
  return SA_OK;
}

namespace {
// A validated Jacobi call: refusals, the inverse filter, the curve length and
// the scratch layout (descriptors, filter, iterate, two index buffers).
struct ApproxPlan {
  int64_t w = 0, h = 0, nb = 0;
  bool gap_in = false;
  int n = 0, iterations = 0, taps = 0, fp64 = 0, use_tree = 0, assoc = -1, max_img = 0;
  int per = 4;  // tuned launch shape
  std::vector<sa::BlockFrame> fr;
  std::vector<double> gd;
  std::vector<float> gf;
  size_t b_fr = 0, b_g = 0, b_cx = 0, b_q = 0;
  size_t scratch() const { return b_fr + b_g + b_cx + 2 * b_q; }
  const void* g_host() const { return fp64 ? (const void*)gd.data() : (const void*)gf.data(); }
  size_t g_bytes() const { return fp64 ? 8 * gd.size() : 4 * gf.size(); }
};

int plan_approx(const sa_image* frames, int n, int iterations, int taps, int fp64, int use_tree,
                double diffusion, ApproxPlan& p, char* err, size_t errlen) {
  if (n < 1 || !frames) return fail(err, errlen, SA_EINVAL, "no frames");
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  p.n = n;
  p.iterations = iterations;
  p.taps = taps;
  p.fp64 = fp64;
  p.use_tree = use_tree;
  p.per = g_knobs.approx_per.load();
  p.w = frames[0].width;
  p.h = frames[0].height;
  for (int i = 0; i < n; ++i) {
    const sa_image& f = frames[i];
    if (!f.src || !f.dst || !f.palette) return fail(err, errlen, SA_EINVAL, "frame %d: null pointer", i);
    if (f.width != p.w || f.height != p.h)
      return fail(err, errlen, SA_EINVAL, "frame %d: geometry differs", i);
    if (f.dst == f.src) return fail(err, errlen, SA_EINVAL, "frame %d: dst aliases src", i);
    if (!valid_in_fmt(f.src_fmt) || !valid_out_fmt(f.dst_fmt))
      return fail(err, errlen, SA_EINVAL, "frame %d: bad pixel format", i);
    if (f.palette->count < 1) return fail(err, errlen, SA_EREFUSED, "empty palette");
    if (f.palette->count > 256)
      return fail(err, errlen, SA_EREFUSED,
                  "the approx engine carries one palette index per byte, so %d colours would alias "
                  "onto the first 256", f.palette->count);
    if (p.assoc < 0) p.assoc = f.palette->assoc;
    if (p.assoc != f.palette->assoc)
      return fail(err, errlen, SA_EUNSUPPORTED, "frames mix associated and plain alpha");
    p.max_img = std::max(p.max_img, f.palette->hdr.bytes);
  }
  if (taps < 1 || taps > 4096) return fail(err, errlen, SA_EREFUSED, "--approx-taps out of range");
  if (iterations < 0) return fail(err, errlen, SA_EINVAL, "negative iterations");
  const int level = curve_level(p.w, p.h);
  p.gap_in = level > 0 && p.w == ((int64_t)1 << level);
  p.nb = level > 0 ? p.w * p.h - (p.gap_in ? 1 : 0) + 1 : 0;
  if (p.nb < p.w * p.h)  // src/rd_approx_cuda.cu:271-274 (a 1x1 image has an empty curve)
    return fail(err, errlen, SA_EREFUSED,
                "the compacted curve is shorter than the image; the Hilbert transcription is wrong");
  const size_t lds = sa::approx_lds_bytes(p.max_img, taps, fp64, p.assoc, p.per);
  if (lds > g_lds_per_cu)
    return fail(err, errlen, SA_EUNSUPPORTED, "taps %d with this palette need %zu B of LDS", taps, lds);
  inverse_filter(taps, diffusion, p.gd);
  p.gf.resize(p.gd.size());
  for (size_t i = 0; i < p.gd.size(); ++i) p.gf[i] = (float)p.gd[i];
  const size_t elem = fp64 ? 8 : 4;
  p.fr.resize((size_t)n);
  for (int i = 0; i < n; ++i) {
    p.fr[(size_t)i].src = frames[i].src;
    p.fr[(size_t)i].dst = frames[i].dst;
    p.fr[(size_t)i].pal = frames[i].palette->dev_hdr;
    p.fr[(size_t)i].src_fmt = frames[i].src_fmt;
    p.fr[(size_t)i].dst_fmt = frames[i].dst_fmt;
  }
  auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
  p.b_fr = up(sizeof(sa::BlockFrame) * p.fr.size());
  p.b_g = up(p.g_bytes());
  p.b_cx = up(elem * 4 * (size_t)p.nb * (size_t)n);
  p.b_q = up((size_t)p.nb * (size_t)n);
  return SA_OK;
}

// The sweeps and the gap fill on s, over a scratch laid out as the plan says
// (kernel launches only).
hipError_t launch_approx_plan(const ApproxPlan& p, unsigned char* scratch, const int32_t* curve,
                              hipStream_t s) {
  sa::BlockFrame* d_fr = (sa::BlockFrame*)scratch;
  void* d_g = scratch + p.b_fr;
  void* d_cx = scratch + p.b_fr + p.b_g;
  uint8_t* q0 = scratch + p.b_fr + p.b_g + p.b_cx;
  uint8_t* q1 = q0 + p.b_q;
  uint8_t* result = nullptr;
  hipError_t e = sa::launch_approx(d_fr, p.n, curve, p.nb, p.iterations, p.taps, d_g, p.fp64,
                                   p.use_tree, p.assoc, p.max_img, d_cx, q0, q1, &result, p.per,
                                   s);
  if (e == hipSuccess && p.gap_in) e = sa::launch_blocks_fill(d_fr, p.n, p.w - 1, p.w, p.h, s);
  return e;
}
}  // namespace

extern "C" int sa_walk_approx(const sa_image* frames, int n, int iterations, int taps, int fp64,
                              int use_tree, double diffusion, void* stream, char* err,
                              size_t errlen) {
  hipStream_t s = (hipStream_t)stream;
  ApproxPlan p;
  int rc = plan_approx(frames, n, iterations, taps, fp64, use_tree, diffusion, p, err, errlen);
  if (rc) return rc;
  const int32_t* curve = nullptr;
  rc = get_curve(p.w, p.h, s, &curve, err, errlen);
  if (rc) return rc;
  unsigned char* scratch = nullptr;
  HIP_TRY(sa::scratch_alloc((void**)&scratch, p.scratch(), s));
  HIP_TRY(staged_copy(scratch, p.fr.data(), sizeof(sa::BlockFrame) * p.fr.size(), s));
  HIP_TRY(staged_copy(scratch + p.b_fr, p.g_host(), p.g_bytes(), s));
  HIP_TRY(timed("approx", s, [&] { return launch_approx_plan(p, scratch, curve, s); }));
  HIP_TRY(sa::scratch_free(scratch, s));
  return SA_OK;
}

// ---- frame chains (hipGraph) -----------------------------------------------------------
// One engine's work for a fixed set of frame buffers, captured once into a
// graph that owns its scratch, curve and weights; a video loop that refills
// the same buffers replays it per batch of frames (one launch on the host).
struct sa_chain_s {
  void* mem = nullptr;  // descriptors, curve copy, weights, engine scratch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {
template <typename Launch>
int capture_chain(sa_chain c, Launch&& launch, char* err, size_t errlen) {
  hipStream_t cs = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    const hipError_t le = launch(cs);
    e = hipStreamEndCapture(cs, &c->graph);
    if (le != hipSuccess) e = le;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&c->exec, c->graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess)
    return fail(err, errlen, SA_EDEVICE, "graph capture: %s", hipGetErrorString(e));
  return SA_OK;
}
}  // namespace

extern "C" int sa_chain_destroy(sa_chain c) {
  if (!c) return SA_OK;
  (void)hipDeviceSynchronize();
  if (c->exec) (void)hipGraphExecDestroy(c->exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->mem) (void)hipFree(c->mem);
  delete c;
  return SA_OK;
}

extern "C" int sa_chain_create_blocks(const sa_image* frames, int n, int block, double diffusion,
                                      void* stream, sa_chain* out, char* err, size_t errlen) {
  if (!out) return fail(err, errlen, SA_EINVAL, "null output handle");
  *out = nullptr;
  hipStream_t s = (hipStream_t)stream;
  BlocksPlan p;
  int rc = plan_blocks(frames, n, block, diffusion, s, p, err, errlen);
  if (rc) return rc;
  const int32_t* curve = nullptr;
  rc = get_curve(p.w, p.h, s, &curve, err, errlen);
  if (rc) return rc;
  const size_t b_fr = (sizeof(sa::BlockFrame) * p.fr.size() + 255) & ~(size_t)255;
  const size_t b_sw = 256, b_curve = sizeof(int32_t) * (size_t)std::max<int64_t>(p.w * p.h, 1);
  sa_chain c = new sa_chain_s();
  auto bail = [&](int code) {
    sa_chain_destroy(c);
    return code;
  };
  if (hipMalloc(&c->mem, b_fr + b_sw + b_curve) != hipSuccess)
    return bail(fail(err, errlen, SA_ENOMEM, "chain: device allocation failed"));
  unsigned char* m = (unsigned char*)c->mem;
  sa::BlockFrame* d_fr = (sa::BlockFrame*)m;
  double* d_sw = (double*)(m + b_fr);
  int32_t* d_curve = (int32_t*)(m + b_fr + b_sw);
  hipError_t e = hipMemcpyAsync(d_fr, p.fr.data(), sizeof(sa::BlockFrame) * p.fr.size(),
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_sw, p.sw, 16 * sizeof(double), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess && p.w * p.h > 1)
    e = hipMemcpyAsync(d_curve, curve, sizeof(int32_t) * (size_t)(p.w * p.h), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return bail(fail(err, errlen, SA_EDEVICE, "chain setup: %s", hipGetErrorString(e)));
  rc = capture_chain(c, [&](hipStream_t cs) { return launch_blocks_plan(p, d_fr, d_curve, d_sw, cs); },
                     err, errlen);
  if (rc) return bail(rc);
  *out = c;
  return SA_OK;
}

extern "C" int sa_chain_create_approx(const sa_image* frames, int n, int iterations, int taps,
                                      int fp64, int use_tree, double diffusion, void* stream,
                                      sa_chain* out, char* err, size_t errlen) {
  if (!out) return fail(err, errlen, SA_EINVAL, "null output handle");
  *out = nullptr;
  hipStream_t s = (hipStream_t)stream;
  ApproxPlan p;
  int rc = plan_approx(frames, n, iterations, taps, fp64, use_tree, diffusion, p, err, errlen);
  if (rc) return rc;
  const int32_t* curve = nullptr;
  rc = get_curve(p.w, p.h, s, &curve, err, errlen);
  if (rc) return rc;
  const size_t b_curve = ((sizeof(int32_t) * (size_t)(p.w * p.h)) + 255) & ~(size_t)255;
  sa_chain c = new sa_chain_s();
  auto bail = [&](int code) {
    sa_chain_destroy(c);
    return code;
  };
  if (hipMalloc(&c->mem, p.scratch() + b_curve) != hipSuccess)
    return bail(fail(err, errlen, SA_ENOMEM, "chain: device allocation failed"));
  unsigned char* scratch = (unsigned char*)c->mem;
  int32_t* d_curve = (int32_t*)(scratch + p.scratch());
  hipError_t e = hipMemcpyAsync(scratch, p.fr.data(), sizeof(sa::BlockFrame) * p.fr.size(),
                                hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(scratch + p.b_fr, p.g_host(), p.g_bytes(), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_curve, curve, sizeof(int32_t) * (size_t)(p.w * p.h), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return bail(fail(err, errlen, SA_EDEVICE, "chain setup: %s", hipGetErrorString(e)));
  rc = capture_chain(c, [&](hipStream_t cs) { return launch_approx_plan(p, scratch, d_curve, cs); },
                     err, errlen);
  if (rc) return bail(rc);
  *out = c;
  return SA_OK;
}

extern "C" int sa_chain_run(sa_chain c, void* stream, char* err, size_t errlen) {
  if (!c || !c->exec) return fail(err, errlen, SA_EINVAL, "null chain");
  HIP_TRY(hipGraphLaunch(c->exec, (hipStream_t)stream));
  return SA_OK;
}

// ---- tuning -----------------------------------------------------------------------
extern "C" int sa_tune_set(const char* knob, int value, char* err, size_t errlen) {
  if (!knob) return fail(err, errlen, SA_EINVAL, "null knob");
  if (!std::strcmp(knob, "blocks.threads")) {
    if (value != 64 && value != 128 && value != 256 && value != 512)
      return fail(err, errlen, SA_EINVAL, "blocks.threads must be 64, 128, 256 or 512");
    g_knobs.blocks_threads = value;
  } else if (!std::strcmp(knob, "approx.per")) {
    if (value != 2 && value != 4 && value != 8)
      return fail(err, errlen, SA_EINVAL, "approx.per must be 2, 4 or 8");
    g_knobs.approx_per = value;
  } else if (!std::strcmp(knob, "seq.memo_bits")) {
    if (value != 0 && (value < 4 || value > 10))
      return fail(err, errlen, SA_EINVAL, "seq.memo_bits must be 0 (largest fit) or 4..10");
    g_knobs.seq_memo_bits = value;
  } else if (!std::strcmp(knob, "octree.cells")) {
    if (value != 0 && value != 1) return fail(err, errlen, SA_EINVAL, "octree.cells must be 0 or 1");
    sa::octree_set_cells(value);
  } else if (!std::strcmp(knob, "octree.exact_root")) {
    if (value != 0 && value != 1)
      return fail(err, errlen, SA_EINVAL, "octree.exact_root must be 0 or 1");
    sa::octree_set_exact_root(value);
  } else {
    return fail(err, errlen, SA_EINVAL, "unknown knob '%s'", knob);
  }
  return SA_OK;
}

extern "C" int sa_tune_get(const char* knob, int* value) {
  if (!knob || !value) return SA_EINVAL;
  if (!std::strcmp(knob, "blocks.threads")) *value = g_knobs.blocks_threads;
  else if (!std::strcmp(knob, "approx.per")) *value = g_knobs.approx_per;
  else if (!std::strcmp(knob, "seq.memo_bits")) *value = g_knobs.seq_memo_bits;
  else if (!std::strcmp(knob, "octree.cells")) *value = sa::octree_cells();
  else if (!std::strcmp(knob, "octree.exact_root")) *value = sa::octree_exact_root();
  else return SA_EINVAL;
  return SA_OK;
}

// ---- timing ---------------------------------------------------------------------
extern "C" int sa_timing_enable(int on) {
  std::lock_guard<std::mutex> g(g_t_mu);
  g_timing = on != 0;
  return SA_OK;
}

extern "C" int sa_timing_collect(const char* name, double* total_ms, int* launches, char* err,
                                 size_t errlen) {
  std::lock_guard<std::mutex> g(g_t_mu);
  double sum = 0.0;
  int n = 0;
  std::vector<Timed> keep;
  for (Timed& t : g_timed) {
    if (name && std::strcmp(name, t.name) != 0) {
      keep.push_back(t);
      continue;
    }
    HIP_TRY(hipEventSynchronize(t.b));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
    sum += ms;
    ++n;
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  g_timed.swap(keep);
  if (total_ms) *total_ms = sum;
  if (launches) *launches = n;
  return SA_OK;
}

extern "C" int sa_seq_stats(uint64_t* out, char* err, size_t errlen) {
  int rc = ensure_init(err, errlen);
  if (rc) return rc;
  uint32_t h[4] = {0, 0, 0, 0};
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(h, g_stats, 16, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemset(g_stats, 0, 16));
  for (int i = 0; i < 4; ++i) out[i] = h[i];
  return SA_OK;
}
