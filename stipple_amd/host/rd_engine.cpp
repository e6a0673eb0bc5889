// rd_engine.cpp -- the reference's engine API (include/rd_riemersma.h,
// include/rd_octree.h, include/rd_plugin.h) over the C-ABI of
// include/stipple_amd.h, as this repository's drop-in library (librd_amd.so).
//
// On top of the GPU engine entry points it shares with the reference-side
// adapter (rd_gpu_engines.inc), it holds what only a stand-alone build needs:
// ColorTree built on the device, the curve functions, RiemersmaWalkBatch and
// the plugin registry with the device screen dithers.
//
// Host code only moves memory and arguments: every result is computed by the
// HIP kernels behind the C-ABI.  Engines keep the reference's conventions:
// an error string, empty on success; synchronous on return; host pixels
// dithered in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rd_im.h"
#include "rd_octree.h"
#include "rd_plugin.h"
#include "rd_riemersma.h"
#include "stipple_amd.h"

namespace rd {
namespace {
// The tree's own device search structure when `pal` is its colormap.
sa_palette resident_palette(const Palette& pal, const ColorTree& tree) {
  if (tree.device_palette() && tree.color_count() == pal.count &&
      std::memcmp(tree.colormap(), pal.entries, sizeof(PaletteEntry) * pal.count) == 0)
    return tree.device_palette();
  return nullptr;
}
}  // namespace

// RiemersmaWalkCuda, RiemersmaBlocksCuda, RiemersmaApproxCuda, ApproxInverseFilter,
// CudaAvailable, CudaMaxFrames, CudaAllocPinned/FreePinned: shared with the
// adapter compiled inside the reference's build (rd_hip_engine.cpp).
#include "rd_gpu_engines.inc"

// ---- ColorTree ----------------------------------------------------------------

ColorTree::ColorTree() = default;
ColorTree::~ColorTree() { Reset(); }
ColorTree::ColorTree(ColorTree&& o) noexcept { *this = std::move(o); }
ColorTree& ColorTree::operator=(ColorTree&& o) noexcept {
  if (this != &o) {
    Reset();
    nodes_ = std::move(o.nodes_);
    colormap_ = std::move(o.colormap_);
    color_count_ = o.color_count_;
    depth_ = o.depth_;
    associate_alpha_ = o.associate_alpha_;
    device_ = o.device_;
    error_ = std::move(o.error_);
    o.device_ = nullptr;
    o.color_count_ = 0;
  }
  return *this;
}

void ColorTree::Reset() {
  if (device_) sa_palette_destroy(device_);
  device_ = nullptr;
  nodes_.clear();
  colormap_.clear();
  color_count_ = 0;
  error_.clear();
}

void ColorTree::Build(const RgbaF* pixels, std::size_t width, std::size_t height, int max_colors,
                      bool associate_alpha, bool grayscale) {
  Reset();
  associate_alpha_ = associate_alpha;
  if (!ensure_device(&error_)) return;
  const size_t bytes = sizeof(RgbaF) * width * height;
  DevBuf src;
  if (!src.reserve(bytes)) {
    error_ = "device allocation failed";
    return;
  }
  hipError_t e = hipMemcpy(src.p, pixels, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    error_ = hip_fail("upload", e);
    return;
  }
  char err[kErr] = {0};
  sa_palette p = nullptr;
  int rc = sa_palette_build(src.p, SA_PIX_RGBAF32, (int64_t)width, (int64_t)height, max_colors,
                            associate_alpha ? 1 : 0, grayscale ? 1 : 0, 1, nullptr, &p, err,
                            sizeof(err));
  if (rc != SA_OK) {
    error_ = prefixed("ColorTree::Build", err);
    return;
  }
  device_ = p;
  int count = 0, assoc = 0, depth = 0, exact = 0;
  int64_t n = 0;
  sa_palette_info(p, &count, &assoc, &depth, &n, &exact);
  std::vector<int32_t> child(16 * (size_t)n), parent(n);
  std::vector<uint32_t> cn(n), id(n), level(n);
  std::vector<uint64_t> nu(n);
  std::vector<double> total(4 * (size_t)n), qe(n), entries(4 * (size_t)count);
  rc = sa_palette_tree(p, child.data(), parent.data(), cn.data(), nu.data(), total.data(),
                       qe.data(), id.data(), level.data(), err, sizeof(err));
  if (rc == SA_OK) rc = sa_palette_entries(p, entries.data(), err, sizeof(err));
  if (rc != SA_OK) {
    error_ = prefixed("ColorTree::Build readback", err);
    return;
  }
  nodes_.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    QNode& q = nodes_[i];
    for (int k = 0; k < 16; ++k) q.child[k] = child[16 * i + k];
    q.parent = parent[i];
    for (int c = 0; c < 4; ++c) q.total_color[c] = total[4 * i + c];
    q.quantize_error = qe[i];
    q.number_unique = nu[i];
    q.color_number = cn[i];
    q.id = id[i];
    q.level = level[i];
  }
  colormap_.resize(count);
  for (int i = 0; i < count; ++i)
    colormap_[i] = PaletteEntry{entries[4 * i], entries[4 * i + 1], entries[4 * i + 2],
                                entries[4 * i + 3]};
  color_count_ = count;
  depth_ = depth;
}

// ---- curve ---------------------------------------------------------------------

void HilbertPoint(std::int64_t n, int level, int* out_x, int* out_y) {
  int64_t x = 0, y = 0;
  for (int b = 0; b < level; ++b) {
    const int64_t s = (int64_t)1 << b;
    const int64_t rx = 1 & (n >> 1);
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
    n >>= 2;
  }
  *out_x = (int)x;
  *out_y = (int)y;
}

std::size_t BuildCurveIndex(int level, std::size_t width, std::size_t height,
                            std::vector<int>* out, std::vector<int>* owner) {
  if (out) out->clear();
  if (owner) owner->clear();
  const size_t n = width * height;
  if (n == 0 || level != ComputeCurveLevel(width, height)) return 0;
  if (level == 0) {  // a 1x1 image: the block curve is empty (src/rd_riemersma_cpu.cpp:275)
    if (owner) owner->assign(n, -1);
    return 0;
  }
  if (!ensure_device(nullptr)) return 0;
  DevBuf d;
  if (!d.reserve(sizeof(int32_t) * n)) return 0;
  char err[kErr] = {0};
  if (sa_curve_order((int64_t)width, (int64_t)height, (int32_t*)d.p, nullptr, err,
                     sizeof(err)) != SA_OK)
    return 0;
  std::vector<int> seq(n);
  if (hipMemcpy(seq.data(), d.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
    return 0;
  // The sequential order is HilbertPoint(k) for k in [0, 4^L) in bounds; its last
  // point (2^L - 1, 0) is in bounds exactly when width == 2^L.  The block curve
  // stops before it and ends with the origin (src/rd_riemersma_cpu.cpp:269-356).
  const bool drop_last = width == ((size_t)1 << level);
  std::vector<int> order(seq.begin(), seq.end() - (drop_last ? 1 : 0));
  order.push_back(0);
  if (owner) {
    owner->assign(n, -1);
    for (size_t i = 0; i < order.size(); ++i) (*owner)[order[i]] = (int)i;
  }
  const size_t len = order.size();
  if (out) *out = std::move(order);
  return len;
}

namespace {
struct CurvePair {
  size_t w = 0, h = 0;
  bool valid = false;
  std::vector<int> order, owner;
};
CurvePair& curve_cache(size_t w, size_t h) {
  thread_local CurvePair c;
  if (c.w != w || c.h != h || !c.valid) {
    c.w = w;
    c.h = h;
    BuildCurveIndex(ComputeCurveLevel(w, h), w, h, &c.order, &c.owner);
    c.valid = !c.owner.empty();
  }
  return c;
}
}  // namespace

const std::vector<int>* BlockCurveOrder(std::size_t width, std::size_t height) {
  CurvePair& c = curve_cache(width, height);
  return c.valid ? &c.order : nullptr;
}

const std::vector<int>* BlockCurveOwner(std::size_t width, std::size_t height) {
  CurvePair& c = curve_cache(width, height);
  return c.valid ? &c.owner : nullptr;
}

// ---- engines -------------------------------------------------------------------------

std::string RiemersmaWalkBatch(const std::vector<WalkItem>& items, const DitherParams& params,
                               std::string* device_name) {
  std::string why;
  if (!ensure_device(&why)) return why;
  fill_name(device_name);
  std::vector<PaletteRef> pals(items.size());
  std::vector<size_t> off(items.size() + 1, 0);
  for (size_t i = 0; i < items.size(); ++i) {
    const WalkItem& it = items[i];
    if (!it.tree || !it.pixels) return "walk item without pixels or tree";
    if (it.tree->color_count() <= 0) return "walk item's tree has no colormap";
    if (!it.tree->device_palette()) return "walk item's tree was not built on the device";
    pals[i].p = it.tree->device_palette();
    off[i + 1] = off[i] + sizeof(RgbaF) * it.width * it.height;
  }
  if (off.back() == 0) return "";
  DevBuf d;
  if (!d.reserve(off.back())) return "device allocation failed";
  std::vector<sa_image> im;
  for (size_t i = 0; i < items.size(); ++i) {
    const WalkItem& it = items[i];
    if (it.width * it.height == 0) continue;
    unsigned char* base = (unsigned char*)d.p + off[i];
    hipError_t he = hipMemcpy(base, it.pixels, off[i + 1] - off[i], hipMemcpyHostToDevice);
    if (he != hipSuccess) return hip_fail("upload", he);
    im.push_back(sa_image{base, base, (int64_t)it.width, (int64_t)it.height, SA_PIX_RGBAF32,
                          SA_PIX_RGBAF32, pals[i].p});
  }
  sa_seq_options o{params.diffusion, params.use_cache ? 1 : 0, 0, 0, 0};
  char err[kErr] = {0};
  if (sa_walk_seq_batch(im.data(), (int)im.size(), &o, nullptr, err, sizeof(err)) != SA_OK)
    return prefixed("RiemersmaWalkBatch", err);
  for (size_t i = 0; i < items.size(); ++i) {
    if (off[i + 1] == off[i]) continue;
    hipError_t he = hipMemcpy(items[i].pixels, (unsigned char*)d.p + off[i], off[i + 1] - off[i],
                              hipMemcpyDeviceToHost);
    if (he != hipSuccess) return hip_fail("download", he);
  }
  return "";
}

// ---- palette files and the pixel-buffer palette ------------------------------------------

bool ImWritePalette(const Palette& palette, const std::string& path, std::string* error) {
  char err[kErr] = {0};
  if (sa_palette_write_txt(&palette.entries[0].r, palette.count, path.c_str(), err,
                           sizeof(err)) != SA_OK) {
    if (error) *error = err;
    return false;
  }
  return true;
}

bool ImReadPalette(const std::string& path, int expect_colors, Palette* palette,
                   ColorTree* tree, std::string* error) {
  char err[kErr] = {0};
  int count = 0, assoc = 0;
  if (sa_palette_read_txt(path.c_str(), expect_colors, &palette->entries[0].r, &count, &assoc,
                          err, sizeof(err)) != SA_OK) {
    if (error) *error = err;
    return false;
  }
  palette->count = count;
  palette->associate_alpha = assoc != 0;
  // The tree of the palette itself: a count x 1 strip at `count` colours.
  std::vector<RgbaF> strip((size_t)count);
  bool gray = true;
  for (int i = 0; i < count; ++i) {
    const PaletteEntry& e = palette->entries[i];
    strip[(size_t)i] = RgbaF{(float)e.r, (float)e.g, (float)e.b, (float)e.a};
    if (e.r != e.g || e.g != e.b) gray = false;
  }
  tree->Build(strip.data(), (size_t)count, 1, count, false, gray);
  if (!tree->last_error().empty()) {
    if (error) *error = path + ": " + tree->last_error();
    return false;
  }
  if (tree->color_count() != count) {
    if (error)
      *error = path + ": the tree reduced " + std::to_string(count) + " colours to " +
               std::to_string(tree->color_count()) +
               "; the palette holds duplicate or unreachable entries";
    return false;
  }
  for (int i = 0; i < count; ++i) palette->entries[i] = tree->colormap()[i];
  return true;
}

bool ImBuildPaletteFromPixels(const RgbaF* pixels, std::size_t width, std::size_t height,
                              int colors, Palette* palette, ColorTree* tree, std::string* error) {
  bool gray = true;
  for (size_t i = 0; i < width * height && gray; ++i)
    gray = pixels[i].r == pixels[i].g && pixels[i].g == pixels[i].b;
  tree->Build(pixels, width, height, colors, false, gray);
  if (!tree->last_error().empty() || tree->color_count() <= 0) {
    if (error) *error = tree->last_error().empty() ? "empty colormap" : tree->last_error();
    return false;
  }
  // HarvestColormap: the colormap with exact (r, g, b) duplicates dropped,
  // first occurrence kept.
  palette->associate_alpha = false;
  int out = 0;
  for (int i = 0; i < tree->color_count(); ++i) {
    const PaletteEntry& e = tree->colormap()[i];
    bool dup = false;
    for (int k = 0; k < out && !dup; ++k)
      dup = palette->entries[k].r == e.r && palette->entries[k].g == e.g &&
            palette->entries[k].b == e.b;
    if (!dup) palette->entries[out++] = e;
  }
  palette->count = out;
  return true;
}

// ---- plugin dithers on the device ---------------------------------------------------------
// The reference's error-diffusion dithers (src/rd_dithers.cc:362-369) and its
// four threshold-screen dithers (src/rd_dithers_ordered.cc:488-502,
// src/rd_dithers.cc:370-371) registered under their own names: linking this
// library gives a caller the same --dither names, run by sa_dither_diffuse and
// sa_dither_screen.
namespace {

// One registered device dither: `launch` is sa_dither_screen or sa_dither_diffuse
// with its kind.
using DeviceDither = int (*)(const sa_image*, int, int, double, void*, char*, size_t);

DitherResult device_run(DitherJob& job, RgbaF* px, DeviceDither launch, int kind,
                        const char* who) {
  DitherResult res;
  if (!job.palette || job.palette->count <= 0) {  // Prepare, src/rd_dither_common.h:108-120
    res.error = std::string(who) + ": no palette";
    return res;
  }
  if (job.width == 0 || job.height == 0) {
    res.error = std::string(who) + ": empty frame";
    return res;
  }
  if (!ensure_device(&res.error)) return res;
  // A screen reads only the entries: a one-leaf tree carries them.
  int32_t child[32];
  for (int k = 0; k < 32; ++k) child[k] = -1;
  child[0] = 1;
  const int32_t parent[2] = {-1, 0};
  const uint32_t cn[2] = {0, 0};
  const uint64_t nu[2] = {1, 1};
  char err[kErr] = {0};
  sa_palette pal = nullptr;
  if (sa_palette_create(&job.palette->entries[0].r, job.palette->count, 0, child, parent, cn, nu, 2,
                        &pal, err, sizeof(err)) != SA_OK) {
    res.error = prefixed(who, err);
    return res;
  }
  const int frames = std::max(job.frames, 1);
  const size_t px_n = job.width * job.height, bytes = sizeof(RgbaF) * px_n * (size_t)frames;
  DevBuf buf;
  std::vector<sa_image> fr((size_t)frames);
  hipError_t he = buf.reserve(bytes) ? hipMemcpy(buf.p, px, bytes, hipMemcpyHostToDevice)
                                     : hipErrorOutOfMemory;
  if (he == hipSuccess) {
    for (int f = 0; f < frames; ++f) {
      unsigned char* d = (unsigned char*)buf.p + sizeof(RgbaF) * px_n * (size_t)f;
      fr[(size_t)f] = sa_image{d, d, (int64_t)job.width, (int64_t)job.height, SA_PIX_RGBAF32,
                               SA_PIX_RGBAF32, pal};
    }
    if (launch(fr.data(), frames, kind, job.diffusion, nullptr, err, sizeof(err)) != SA_OK)
      res.error = prefixed(who, err);
    else
      he = hipMemcpy(px, buf.p, bytes, hipMemcpyDeviceToHost);
  }
  if (he != hipSuccess && res.error.empty()) res.error = hip_fail(who, he);
  sa_palette_destroy(pal);
  return res;
}

DitherResult run_bayer(DitherJob& j, RgbaF* p) {
  return device_run(j, p, sa_dither_screen, SA_SCREEN_BAYER8, "bayer-ordered");
}
DitherResult run_void64(DitherJob& j, RgbaF* p) {
  return device_run(j, p, sa_dither_screen, SA_SCREEN_VOID64, "void-and-cluster");
}
DitherResult run_void32(DitherJob& j, RgbaF* p) {
  return device_run(j, p, sa_dither_screen, SA_SCREEN_VOID32, "void-and-cluster-fast");
}
DitherResult run_cluster(DitherJob& j, RgbaF* p) {
  return device_run(j, p, sa_dither_screen, SA_SCREEN_CLUSTERED8, "clustered-dot");
}
DitherResult run_floyd(DitherJob& j, RgbaF* p) {
  return device_run(j, p, sa_dither_diffuse, SA_DIFFUSE_FLOYD, "floyd-steinberg");
}
DitherResult run_atkinson(DitherJob& j, RgbaF* p) {
  return device_run(j, p, sa_dither_diffuse, SA_DIFFUSE_ATKINSON, "atkinson");
}
DitherResult run_jarvis(DitherJob& j, RgbaF* p) {
  return device_run(j, p, sa_dither_diffuse, SA_DIFFUSE_JARVIS, "jarvis");
}

}  // namespace

// ---- plugin registry ---------------------------------------------------------------------

namespace {
struct Registry {
  std::mutex mu;
  std::vector<DitherAlgorithm> algos;
};
Registry& reg() {
  static Registry r;
  return r;
}
}  // namespace

bool RegisterDither(const DitherAlgorithm& algo) {
  if (!algo.name || !algo.name[0] || !algo.run) return false;
  Registry& r = reg();
  std::lock_guard<std::mutex> g(r.mu);
  for (const DitherAlgorithm& a : r.algos)
    if (std::strcmp(a.name, algo.name) == 0) return false;  // duplicates are refused
  r.algos.push_back(algo);
  return true;
}

const DitherAlgorithm* FindDither(const char* name) {
  if (!name) return nullptr;
  Registry& r = reg();
  std::lock_guard<std::mutex> g(r.mu);
  for (const DitherAlgorithm& a : r.algos)
    if (std::strcmp(a.name, name) == 0) return &a;
  return nullptr;
}

const std::vector<DitherAlgorithm>& RegisteredDithers() { return reg().algos; }

namespace {
const bool kScreensRegistered = [] {
  RegisterDither({"floyd-steinberg", "error diffusion, 4 neighbours (7/16 3/16 5/16 1/16)",
                  run_floyd});
  RegisterDither({"atkinson", "error diffusion, 6 neighbours at 1/8; discards 1/4 of the error",
                  run_atkinson});
  RegisterDither({"jarvis", "Jarvis/Judice/Ninke, 12 neighbours, 48-denominator weights",
                  run_jarvis});
  RegisterDither({"bayer-ordered", "ordered 8x8 Bayer, no error diffusion", run_bayer});
  RegisterDither({"void-and-cluster", "blue-noise thresholds, 64x64 matrix", run_void64});
  RegisterDither({"void-and-cluster-fast", "blue-noise thresholds, 32x32 matrix (quicker)",
                  run_void32});
  RegisterDither({"clustered-dot", "8x8 halftone screen, no diffusion; shows the 8x8 period",
                  run_cluster});
  return true;
}();
}  // namespace

std::string DitherListText() {
  Registry& r = reg();
  std::lock_guard<std::mutex> g(r.mu);
  std::string out;
  for (size_t i = 0; i < r.algos.size(); ++i) {
    const DitherAlgorithm& a = r.algos[i];
    if (i) out += "\n";
    const size_t len = std::strlen(a.name);
    out += "  ";
    out += a.name;
    out += std::string(len < 16 ? 16 - len : 1, ' ');
    out += a.description ? a.description : "";
  }
  return out;
}

}  // namespace rd
