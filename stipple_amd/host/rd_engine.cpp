// rd_engine.cpp -- the reference's engine API (include/rd_riemersma.h,
// include/rd_octree.h, include/rd_plugin.h) over the C-ABI of
// include/stipple_amd.h.
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

#include "rd_octree.h"
#include "rd_plugin.h"
#include "rd_riemersma.h"
#include "stipple_amd.h"

namespace rd {
namespace {

constexpr size_t kErr = 512;

std::string prefixed(const char* what, const char* err) {
  return std::string(what) + ": " + (err[0] ? err : "failed");
}

bool ensure_device(std::string* why) {
  static std::once_flag once;
  static int rc = SA_EDEVICE;
  static char msg[kErr] = {0};
  std::call_once(once, [] {
    if (!sa_available()) {
      std::snprintf(msg, sizeof(msg), "no MI355X (gfx950) device available");
      return;
    }
    rc = sa_init(0, msg, sizeof(msg));
  });
  if (rc != SA_OK && why) *why = msg[0] ? msg : "device initialisation failed";
  return rc == SA_OK;
}

void fill_name(std::string* device_name) {
  if (!device_name) return;
  char b[256] = {0};
  sa_device_name(b, sizeof(b));
  *device_name = b;
}

// Device memory owned by one call (or one state slot).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  bool reserve(size_t n) {
    if (n <= bytes) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, std::max<size_t>(n, 16)) != hipSuccess) return false;
    bytes = n;
    return true;
  }
};

std::string hip_fail(const char* what, hipError_t e) {
  return std::string(what) + ": " + hipGetErrorString(e);
}

// The device search structure for (palette, tree): the tree's own when the
// palette is its colormap, else uploaded from the tree's node arrays.
struct PaletteRef {
  sa_palette p = nullptr;
  bool owned = false;
  ~PaletteRef() {
    if (owned && p) sa_palette_destroy(p);
  }
};

std::string palette_for(const Palette& pal, const ColorTree& tree, PaletteRef* out) {
  if (pal.count <= 0 || pal.count > kMaxColormapSize) return "palette is empty or too large";
  if (tree.node_count() == 0) return "colour tree is empty (ColorTree::Build not run)";
  if (tree.device_palette() && tree.color_count() == pal.count &&
      std::memcmp(tree.colormap(), pal.entries, sizeof(PaletteEntry) * pal.count) == 0) {
    out->p = tree.device_palette();
    out->owned = false;
    return "";
  }
  const std::vector<QNode>& nodes = tree.nodes();
  const size_t n = nodes.size();
  std::vector<int32_t> child(16 * n), parent(n);
  std::vector<uint32_t> cn(n);
  std::vector<uint64_t> nu(n);
  for (size_t i = 0; i < n; ++i) {
    for (int k = 0; k < 16; ++k) child[16 * i + k] = nodes[i].child[k];
    parent[i] = nodes[i].parent;
    cn[i] = nodes[i].color_number;
    nu[i] = nodes[i].number_unique;
  }
  char err[kErr] = {0};
  const int rc = sa_palette_create(&pal.entries[0].r, pal.count, tree.associate_alpha() ? 1 : 0,
                                   child.data(), parent.data(), cn.data(), nu.data(),
                                   (int64_t)n, &out->p, err, sizeof(err));
  if (rc != SA_OK) return prefixed("palette upload", err);
  out->owned = true;
  return "";
}

}  // namespace

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

// ---- runtime ---------------------------------------------------------------------

bool CudaAvailable() { return ensure_device(nullptr); }

int CudaMaxFrames(std::size_t width, std::size_t height, std::size_t vram_budget) {
  if (!ensure_device(nullptr)) return 0;
  const size_t px = width * height;
  if (px == 0) return 0;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 1;
  const size_t budget = vram_budget ? std::min(vram_budget, free_b) : free_b;
  // Per frame: the image (16 B/px), the staged inputs (32 B/px), the index
  // per visit (2 B/px) and the memo table (2^24 entries of 2 B with alpha).
  const size_t per_frame = px * (16 + 32 + 2) + ((size_t)2 << 24);
  const size_t fixed = px * 4 + ((size_t)64 << 20);  // curve + slack
  if (budget <= fixed) return 0;
  return (int)std::min<size_t>((budget - fixed) / per_frame, 1 << 20);
}

void* CudaAllocPinned(std::size_t bytes) {
  if (!ensure_device(nullptr)) return nullptr;
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}

void CudaFreePinned(void* p) {
  if (p) (void)hipHostFree(p);
}

// ---- engines -------------------------------------------------------------------------

std::string RiemersmaWalkCuda(const Palette& palette, const DitherParams& params,
                              const ColorTree& tree, std::size_t width, std::size_t height,
                              RgbaF* pixels, int frames, std::string* device_name) {
  std::string why;
  if (!ensure_device(&why)) return why;
  if (frames < 1) frames = 1;
  fill_name(device_name);
  const size_t px = width * height;
  if (px == 0) return "";
  PaletteRef pr;
  std::string e = palette_for(palette, tree, &pr);
  if (!e.empty()) return e;
  // `pixels` holds ONE frame: the reference uploads frame 0 as the shared
  // source of its `frames` walks and downloads one frame
  // (src/rd_riemersma_cuda.cu:688, :747).  Those walks read the same source
  // with the same palette and state, so they are identical, and one walk
  // yields the frame it returns.  Distinct images walk concurrently through
  // RiemersmaWalkBatch.
  const size_t bytes = sizeof(RgbaF) * px;
  DevBuf d;
  if (!d.reserve(bytes)) return "device allocation failed";
  hipError_t he = hipMemcpy(d.p, pixels, bytes, hipMemcpyHostToDevice);
  if (he != hipSuccess) return hip_fail("upload", he);
  const sa_image item{d.p, d.p, (int64_t)width, (int64_t)height, SA_PIX_RGBAF32,
                      SA_PIX_RGBAF32, pr.p};
  sa_seq_options o{params.diffusion, params.use_cache ? 1 : 0, 0, 0, 0};
  char err[kErr] = {0};
  if (sa_walk_seq_batch(&item, 1, &o, nullptr, err, sizeof(err)) != SA_OK)
    return prefixed("RiemersmaWalkCuda", err);
  he = hipMemcpy(pixels, d.p, bytes, hipMemcpyDeviceToHost);
  if (he != hipSuccess) return hip_fail("download", he);
  return "";
}

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

std::vector<double> ApproxInverseFilter(int taps, double diffusion) {
  std::vector<double> g;
  if (taps < 1 || taps > 4096) return g;
  g.resize((size_t)taps + 1);
  char err[kErr] = {0};
  if (sa_inverse_filter(taps, diffusion, g.data(), err, sizeof(err)) != SA_OK) g.clear();
  return g;
}

std::string RiemersmaApproxCuda(const Palette& palette, const DitherParams& params,
                                const ColorTree& tree, std::size_t width, std::size_t height,
                                RgbaF* pixels, const ApproxOptions& options,
                                std::string* device_name) {
  std::string why;
  if (!ensure_device(&why)) return why;
  fill_name(device_name);
  const size_t px = width * height;
  if (px == 0) return "";
  PaletteRef pr;
  std::string e = palette_for(palette, tree, &pr);
  if (!e.empty()) return e;
  const size_t bytes = sizeof(RgbaF) * px;
  DevBuf src, dst;
  if (!src.reserve(bytes) || !dst.reserve(bytes)) return "device allocation failed";
  hipError_t he = hipMemcpy(src.p, pixels, bytes, hipMemcpyHostToDevice);
  if (he == hipSuccess) he = hipMemcpy(dst.p, src.p, bytes, hipMemcpyDeviceToDevice);
  if (he != hipSuccess) return hip_fail("upload", he);
  sa_image im{src.p, dst.p, (int64_t)width, (int64_t)height, SA_PIX_RGBAF32, SA_PIX_RGBAF32,
              pr.p};
  char err[kErr] = {0};
  if (sa_walk_approx(&im, 1, options.iterations, options.taps, options.fp64 ? 1 : 0,
                     options.use_tree ? 1 : 0, params.diffusion, nullptr, err,
                     sizeof(err)) != SA_OK)
    return prefixed("RiemersmaApproxCuda", err);
  he = hipMemcpy(pixels, dst.p, bytes, hipMemcpyDeviceToHost);
  if (he != hipSuccess) return hip_fail("download", he);
  return "";
}

namespace {
struct BlockSlot {
  std::mutex mu;
  DevBuf src, dst;
};
BlockSlot g_slots[kGpuStateSlots];
}  // namespace

std::string RiemersmaBlocksCuda(const Palette& palette, const DitherParams& params,
                                const ColorTree& tree, std::size_t width, std::size_t height,
                                RgbaF* batch, const BlockOptions& options,
                                std::string* device_name, std::uint16_t* out_u16, int state_slot,
                                const std::uint16_t* in_u16) {
  std::string why;
  if (!ensure_device(&why)) return why;
  fill_name(device_name);
  if (options.frames < 1) return "RiemersmaBlocksCuda: frames must be >= 1";
  if (state_slot < 0 || state_slot >= kGpuStateSlots)
    return "RiemersmaBlocksCuda: state_slot out of range";
  // The planar modes pass 8-bit planes through the uint16 pointers, as the
  // reference does (src/rd_blocks_cuda.cu:1066-1081, :1532-1545).
  const bool in_planar = options.upload_u16 && options.in_mode != BlockOptions::InMode::Interleaved16;
  if (options.upload_u16 && !in_u16) return "RiemersmaBlocksCuda: upload_u16 needs in_u16";
  if (options.upload_u16 && !in_planar && options.in_channels != 3 && options.in_channels != 4)
    return "RiemersmaBlocksCuda: upload_u16 needs in_u16 with 3 or 4 channels";
  if (options.emit_u16 && !out_u16) return "RiemersmaBlocksCuda: emit_u16 needs out_u16";
  if (!options.upload_u16 && !batch) return "RiemersmaBlocksCuda: null batch";
  const size_t px = width * height;
  if (px == 0) return "";
  PaletteRef pr;
  std::string e = palette_for(palette, tree, &pr);
  if (!e.empty()) return e;
  BlockSlot& slot = g_slots[state_slot];
  std::lock_guard<std::mutex> lock(slot.mu);
  const int frames = options.frames;
  // Bytes per frame of the source and the result.
  int in_fmt = SA_PIX_RGBAF32;
  size_t in_frame = sizeof(RgbaF) * px;
  if (options.upload_u16) {
    switch (options.in_mode) {
      case BlockOptions::InMode::PlanarYuv444:
      case BlockOptions::InMode::PlanarYuv444Prepass:  // the same values, widened in one pass
        in_fmt = SA_PIX_YUV444P8;
        in_frame = 3 * px;
        break;
      case BlockOptions::InMode::PlanarYuv420:
        in_fmt = SA_PIX_YUV420P8;
        in_frame = px + 2 * (width / 2) * (height / 2);
        break;
      default:
        in_fmt = options.in_channels == 3 ? SA_PIX_RGB16 : SA_PIX_RGBA16;
        in_frame = 2 * (size_t)options.in_channels * px;
        break;
    }
  }
  int out_fmt = SA_PIX_RGBAF32;
  size_t out_frame = sizeof(RgbaF) * px;
  if (options.emit_u16) {
    out_fmt = options.emit_yuv444 ? SA_PIX_YUV444P8 : SA_PIX_RGBA16;
    out_frame = options.emit_yuv444 ? 3 * px : 8 * px;
  }
  if (!slot.src.reserve(in_frame * frames) || !slot.dst.reserve(out_frame * frames))
    return "device allocation failed";
  const void* host_in = options.upload_u16 ? (const void*)in_u16 : (const void*)batch;
  hipError_t he = hipMemcpy(slot.src.p, host_in, in_frame * frames, hipMemcpyHostToDevice);
  if (he != hipSuccess) return hip_fail("upload", he);
  std::vector<sa_image> fr(frames);
  for (int f = 0; f < frames; ++f)
    fr[f] = sa_image{(unsigned char*)slot.src.p + in_frame * f,
                     (unsigned char*)slot.dst.p + out_frame * f, (int64_t)width,
                     (int64_t)height, in_fmt, out_fmt, pr.p};
  char err[kErr] = {0};
  if (sa_walk_blocks(fr.data(), frames, options.block, params.diffusion, nullptr, err,
                     sizeof(err)) != SA_OK)
    return prefixed("RiemersmaBlocksCuda", err);
  void* host_out = options.emit_u16 ? (void*)out_u16 : (void*)batch;
  if (!host_out) return "RiemersmaBlocksCuda: nowhere to write the result";
  he = hipMemcpy(host_out, slot.dst.p, out_frame * frames, hipMemcpyDeviceToHost);
  if (he != hipSuccess) return hip_fail("download", he);
  return "";
}

// ---- threshold-screen plugins on the device ----------------------------------------------
// The reference's four threshold-screen dithers (src/rd_dithers_ordered.cc:488-502,
// src/rd_dithers.cc:370-371) registered under their own names: linking this
// library gives a caller the same --dither names, run by sa_dither_screen.
namespace {

DitherResult screen_run(DitherJob& job, RgbaF* px, int screen, const char* who) {
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
    if (sa_dither_screen(fr.data(), frames, screen, job.diffusion, nullptr, err, sizeof(err)) !=
        SA_OK)
      res.error = prefixed(who, err);
    else
      he = hipMemcpy(px, buf.p, bytes, hipMemcpyDeviceToHost);
  }
  if (he != hipSuccess && res.error.empty()) res.error = hip_fail(who, he);
  sa_palette_destroy(pal);
  return res;
}

DitherResult run_bayer(DitherJob& j, RgbaF* p) {
  return screen_run(j, p, SA_SCREEN_BAYER8, "bayer-ordered");
}
DitherResult run_void64(DitherJob& j, RgbaF* p) {
  return screen_run(j, p, SA_SCREEN_VOID64, "void-and-cluster");
}
DitherResult run_void32(DitherJob& j, RgbaF* p) {
  return screen_run(j, p, SA_SCREEN_VOID32, "void-and-cluster-fast");
}
DitherResult run_cluster(DitherJob& j, RgbaF* p) {
  return screen_run(j, p, SA_SCREEN_CLUSTERED8, "clustered-dot");
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
