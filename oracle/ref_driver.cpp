// ref_driver.cpp -- a C entry layer over the reference's own CPU sources.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (oracle/Makefile, target `ref`) together
// with the reference's IM-free translation units src/rd_octree.cpp and
// src/rd_riemersma_cpu.cpp, read in place from the reference checkout, into
// oracle/_ref/libref.so.  This file is harness code written for this
// repository: it adds no stand-in for anything the reference lacks, it only
// exposes the reference's functions with plain-C signatures so ctypes can call
// them.  The output library is used to pin the CPU restatement (oracle.cpp)
// and, on the GPU host, as the "reference" CPU baseline.
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rd_octree.h"
#include "rd_plugin.h"
#include "rd_riemersma.h"
#include "rd_types.h"

namespace {

struct RefTree {
  rd::ColorTree tree;
  std::unique_ptr<rd::Palette> palette;
};

void fill_palette(RefTree* t) {
  t->palette.reset(new rd::Palette());
  t->palette->count = t->tree.color_count();
  t->palette->associate_alpha = t->tree.associate_alpha();
  for (int i = 0; i < t->palette->count; ++i) t->palette->entries[i] = t->tree.colormap()[i];
}

}  // namespace

extern "C" {

void* ref_tree_build(const float* rgba, int64_t w, int64_t h, int colors, int assoc,
                     int gray) {
  RefTree* t = new RefTree();
  t->tree.Build(reinterpret_cast<const rd::RgbaF*>(rgba), (std::size_t)w, (std::size_t)h,
                colors, assoc != 0, gray != 0);
  fill_palette(t);
  return t;
}

void ref_tree_free(void* p) { delete static_cast<RefTree*>(p); }
int ref_tree_color_count(void* p) { return static_cast<RefTree*>(p)->tree.color_count(); }
int64_t ref_tree_node_count(void* p) {
  return (int64_t) static_cast<RefTree*>(p)->tree.node_count();
}
int ref_tree_depth(void* p) { return static_cast<RefTree*>(p)->tree.depth(); }

void ref_tree_colormap(void* p, double* out) {
  RefTree* t = static_cast<RefTree*>(p);
  for (int i = 0; i < t->tree.color_count(); ++i) {
    const rd::PaletteEntry& e = t->tree.colormap()[i];
    out[4 * i + 0] = e.r;
    out[4 * i + 1] = e.g;
    out[4 * i + 2] = e.b;
    out[4 * i + 3] = e.a;
  }
}

void ref_tree_nodes(void* p, int32_t* child16, int32_t* parent, uint32_t* color_number,
                    uint64_t* number_unique, double* total4, double* qe, uint32_t* id,
                    uint32_t* level) {
  const std::vector<rd::QNode>& nodes = static_cast<RefTree*>(p)->tree.nodes();
  for (std::size_t i = 0; i < nodes.size(); ++i) {
    const rd::QNode& n = nodes[i];
    for (int c = 0; c < 16; ++c) child16[16 * i + c] = n.child[c];
    parent[i] = n.parent;
    color_number[i] = n.color_number;
    number_unique[i] = n.number_unique;
    for (int c = 0; c < 4; ++c) total4[4 * i + c] = n.total_color[c];
    qe[i] = n.quantize_error;
    id[i] = n.id;
    level[i] = n.level;
  }
}

void ref_walk_cpu(void* p, float* rgba, int64_t w, int64_t h, double diffusion,
                  int use_cache) {
  RefTree* t = static_cast<RefTree*>(p);
  rd::DitherParams params;
  params.diffusion = diffusion;
  params.use_cache = use_cache != 0;
  rd::RiemersmaWalkCpu(*t->palette, params, t->tree, (std::size_t)w, (std::size_t)h,
                       reinterpret_cast<rd::RgbaF*>(rgba), nullptr);
}

int ref_blocks_cpu(void* p, float* batch, int64_t w, int64_t h, int frames, int block,
                   double diffusion) {
  RefTree* t = static_cast<RefTree*>(p);
  rd::DitherParams params;
  params.diffusion = diffusion;
  std::string err;
  rd::RiemersmaBlocksCpu(*t->palette, params, t->tree, (std::size_t)w, (std::size_t)h,
                         reinterpret_cast<rd::RgbaF*>(batch), frames, block, nullptr, 0,
                         true, &err);
  return err.empty() ? 0 : -1;
}

int ref_blocks_cpu_u16(void* p, float* batch, int64_t w, int64_t h, int frames, int block,
                       double diffusion, uint16_t* out) {
  RefTree* t = static_cast<RefTree*>(p);
  rd::DitherParams params;
  params.diffusion = diffusion;
  std::string err;
  rd::RiemersmaBlocksCpu(*t->palette, params, t->tree, (std::size_t)w, (std::size_t)h,
                         reinterpret_cast<rd::RgbaF*>(batch), frames, block,
                         reinterpret_cast<unsigned char*>(out), (std::size_t)(w * h),
                         false, &err);
  return err.empty() ? 0 : -1;
}

// The fused gather and scatter of RiemersmaBlocksCpu: gather_layout 3 reads
// planar 4:4:4 from gather_raw (4 = interleaved uint16 of gather_channels),
// 0 reads `batch`; raw_out receives planar 4:4:4 (out_yuv444) or rgba64le.
int ref_blocks_cpu_fused(void* p, float* batch, int64_t w, int64_t h, int frames, int block,
                         double diffusion, uint8_t* raw_out, int out_yuv444,
                         const uint8_t* gather_raw, int64_t gather_frame_bytes,
                         int gather_layout, int gather_channels) {
  RefTree* t = static_cast<RefTree*>(p);
  rd::DitherParams params;
  params.diffusion = diffusion;
  std::string err;
  rd::RiemersmaBlocksCpu(*t->palette, params, t->tree, (std::size_t)w, (std::size_t)h,
                         reinterpret_cast<rd::RgbaF*>(batch), frames, block, raw_out,
                         (std::size_t)(w * h), out_yuv444 != 0, &err, gather_raw,
                         (std::size_t)gather_frame_bytes, gather_layout, gather_channels);
  return err.empty() ? 0 : -1;
}

// A registered plugin dither of the reference (FindDither(name)->run) in
// place on RgbaF frames; 0, or -1 with the plugin's error.
int ref_dither(const char* name, const double* ent, int count, float* rgba, int64_t w,
               int64_t h, int frames, double diffusion) {
  const rd::DitherAlgorithm* a = rd::FindDither(name);
  if (!a) return -2;
  static rd::Palette pal;  // 2 MiB: not on the stack
  pal.count = count;
  pal.associate_alpha = false;
  for (int i = 0; i < count; ++i) {
    pal.entries[i].r = ent[4 * i];
    pal.entries[i].g = ent[4 * i + 1];
    pal.entries[i].b = ent[4 * i + 2];
    pal.entries[i].a = ent[4 * i + 3];
  }
  rd::DitherJob job;
  job.palette = count > 0 ? &pal : nullptr;
  job.width = (std::size_t)w;
  job.height = (std::size_t)h;
  job.frames = frames;
  job.diffusion = diffusion;
  const rd::DitherResult r = a->run(job, reinterpret_cast<rd::RgbaF*>(rgba));
  return r.error.empty() ? 0 : -1;
}

int64_t ref_build_curve_index(int level, int64_t w, int64_t h, int32_t* out,
                              int32_t* owner) {
  std::vector<int> o, ow;
  const std::size_t n = rd::BuildCurveIndex(level, (std::size_t)w, (std::size_t)h, &o, &ow);
  for (std::size_t i = 0; i < n; ++i) out[i] = o[i];
  for (std::size_t i = 0; i < ow.size(); ++i) owner[i] = ow[i];
  return (int64_t)n;
}

void ref_hilbert_point(int64_t n, int level, int* x, int* y) {
  rd::HilbertPoint(n, level, x, y);
}

int64_t ref_curve_mismatches(int level, int transform) {
  return (int64_t)rd::CurveTransformMismatches(level, transform);
}

int ref_curve_level(int64_t w, int64_t h) {
  return rd::ComputeCurveLevel((std::size_t)w, (std::size_t)h);
}

void ref_error_weights(double* out) {
  double w[rd::kErrorQueueLength];
  rd::build_error_weights(w);
  std::memcpy(out, w, sizeof(w));
}

unsigned ref_scale_char(float q) { return rd::scale_quantum_to_char(q); }

// CPU baseline over many images: one image per worker at a time, largest
// first, as the reference runs RiemersmaWalkCpu (one single-threaded walk per
// image).
int ref_walk_cpu_many(void** trees, float** images, const int64_t* ws, const int64_t* hs,
                      int n, double diffusion, int threads) {
  std::vector<int> ord((std::size_t)n);
  for (int i = 0; i < n; ++i) ord[(std::size_t)i] = i;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int a, int b) { return ws[a] * hs[a] > ws[b] * hs[b]; });
  std::mutex mu;
  int next = 0;
  std::vector<std::thread> pool;
  if (threads < 1) threads = 1;
  for (int k = 0; k < threads; ++k) {
    pool.emplace_back([&]() {
      for (;;) {
        int j;
        {
          std::lock_guard<std::mutex> g(mu);
          if (next >= n) return;
          j = ord[(std::size_t)next++];
        }
        ref_walk_cpu(trees[j], images[j], ws[j], hs[j], diffusion, 1);
      }
    });
  }
  for (auto& th : pool) th.join();
  return 0;
}

}  // extern "C"
