// dropin_test.cpp -- the C++ drop-in API (include/rd_riemersma.h, rd_octree.h,
// rd_plugin.h) against the oracle, on the GPU.
//
// Written the way a caller of the reference's engine API would use it: build a
// ColorTree, hand Palette + tree + host RgbaF pixels to the engines, compare
// with the CPU restatement (oracle/oracle.h) on the same inputs.  Prints one
// line per check and exits nonzero on any failure.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../oracle/oracle.h"
#include "rd_im.h"
#include "rd_plugin.h"
#include "rd_riemersma.h"

namespace {

int g_fail = 0;

void check(bool ok, const char* what, const std::string& detail = "") {
  std::printf("%s %s%s%s\n", ok ? "PASS" : "FAIL", what, detail.empty() ? "" : ": ",
              detail.c_str());
  if (!ok) ++g_fail;
}

// Seeded RGBA Q16 image: sinusoids + LCG noise, one transparent rectangle.
std::vector<rd::RgbaF> make_image(int w, int h, unsigned seed, bool alpha) {
  std::vector<rd::RgbaF> im((size_t)w * h);
  unsigned s = seed * 2654435761u + 12345u;
  auto rnd = [&]() {
    s = s * 1664525u + 1013904223u;
    return (double)(s >> 8) / 16777216.0;
  };
  const double fx = 0.01 + 0.03 * rnd(), fy = 0.01 + 0.03 * rnd(), ph = 6.28 * rnd();
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      rd::RgbaF& p = im[(size_t)y * w + x];
      const double base = 127.5 + 100.0 * std::sin(fx * x + ph) * std::cos(fy * y);
      auto ch = [&](double v) {
        const double q = std::round(v + 12.0 * (rnd() - 0.5));
        return (float)(std::fmin(255.0, std::fmax(0.0, q)) * 257.0);
      };
      p.r = ch(base);
      p.g = ch(255.0 - base);
      p.b = ch(0.5 * base + 60.0);
      p.a = (alpha && x > w / 3 && x < w / 2 && y > h / 4) ? 0.0f : 65535.0f;
    }
  return im;
}

bool same_bytes(const void* a, const void* b, size_t n) { return std::memcmp(a, b, n) == 0; }

void palette_of(const rd::ColorTree& t, rd::Palette* p) {
  p->count = t.color_count();
  p->associate_alpha = t.associate_alpha();
  for (int i = 0; i < p->count; ++i) p->entries[i] = t.colormap()[i];
}

rd::DitherResult noop_run(rd::DitherJob&, rd::RgbaF*) { return {}; }

}  // namespace

int main() {
  check(rd::CudaAvailable(), "CudaAvailable");
  if (!rd::CudaAvailable()) return 1;

  // ---- curve -------------------------------------------------------------------
  {
    bool ok = true;
    for (int64_t n = 0; n < (1 << 10) && ok; ++n) {
      int x, y, ox, oy;
      rd::HilbertPoint(n, 5, &x, &y);
      orc_hilbert_point(n, 5, &ox, &oy);
      ok = x == ox && y == oy;
    }
    check(ok, "HilbertPoint == oracle at level 5");
    const int geoms[][2] = {{64, 40}, {37, 53}, {128, 128}, {200, 150}, {1, 1}};
    for (auto& g : geoms) {
      std::vector<int> order, owner;
      const size_t n = rd::BuildCurveIndex(rd::ComputeCurveLevel(g[0], g[1]), g[0], g[1], &order,
                                           &owner);
      std::vector<int32_t> eo((size_t)g[0] * g[1] + 1), ew((size_t)g[0] * g[1]);
      const int64_t en = orc_block_order(g[0], g[1], eo.data(), ew.data());
      bool same = (int64_t)n == en && order.size() == (size_t)en &&
                  same_bytes(order.data(), eo.data(), 4 * (size_t)en) &&
                  same_bytes(owner.data(), ew.data(), 4 * ew.size());
      const std::vector<int>* bo = rd::BlockCurveOrder(g[0], g[1]);
      same = same && bo && *bo == order;
      char name[96];
      std::snprintf(name, sizeof(name), "BuildCurveIndex/BlockCurveOrder %dx%d", g[0], g[1]);
      check(same, name);
    }
    std::vector<double> g = rd::ApproxInverseFilter(96, 1.0), eg(97);
    orc_inverse_filter(96, 1.0, eg.data());
    check(g.size() == 97 && same_bytes(g.data(), eg.data(), 97 * 8), "ApproxInverseFilter");
    check(rd::ApproxInverseFilter(0, 1.0).empty(), "ApproxInverseFilter refuses taps=0");
  }

  // ---- ColorTree::Build ----------------------------------------------------------
  const int W = 173, H = 119;
  std::vector<rd::RgbaF> img = make_image(W, H, 1, true);
  rd::ColorTree tree;
  tree.Build(img.data(), W, H, 24, true, false);
  OrcTree* ot = orc_tree_build(&img[0].r, W, H, 24, 1, 0);
  {
    const int n = (int)orc_tree_node_count(ot);
    bool ok = tree.last_error().empty() && tree.color_count() == orc_tree_color_count(ot) &&
              (int)tree.node_count() == n && tree.depth() == orc_tree_depth(ot);
    std::vector<double> cm(4 * (size_t)orc_tree_color_count(ot));
    orc_tree_colormap(ot, cm.data());
    ok = ok && same_bytes(tree.colormap(), cm.data(), 8 * cm.size());
    std::vector<int32_t> child(16 * (size_t)n), parent(n);
    std::vector<uint32_t> cn(n), id(n), level(n);
    std::vector<uint64_t> nu(n);
    std::vector<double> tot(4 * (size_t)n), qe(n);
    orc_tree_nodes(ot, child.data(), parent.data(), cn.data(), nu.data(), tot.data(), qe.data(),
                   id.data(), level.data());
    for (int i = 0; i < n && ok; ++i) {
      const rd::QNode& q = tree.nodes()[i];
      ok = same_bytes(q.child, &child[16 * (size_t)i], 64) && q.parent == parent[i] &&
           q.color_number == cn[i] && q.number_unique == nu[i] && q.id == id[i] &&
           q.level == level[i] && same_bytes(q.total_color, &tot[4 * (size_t)i], 32) &&
           (i == 0 || std::memcmp(&q.quantize_error, &qe[i], 8) == 0);
    }
    check(ok, "ColorTree::Build == oracle (colormap, every QNode)", tree.last_error());
  }
  rd::Palette* pal = new rd::Palette();
  palette_of(tree, pal);
  rd::DitherParams params;

  // ---- RiemersmaWalkCuda: one host frame, frames=2 ------------------------------------
  // The reference uploads and downloads exactly one frame whatever `frames` says
  // (src/rd_riemersma_cuda.cu:688, :747), so the caller's buffer holds one frame:
  // a guard band after it must come back untouched.
  {
    const size_t px = (size_t)W * H, guard = 4096;
    std::vector<rd::RgbaF> buf(px + guard);
    std::copy(img.begin(), img.end(), buf.begin());
    for (size_t i = px; i < buf.size(); ++i) buf[i] = rd::RgbaF{-7.0f, -7.0f, -7.0f, -7.0f};
    const std::vector<rd::RgbaF> guard_copy(buf.begin() + px, buf.end());
    std::string dev;
    const std::string e = rd::RiemersmaWalkCuda(*pal, params, tree, W, H, buf.data(), 2, &dev);
    std::vector<rd::RgbaF> e1 = img;
    orc_walk_seq(ot, &e1[0].r, W, H, 1.0, 1);
    check(e.empty() && same_bytes(buf.data(), e1.data(), 16 * px) && !dev.empty(),
          "RiemersmaWalkCuda frames=2 on a one-frame buffer == RiemersmaWalkCpu", e + " " + dev);
    check(same_bytes(buf.data() + px, guard_copy.data(), 16 * guard),
          "RiemersmaWalkCuda frames=2 touches one frame only");
    rd::DitherParams nc;
    nc.use_cache = false;
    nc.diffusion = 0.75;
    std::vector<rd::RgbaF> a = img, b = img;
    const std::string e3 = rd::RiemersmaWalkCuda(*pal, nc, tree, W, H, a.data(), 1, nullptr);
    orc_walk_seq(ot, &b[0].r, W, H, 0.75, 0);
    check(e3.empty() && same_bytes(a.data(), b.data(), 16 * a.size()),
          "RiemersmaWalkCuda use_cache=false diffusion=0.75", e3);
  }

  // ---- RiemersmaWalkBatch: uneven items, own trees ------------------------------------
  {
    const int sz[][2] = {{1, 1}, {3, 5}, {17, 13}, {90, 41}, {64, 64}};
    std::vector<std::vector<rd::RgbaF>> ims, exp;
    std::vector<rd::ColorTree> trees(5);
    std::vector<OrcTree*> ots;
    std::vector<rd::WalkItem> items;
    for (int k = 0; k < 5; ++k) {
      ims.push_back(make_image(sz[k][0], sz[k][1], 10 + k, k % 2 == 0));
      trees[k].Build(ims[k].data(), sz[k][0], sz[k][1], 4 + 7 * k, true, false);
      ots.push_back(orc_tree_build(&ims[k][0].r, sz[k][0], sz[k][1], 4 + 7 * k, 1, 0));
      exp.push_back(ims[k]);
      orc_walk_seq(ots[k], &exp[k][0].r, sz[k][0], sz[k][1], 1.0, 1);
    }
    for (int k = 0; k < 5; ++k)
      items.push_back(rd::WalkItem{ims[k].data(), (size_t)sz[k][0], (size_t)sz[k][1], &trees[k]});
    const std::string e = rd::RiemersmaWalkBatch(items, params, nullptr);
    bool ok = e.empty();
    for (int k = 0; k < 5 && ok; ++k) ok = same_bytes(ims[k].data(), exp[k].data(), 16 * exp[k].size());
    check(ok, "RiemersmaWalkBatch (1x1, 3x5, 17x13, 90x41, 64x64) == RiemersmaWalkCpu", e);
    for (OrcTree* t : ots) orc_tree_free(t);
  }

  // ---- RiemersmaBlocksCuda --------------------------------------------------------------
  {
    std::vector<rd::RgbaF> batch = img;
    std::vector<rd::RgbaF> f2 = make_image(W, H, 3, true);
    batch.insert(batch.end(), f2.begin(), f2.end());
    std::vector<rd::RgbaF> expect = batch;
    rd::BlockOptions bo;
    bo.block = 64;
    bo.frames = 2;
    const int rc = orc_walk_blocks(ot, &expect[0].r, W, H, 2, 64, 1.0, nullptr);
    const std::string e = rd::RiemersmaBlocksCuda(*pal, params, tree, W, H, batch.data(), bo,
                                                  nullptr);
    check(rc == 0 && e.empty() && same_bytes(batch.data(), expect.data(), 16 * batch.size()),
          "RiemersmaBlocksCuda B=64 frames=2 == RiemersmaBlocksCpu", e);
    std::vector<rd::RgbaF> src = img;
    src.insert(src.end(), f2.begin(), f2.end());
    std::vector<uint16_t> out(4 * src.size()), eu(4 * src.size());
    orc_walk_blocks_u16(ot, &src[0].r, W, H, 2, 64, 1.0, eu.data());
    bo.emit_u16 = true;
    const std::string e2 = rd::RiemersmaBlocksCuda(*pal, params, tree, W, H, src.data(), bo,
                                                   nullptr, out.data(), 1);
    check(e2.empty() && out == eu, "RiemersmaBlocksCuda emit_u16 (state slot 1) == rgba64le", e2);
    rd::BlockOptions bad = bo;
    bad.block = 8;
    check(!rd::RiemersmaBlocksCuda(*pal, params, tree, W, H, src.data(), bad, nullptr, out.data())
               .empty(),
          "RiemersmaBlocksCuda refuses block < 16");
    check(!rd::RiemersmaBlocksCuda(*pal, params, tree, W, H, src.data(), bo, nullptr, out.data(),
                                   5)
               .empty(),
          "RiemersmaBlocksCuda refuses state_slot out of range");
    bad = bo;
    bad.frames = 0;
    check(!rd::RiemersmaBlocksCuda(*pal, params, tree, W, H, src.data(), bad, nullptr, out.data())
               .empty(),
          "RiemersmaBlocksCuda refuses frames < 1");
  }

  // ---- RiemersmaBlocksCuda: planar YUV through the uint16 pointers -------------------------
  {
    const int w = 174, h = 120, F = 2;
    const size_t pp = (size_t)w * h;
    unsigned s = 77u;
    auto rnd = [&]() {
      s = s * 1664525u + 1013904223u;
      return (uint8_t)(s >> 24);
    };
    std::vector<uint8_t> y444(3 * pp * F), y420(pp * 3 / 2 * F);
    for (uint8_t& b : y444) b = rnd();
    for (uint8_t& b : y420) b = rnd();
    std::vector<float> wide(4 * pp * F);
    rd::BlockOptions bo;
    bo.block = 64;
    bo.frames = F;
    bo.upload_u16 = true;
    bo.emit_u16 = true;
    bo.emit_yuv444 = true;
    std::vector<uint8_t> out(3 * pp * F), expect(3 * pp * F);
    orc_yuv_to_rgbaf(y444.data(), w, h, F, 7, wide.data());
    orc_walk_blocks_yuv444(ot, wide.data(), w, h, F, 64, 1.0, expect.data());
    for (rd::BlockOptions::InMode m : {rd::BlockOptions::InMode::PlanarYuv444,
                                       rd::BlockOptions::InMode::PlanarYuv444Prepass}) {
      bo.in_mode = m;
      std::fill(out.begin(), out.end(), 0);
      const std::string e = rd::RiemersmaBlocksCuda(
          *pal, params, tree, w, h, nullptr, bo, nullptr, reinterpret_cast<uint16_t*>(out.data()),
          0, reinterpret_cast<const uint16_t*>(y444.data()));
      check(e.empty() && out == expect, "RiemersmaBlocksCuda planar 4:4:4 in, emit_yuv444 out", e);
    }
    bo.in_mode = rd::BlockOptions::InMode::PlanarYuv420;
    bo.emit_yuv444 = false;
    std::vector<uint16_t> o16(4 * pp * F), e16(4 * pp * F);
    orc_yuv_to_rgbaf(y420.data(), w, h, F, 8, wide.data());
    orc_walk_blocks_u16(ot, wide.data(), w, h, F, 64, 1.0, e16.data());
    const std::string e = rd::RiemersmaBlocksCuda(*pal, params, tree, w, h, nullptr, bo, nullptr,
                                                  o16.data(), 1,
                                                  reinterpret_cast<const uint16_t*>(y420.data()));
    check(e.empty() && o16 == e16, "RiemersmaBlocksCuda planar 4:2:0 in, rgba64le out", e);
    check(!rd::RiemersmaBlocksCuda(*pal, params, tree, w - 1, h, nullptr, bo, nullptr,
                                   o16.data(), 1, reinterpret_cast<const uint16_t*>(y420.data()))
               .empty(),
          "RiemersmaBlocksCuda refuses 4:2:0 with an odd side");
  }

  // ---- RiemersmaApproxCuda --------------------------------------------------------------
  {
    std::vector<rd::RgbaF> a = img, b = img;
    rd::ApproxOptions ao;
    ao.iterations = 6;
    ao.taps = 48;
    const std::string e = rd::RiemersmaApproxCuda(*pal, params, tree, W, H, a.data(), ao, nullptr);
    const int rc = orc_jacobi(ot, &b[0].r, W, H, 6, 48, 0, 1, 1.0, nullptr);
    check(rc == 0 && e.empty() && same_bytes(a.data(), b.data(), 16 * a.size()),
          "RiemersmaApproxCuda fp32 == the Jacobi restatement", e);
    ao.taps = 5000;
    check(!rd::RiemersmaApproxCuda(*pal, params, tree, W, H, a.data(), ao, nullptr).empty(),
          "RiemersmaApproxCuda refuses taps > 4096");
  }

  // ---- palette above 256 colours: blocks and approx refuse --------------------------------
  {
    std::vector<rd::RgbaF> big = make_image(211, 197, 7, false);
    for (size_t i = 0; i < big.size(); ++i) big[i].b = (float)((i * 7919) % 65536);
    rd::ColorTree t2;
    t2.Build(big.data(), 211, 197, 300, false, false);
    rd::Palette* p2 = new rd::Palette();
    palette_of(t2, p2);
    rd::BlockOptions bo;
    bo.block = 64;
    const bool over = p2->count > 256;
    check(over && !rd::RiemersmaBlocksCuda(*p2, params, t2, 211, 197, big.data(), bo, nullptr)
                       .empty(),
          "RiemersmaBlocksCuda refuses a palette above 256 colours");
    delete p2;
  }

  // ---- plugin seam -------------------------------------------------------------------------
  {
    rd::DitherAlgorithm a{"probe", "a test dither", noop_run};
    rd::DitherAlgorithm nul{"", "empty", noop_run};
    const bool r1 = rd::RegisterDither(a);
    const bool r2 = rd::RegisterDither(a);
    const bool r3 = rd::RegisterDither(nul);
    const std::string list = rd::DitherListText();
    const std::string tail = "\n  probe           a test dither";
    check(r1 && !r2 && !r3 && rd::FindDither("probe") && !rd::FindDither("nope") &&
              rd::RegisteredDithers().size() == 8 && list.size() > tail.size() &&
              list.compare(list.size() - tail.size(), tail.size(), tail) == 0 &&
              list.rfind("  floyd-steinberg error diffusion, 4 neighbours", 0) == 0,
          "RegisterDither / FindDither / DitherListText (the seven device dithers registered)");
  }

  // ---- the screen dithers behind DitherAlgorithm, on the device ---------------------------
  {
    const char* names[4] = {"bayer-ordered", "void-and-cluster", "void-and-cluster-fast",
                            "clustered-dot"};
    std::vector<double> ent(4 * (size_t)pal->count);
    for (int i = 0; i < pal->count; ++i) {
      ent[4 * i] = pal->entries[i].r;
      ent[4 * i + 1] = pal->entries[i].g;
      ent[4 * i + 2] = pal->entries[i].b;
      ent[4 * i + 3] = pal->entries[i].a;
    }
    for (int k = 0; k < 4; ++k) {
      const rd::DitherAlgorithm* a = rd::FindDither(names[k]);
      std::vector<rd::RgbaF> got = img, want = img;
      rd::DitherJob job;
      job.palette = pal;
      job.width = W;
      job.height = H;
      job.diffusion = 0.8;
      const rd::DitherResult r = a ? a->run(job, got.data()) : rd::DitherResult{"missing"};
      orc_dither_screen(ent.data(), pal->count, &want[0].r, W, H, 1, k, 0.8);
      check(r.error.empty() && same_bytes(got.data(), want.data(), 16 * got.size()), names[k],
            r.error);
    }
    rd::DitherJob empty;
    empty.palette = pal;
    check(rd::FindDither("clustered-dot")->run(empty, nullptr).error ==
              "clustered-dot: empty frame",
          "screen dither refuses an empty frame");
    // The error-diffusion dithers, two frames per job.
    const char* dnames[3] = {"floyd-steinberg", "atkinson", "jarvis"};
    std::vector<rd::RgbaF> two = img;
    std::vector<rd::RgbaF> img2 = make_image(W, H, 9, false);
    two.insert(two.end(), img2.begin(), img2.end());
    for (int k = 0; k < 3; ++k) {
      const rd::DitherAlgorithm* a = rd::FindDither(dnames[k]);
      std::vector<rd::RgbaF> got = two, want = two;
      rd::DitherJob job;
      job.palette = pal;
      job.width = W;
      job.height = H;
      job.frames = 2;
      job.diffusion = 0.9;
      const rd::DitherResult r = a ? a->run(job, got.data()) : rd::DitherResult{"missing"};
      orc_dither_diffuse(ent.data(), pal->count, &want[0].r, W, H, 2, k, 0.9);
      check(r.error.empty() && same_bytes(got.data(), want.data(), 16 * got.size()), dnames[k],
            r.error);
    }
    check(rd::FindDither("jarvis")->run(empty, nullptr).error == "jarvis: empty frame",
          "diffusion dither refuses an empty frame");
  }

  // ---- rd_im.h: palette files and the pixel-buffer palette --------------------------
  {
    // ImBuildPaletteFromPixels == ColorTree::Build's colormap without exact duplicates.
    std::vector<rd::RgbaF> im = make_image(120, 70, 5, false);
    rd::Palette* p2 = new rd::Palette();
    rd::ColorTree t2;
    std::string e;
    const bool ok = rd::ImBuildPaletteFromPixels(im.data(), 120, 70, 12, p2, &t2, &e);
    OrcTree* o2 = orc_tree_build(&im[0].r, 120, 70, 12, 0, 0);
    std::vector<double> cm(4 * (size_t)orc_tree_color_count(o2));
    orc_tree_colormap(o2, cm.data());
    std::vector<double> uniq;
    for (size_t i = 0; i < cm.size() / 4; ++i) {
      bool dup = false;
      for (size_t k = 0; k < uniq.size() / 4 && !dup; ++k)
        dup = uniq[4 * k] == cm[4 * i] && uniq[4 * k + 1] == cm[4 * i + 1] &&
              uniq[4 * k + 2] == cm[4 * i + 2];
      if (!dup) uniq.insert(uniq.end(), cm.begin() + 4 * i, cm.begin() + 4 * i + 4);
    }
    check(ok && p2->count == (int)(uniq.size() / 4) &&
              same_bytes(&p2->entries[0].r, uniq.data(), 8 * uniq.size()) &&
              t2.color_count() == orc_tree_color_count(o2),
          "ImBuildPaletteFromPixels == ColorTree::Build colormap, duplicates removed", e);
    // ImWritePalette -> ImReadPalette: the tree of the file's entries.
    const char* tmpdir = std::getenv("TMPDIR");
    const std::string path = std::string(tmpdir && *tmpdir ? tmpdir : "/tmp") + "/dropin_palette_test.txt";
    check(rd::ImWritePalette(*p2, path, &e), "ImWritePalette .txt", e);
    rd::Palette* p3 = new rd::Palette();
    rd::ColorTree t3;
    const bool rok = rd::ImReadPalette(path, p2->count, p3, &t3, &e);
    std::vector<rd::RgbaF> strip((size_t)p3->count);
    for (int i = 0; i < p3->count; ++i)
      strip[(size_t)i] = rd::RgbaF{(float)p3->entries[i].r, (float)p3->entries[i].g,
                                   (float)p3->entries[i].b, (float)p3->entries[i].a};
    OrcTree* o3 = orc_tree_build(&strip[0].r, p3->count, 1, p3->count, 0, 0);
    std::vector<double> cm3(4 * (size_t)orc_tree_color_count(o3));
    orc_tree_colormap(o3, cm3.data());
    check(rok && p3->count == p2->count && orc_tree_color_count(o3) == p3->count &&
              same_bytes(&p3->entries[0].r, cm3.data(), 8 * cm3.size()) &&
              t3.color_count() == p3->count,
          "ImReadPalette .txt: the strip's tree colormap adopted", e);
    std::string e2;
    check(!rd::ImReadPalette(path, p2->count + 1, p3, &t3, &e2) && !e2.empty(),
          "ImReadPalette refuses a count that disagrees with --colors", e2);
    check(!rd::ImWritePalette(*p2, "palette.png", &e2), "ImWritePalette refuses the image form");
    std::remove(path.c_str());
    orc_tree_free(o2);
    orc_tree_free(o3);
    delete p2;
    delete p3;
  }

  orc_tree_free(ot);
  delete pal;
  std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
  return g_fail ? 1 : 0;
}
