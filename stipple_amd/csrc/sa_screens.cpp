// sa_screens.cpp -- the threshold screens of the plugin dithers (host, once
// per process): the 8x8 Bayer matrix, the 8x8 clustered-dot cell and the
// void-and-cluster rank tables of side 32 and 64.
#include <cstddef>
#include <cstdint>
#include <vector>

#include "sa_internal.h"

namespace sa {

// Void-and-cluster ranks of an n x n torus, 1-based (BuildVoidCluster,
// src/rd_dithers_ordered.cc:220-330): a seeded 10% initial pattern, swaps of
// the tightest cluster into the largest void until they coincide, then ranks
// by removing clusters down from the settled pattern and filling voids up
// from it.  Energies are a separable 5-tap blur on the torus, summed in tap
// order; ties go to the first index.
std::vector<uint16_t> void_cluster_ranks(int n) {
  const size_t np = (size_t)n * (size_t)n;
  static const double kK[5] = {0.12012, 0.23387, 0.29202, 0.23387, 0.12012};
  std::vector<double> on(np), tmp(np), filt(np);
  uint64_t s = 0x9E3779B97F4A7C15ull;
  for (size_t i = 0; i < np; ++i) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    on[i] = ((s >> 33) % 10 == 0) ? 1.0 : 0.0;
  }
  std::vector<int> nb((size_t)n * 5);  // the 5 torus neighbours of each coordinate
  for (int i = 0; i < n; ++i)
    for (int t = 0; t < 5; ++t) nb[(size_t)i * 5 + t] = ((i + t - 2) % n + n) % n;
  auto blur = [&]() {
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        double acc = 0.0;
        for (int t = 0; t < 5; ++t) acc += kK[t] * on[(size_t)y * n + nb[(size_t)x * 5 + t]];
        tmp[(size_t)y * n + x] = acc;
      }
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x) {
        double acc = 0.0;
        for (int t = 0; t < 5; ++t) acc += kK[t] * tmp[(size_t)nb[(size_t)y * 5 + t] * n + x];
        filt[(size_t)y * n + x] = acc;
      }
  };
  // The densest ON site (want_on) or the emptiest OFF site; np when none.
  auto extreme = [&](bool want_on) {
    size_t best = np;
    for (size_t i = 0; i < np; ++i) {
      if ((on[i] > 0.5) != want_on) continue;
      if (best == np || (want_on ? filt[i] > filt[best] : filt[i] < filt[best])) best = i;
    }
    return best;
  };
  for (size_t it = 0; it < 4 * np; ++it) {
    blur();
    const size_t tight = extreme(true);
    if (tight == np) break;
    on[tight] = 0.0;
    blur();
    const size_t hole = extreme(false);
    if (hole == np) break;
    on[hole] = 1.0;
    if (tight == hole) break;
  }
  const std::vector<double> settled = on;
  size_t dots = 0;
  for (size_t i = 0; i < np; ++i) dots += on[i] > 0.5;
  std::vector<uint16_t> rank(np, 0);
  for (size_t r = dots; r > 0; --r) {
    blur();
    const size_t tight = extreme(true);
    if (tight == np) break;
    on[tight] = 0.0;
    rank[tight] = (uint16_t)r;
  }
  on = settled;
  for (size_t r = dots; r < np; ++r) {
    blur();
    const size_t hole = extreme(false);
    if (hole == np) break;
    on[hole] = 1.0;
    rank[hole] = (uint16_t)(r + 1);
  }
  return rank;
}

// The 8x8 screens (src/rd_dithers_ordered.cc:68-77; src/rd_dithers.cc:302-311).
const uint16_t kBayer8[64] = {0,  32, 8,  40, 2,  34, 10, 42, 48, 16, 56, 24, 50, 18, 58, 26,
                              12, 44, 4,  36, 14, 46, 6,  38, 60, 28, 52, 20, 62, 30, 54, 22,
                              3,  35, 11, 43, 1,  33, 9,  41, 51, 19, 59, 27, 49, 17, 57, 25,
                              15, 47, 7,  39, 13, 45, 5,  37, 63, 31, 55, 23, 61, 29, 53, 21};
const uint16_t kDot8[64] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 1, 1, 1, 1,
                            0, 0, 0, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 1, 0, 0, 0, 1, 1,
                            1, 1, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

}  // namespace sa
