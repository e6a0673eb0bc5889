// Host check of two walk-path identities in sa_math.h:
//   memo_key_part composes to memo_store_key (every channel byte, both modes);
//   scale_char_clamped equals scale_char on every float in [0, 65535], -0, NaN.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

#include "../../stipple_amd/csrc/sa_math.h"

int main() {
  using namespace sa;
  long bad = 0;
  for (unsigned a = 0; a < 256; a += 1)
    for (unsigned b = 0; b < 256; b += 3)
      for (unsigned c = 0; c < 256; c += 5)
        for (unsigned d = 0; d < 256; d += 7)
          for (int assoc = 0; assoc < 2; ++assoc) {
            uint32_t k = memo_key_part(0, a) | memo_key_part(1, b) | memo_key_part(2, c);
            if (assoc) k |= memo_key_part(3, d);
            if (k != memo_store_key(assoc != 0, a, b, c, assoc ? d : 0u)) ++bad;
          }
  // every float in [0, 65535]
  float q = 0.0f;
  long n = 0;
  while (q <= 65535.0f) {
    if (scale_char_clamped(q) != scale_char(q)) ++bad;
    q = std::nextafter(q, 70000.0f);
    ++n;
  }
  if (scale_char_clamped(-0.0f) != scale_char(-0.0f)) ++bad;
  if (scale_char_clamped(std::numeric_limits<float>::quiet_NaN()) != 0u) ++bad;
  std::printf("floats %ld bad %ld\n", n, bad);
  return bad ? 1 : 0;
}
