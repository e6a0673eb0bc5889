// sa_palio.cpp -- the palette text file of the reference, written and read.
//
// The .txt / .text form of ImWritePalette / ImReadPalette (include/rd_im.h:
// 128-186; src/rd_im.cpp:994-1024, :1175-1330): two comment lines, then one
// row per entry, "index<TAB>r<TAB>g<TAB>b<TAB>a<TAB>RRGGBB" with the Q16 values
// at %.6f and an 8-bit hex column for people (ignored on read).  Reading
// accepts the legacy four- and five-token rows (alpha then opaque), requires
// indices contiguous from 0, checks the expected count, narrows r, g, b
// through float (the pixel type of the walks) and sorts the entries
// canonically (r, g, b, a ascending) -- the palette the reference's importer
// builds its tree from.  The image form (an indexed PNG through ImageMagick)
// is not part of this path and is refused.
//
// Host code only: no device work.
#include "../../include/stipple_amd.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

int pfail(char* err, size_t n, int code, const char* fmt, ...) {
  if (err && n) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, n, fmt, ap);
    va_end(ap);
  }
  return code;
}

bool text_path(const char* path) {
  const size_t n = std::strlen(path);
  return (n >= 4 && std::strcmp(path + n - 4, ".txt") == 0) ||
         (n >= 5 && std::strcmp(path + n - 5, ".text") == 0);
}

constexpr int kMaxColours = 65536;      // MaxColormapSize
constexpr double kOpaque = 65535.0;     // OpaqueAlpha == QuantumRange

unsigned hex8(double v) { return (unsigned)(v / 257.0 + 0.5); }

}  // namespace

extern "C" int sa_palette_write_txt(const double* entries, int count, const char* path, char* err,
                                    size_t errlen) {
  if (!path) return pfail(err, errlen, SA_EINVAL, "null path");
  if (count <= 0 || !entries) return pfail(err, errlen, SA_EINVAL, "palette is empty");
  if (!text_path(path))
    return pfail(err, errlen, SA_EUNSUPPORTED,
                 "%s: only the .txt / .text palette form is written here (the image form "
                 "goes through ImageMagick)", path);
  FILE* f = std::fopen(path, "wb");
  if (!f) return pfail(err, errlen, SA_EINVAL, "cannot open %s for writing", path);
  std::fprintf(f, "# rdither palette, %d colours, Q16 (0..65535) and 8-bit hex\n", count);
  std::fprintf(f, "# count\tr\tg\tb\ta\thex\n");
  for (int i = 0; i < count; ++i) {
    const double* e = entries + 4 * (size_t)i;
    std::fprintf(f, "%d\t%.6f\t%.6f\t%.6f\t%.6f\t%02X%02X%02X\n", i, e[0], e[1], e[2], e[3],
                 hex8(e[0]), hex8(e[1]), hex8(e[2]));
  }
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok)
    return pfail(err, errlen, SA_EINVAL, "writing %s failed", path);
  return SA_OK;
}

extern "C" int sa_palette_read_txt(const char* path, int expect_colors, double* entries,
                                   int* count, int* assoc, char* err, size_t errlen) {
  if (!path || !entries || !count) return pfail(err, errlen, SA_EINVAL, "null argument");
  *count = 0;
  if (assoc) *assoc = 0;
  if (!text_path(path))
    return pfail(err, errlen, SA_EUNSUPPORTED,
                 "%s: only the .txt / .text palette form is read here (the image form goes "
                 "through ImageMagick)", path);
  FILE* f = std::fopen(path, "rb");
  if (!f) return pfail(err, errlen, SA_EINVAL, "cannot open %s for reading", path);
  std::vector<char> seen;
  int n = 0;
  bool transparent = false;
  char line[512];  // rows are read in 511-byte pieces, as the reference reads them
  int lineno = 0;
  while (std::fgets(line, sizeof(line), f)) {
    ++lineno;
    const char c0 = line[0];
    if (c0 == '#' || c0 == '\n' || c0 == '\r') continue;
    int index = 0;
    double r = 0.0, g = 0.0, b = 0.0;
    char t5[32] = {0}, t6[32] = {0};
    const int got = std::sscanf(line, "%d %lf %lf %lf %31s %31s", &index, &r, &g, &b, t5, t6);
    if (got < 4) {
      std::fclose(f);
      return pfail(err, errlen, SA_EINVAL,
                   "%s: line %d is not a palette row (want: index r g b [a] [hex])", path,
                   lineno);
    }
    double a = kOpaque;
    if (got >= 6) {  // a five-column row: the fifth token is alpha
      char* end = nullptr;
      const double v = std::strtod(t5, &end);
      if (end == t5 || *end != '\0' || !(v >= 0.0 && v <= kOpaque)) {
        std::fclose(f);
        return pfail(err, errlen, SA_EINVAL, "%s: line %d alpha '%s' is not a number in 0..%s",
                     path, lineno, t5, std::to_string(kOpaque).c_str());
      }
      a = v;
    }
    if (index < 0 || index >= kMaxColours) {
      std::fclose(f);
      return pfail(err, errlen, SA_EINVAL, "%s: line %d index %d is out of range", path, lineno,
                   index);
    }
    if (index >= (int)seen.size()) seen.resize((size_t)index + 1, 0);
    seen[(size_t)index] = 1;
    double* e = entries + 4 * (size_t)index;
    e[0] = r;
    e[1] = g;
    e[2] = b;
    e[3] = a;
    if (a != kOpaque) transparent = true;
    n = std::max(n, index + 1);
  }
  std::fclose(f);
  if (n <= 0) return pfail(err, errlen, SA_EINVAL, "%s contains no palette rows", path);
  for (int i = 0; i < n; ++i)
    if (!seen[(size_t)i])
      return pfail(err, errlen, SA_EINVAL,
                   "%s: no row declares index %d, but the file's highest index is %d; importing "
                   "it would fill the gap with black.  Renumber the rows so the indices are "
                   "contiguous from 0, or export a fresh palette.", path, i, n - 1);
  if (expect_colors > 0 && n != expect_colors)
    return pfail(err, errlen, SA_EREFUSED,
                 "%s holds %d colours but --colors is %d; the two disagree, so the palette was "
                 "not used.  Pass the file's count as --colors, which is the only value that "
                 "works from the command line.", path, n, expect_colors);
  // The walks' pixel type is float: r, g, b narrow through it.  Then the
  // canonical order, r, g, b, a ascending.
  struct E {
    double v[4];
  };
  std::vector<E> es((size_t)n);
  for (int i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) es[(size_t)i].v[c] = (double)(float)entries[4 * (size_t)i + c];
    es[(size_t)i].v[3] = entries[4 * (size_t)i + 3];
  }
  std::sort(es.begin(), es.end(), [](const E& x, const E& y) {
    for (int c = 0; c < 4; ++c)
      if (x.v[c] != y.v[c]) return x.v[c] < y.v[c];
    return false;
  });
  for (int i = 0; i < n; ++i) std::memcpy(entries + 4 * (size_t)i, es[(size_t)i].v, 32);
  *count = n;
  if (assoc) *assoc = transparent ? 1 : 0;
  return SA_OK;
}
