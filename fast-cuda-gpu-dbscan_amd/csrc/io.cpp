// Dataset and label files.
//
// Dataset grammar of ImportDataset (main.cu:469-498): text lines, fields
// separated by ',', every numeric field appended to one row-major array of
// doubles; `dim` consecutive values form a point.  The reference reads at
// most DATASET_COUNT*DIMENSION values into a buffer of that size and leaves
// uninitialised memory when the file is short or a field does not parse;
// here the count comes from the file and a bad field is an error.
//
// Label files: one integer per line, 0 = noise (test-results/
// cluster-check.cpp:201-224 reads this format).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dclust.h"

namespace {

// strtok(",")-equivalent field walk over one line; calls f(value) for each
// field.  Returns false on a field that is not a number.
template <class F>
bool parse_line(char* line, F&& f) {
  char* save = nullptr;
  for (char* tok = strtok_r(line, ",", &save); tok; tok = strtok_r(nullptr, ",", &save)) {
    char* p = tok;
    while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
    if (*p == '\0') continue;  // trailing separator / blank tail
    errno = 0;
    char* end = nullptr;
    double v = std::strtod(p, &end);
    if (end == p) return false;
    while (*end == ' ' || *end == '\t' || *end == '\r' || *end == '\n') ++end;
    if (*end != '\0') return false;
    f(v);
  }
  return true;
}

}  // namespace

extern "C" int dclust_import_dataset(const char* path, int32_t dim, int64_t max_points,
                                     double* out, int64_t* n_points) {
  if (!path || !n_points || (dim != 2 && dim != 3)) return DCLUST_E_PARAM;
  FILE* fp = std::fopen(path, "r");
  if (!fp) return DCLUST_E_IO;
  const int64_t cap = max_points > 0 ? max_points * dim : -1;
  int64_t cnt = 0;
  bool ok = true;
  std::vector<char> buf(1 << 16);
  size_t used = 0;
  for (;;) {
    // read one full line of any length
    if (!std::fgets(buf.data() + used, (int)(buf.size() - used), fp)) {
      if (used == 0) break;
    } else {
      used += std::strlen(buf.data() + used);
      if (used + 1 >= buf.size() && buf[used - 1] != '\n' && !std::feof(fp)) {
        buf.resize(buf.size() * 2);
        continue;
      }
    }
    if (used == 0) break;
    ok = parse_line(buf.data(), [&](double v) {
      if (cap >= 0 && cnt >= cap) return;
      if (out) out[cnt] = v;
      ++cnt;
    });
    used = 0;
    if (!ok) break;
    if (cap >= 0 && cnt >= cap) break;
  }
  std::fclose(fp);
  if (!ok) return DCLUST_E_IO;
  if (cnt % dim != 0) return DCLUST_E_IO;  // a partial last point
  *n_points = cnt / dim;
  return DCLUST_OK;
}

extern "C" int dclust_write_labels(const char* path, const int32_t* labels, int64_t n) {
  if (!path || (n > 0 && !labels) || n < 0) return DCLUST_E_PARAM;
  FILE* fp = std::fopen(path, "w");
  if (!fp) return DCLUST_E_IO;
  for (int64_t i = 0; i < n; ++i)
    if (std::fprintf(fp, "%d\n", labels[i]) < 0) {
      std::fclose(fp);
      return DCLUST_E_IO;
    }
  return std::fclose(fp) == 0 ? DCLUST_OK : DCLUST_E_IO;
}
