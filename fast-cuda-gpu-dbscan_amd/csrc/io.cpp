// Dataset and label files.
//
// Dataset grammar of ImportDataset (main.cu:469-498): text lines, fields
// separated by ',', every numeric field appended to one row-major array of
// doubles; `dim` consecutive values form a point.  The reference reads at
// most DATASET_COUNT*DIMENSION values into a buffer of that size and leaves
// uninitialised memory when the file is short or a field does not parse;
// here the count comes from the file and a bad field is an error.
//
// The text is parsed by several threads: the file is read once, cut into
// chunks at line ends, every chunk is parsed twice (count and validate,
// then write at its offset).  The result is the sequential one: the same
// values (strtod on the same fields) and the same errors, including the
// max_points rule (reading stops after the line that reaches the limit, so
// a bad field later in the file is not an error).
//
// Binary datasets (no counterpart in the reference): a 24-byte header --
// the 8 bytes "DCLUSTF8", the point count (int64), the dimension (int32),
// 4 zero bytes -- then the points as row-major little-endian doubles.
// dclust_import_dataset recognises the header; dclust_write_dataset_f64
// writes the format.
//
// Label files: one integer per line, 0 = noise (test-results/
// cluster-check.cpp:201-224 reads this format).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "dclust.h"

namespace {

constexpr char kMagic[8] = {'D', 'C', 'L', 'U', 'S', 'T', 'F', '8'};
constexpr size_t kHeader = 24;

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

// Fields of one line [b, e) (no '\n' inside; the buffer holds a NUL after
// the file): calls f(value) per non-empty field, returns false on a field
// that is not exactly one number.  Empty fields (",,", trailing ',') are
// skipped, as strtok(",") skips them in the reference.
template <class F>
bool parse_line(const char* b, const char* e, F&& f) {
  const char* p = b;
  while (p < e) {
    const char* q = p;
    while (q < e && *q != ',') ++q;  // field [p, q)
    const char* s = p;
    while (s < q && is_space(*s)) ++s;
    const char* t = q;
    while (t > s && is_space(t[-1])) --t;
    if (s < t) {
      char* end = nullptr;
      const double v = std::strtod(s, &end);
      if (end != t) return false;  // not a number, or more than one token
      f(v);
    }
    p = q + 1;
  }
  return true;
}

struct Chunk {
  const char* b;
  const char* e;
  int64_t count = 0;     // values in the chunk
  const char* bad = nullptr;  // start of the first line with a bad field
};

// Values of [b, e) up to (not beyond) the first bad line; c.bad marks it.
void count_chunk(Chunk& c) {
  const char* p = c.b;
  while (p < c.e) {
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(c.e - p)));
    const char* le = nl ? nl : c.e;
    int64_t k = 0;
    if (!parse_line(p, le, [&](double) { ++k; })) {
      c.bad = p;
      return;
    }
    c.count += k;
    p = le + 1;
  }
}

// Writes the values of [b, e) to out[off ..), at most `room` of them.
void write_chunk(const Chunk& c, double* out, int64_t off, int64_t room) {
  const char* p = c.b;
  int64_t k = 0;
  while (p < c.e && k < room) {
    const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(c.e - p)));
    const char* le = nl ? nl : c.e;
    parse_line(p, le, [&](double v) {
      if (k < room) out[off + k] = v;
      ++k;
    });
    p = le + 1;
  }
}

// The whole file plus a terminating NUL (buf.size() = file size + 1).
bool read_all(const char* path, std::vector<char>& buf) {
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return false;
  bool ok = std::fseek(fp, 0, SEEK_END) == 0;
  const long sz = ok ? std::ftell(fp) : -1;
  ok = ok && sz >= 0 && std::fseek(fp, 0, SEEK_SET) == 0;
  if (ok) {
    buf.assign((size_t)sz + 1, '\0');
    ok = std::fread(buf.data(), 1, (size_t)sz, fp) == (size_t)sz;
  }
  std::fclose(fp);
  return ok;
}

// Binary dataset: the header gives the count; the points are read straight
// into `out` (nothing is read when out is NULL).
int import_binary(FILE* fp, int32_t dim, int64_t max_points, double* out, int64_t* n_points) {
  char head[kHeader];
  if (std::fseek(fp, 0, SEEK_END) != 0) return DCLUST_E_IO;
  const long size = std::ftell(fp);
  if (size < (long)kHeader || std::fseek(fp, 0, SEEK_SET) != 0 ||
      std::fread(head, 1, kHeader, fp) != kHeader)
    return DCLUST_E_IO;
  int64_t n = 0;
  int32_t d = 0;
  std::memcpy(&n, head + 8, 8);
  std::memcpy(&d, head + 16, 4);
  if (d != 2 && d != 3) return DCLUST_E_IO;
  if (d != dim) return DCLUST_E_PARAM;
  const uint64_t body = (uint64_t)size - kHeader;
  if (n < 0 || (uint64_t)n > body / (8ull * (uint64_t)d) || body != (uint64_t)n * (uint64_t)d * 8)
    return DCLUST_E_IO;  // truncated or trailing bytes
  if (max_points > 0 && n > max_points) n = max_points;
  const size_t cnt = (size_t)n * (size_t)d;
  if (out && cnt > 0 && std::fread(out, sizeof(double), cnt, fp) != cnt) return DCLUST_E_IO;
  *n_points = n;
  return DCLUST_OK;
}

// Chunk size: DCLUST_IO_CHUNK bytes when set (a test knob), else the file
// split over the threads, at least 1 MiB each.
size_t chunk_bytes(size_t size, unsigned threads) {
  if (const char* e = std::getenv("DCLUST_IO_CHUNK")) {
    const long v = std::atol(e);
    if (v > 0) return (size_t)v;
  }
  const size_t per = size / std::max(1u, threads) + 1;
  return std::max<size_t>(per, 1u << 20);
}

}  // namespace

extern "C" int dclust_import_dataset(const char* path, int32_t dim, int64_t max_points,
                                     double* out, int64_t* n_points) {
  if (!path || !n_points || (dim != 2 && dim != 3)) return DCLUST_E_PARAM;
  {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return DCLUST_E_IO;
    char magic[8];
    const bool bin = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, kMagic, 8) == 0;
    const int rc = bin ? import_binary(fp, dim, max_points, out, n_points) : DCLUST_OK;
    std::fclose(fp);
    if (bin) return rc;
  }
  std::vector<char> buf;
  if (!read_all(path, buf)) return DCLUST_E_IO;

  // chunks end at line ends
  const char* const b0 = buf.data();
  const char* const e0 = b0 + (buf.size() - 1);
  unsigned hw = std::thread::hardware_concurrency();
  const unsigned threads = std::min(16u, hw ? hw : 1u);
  const size_t step = chunk_bytes(buf.size(), threads);
  std::vector<Chunk> chunks;
  for (const char* p = b0; p < e0;) {
    const char* e = p + std::min(step, (size_t)(e0 - p));
    if (e < e0) {
      const char* nl = static_cast<const char*>(std::memchr(e, '\n', (size_t)(e0 - e)));
      e = nl ? nl + 1 : e0;
    }
    chunks.push_back(Chunk{p, e});
    p = e;
  }
  auto run = [&](auto&& fn) {
    const size_t nt = std::min<size_t>(threads, chunks.size());
    if (nt <= 1) {
      for (size_t i = 0; i < chunks.size(); ++i) fn(i);
      return;
    }
    std::vector<std::thread> pool;
    for (size_t t = 0; t < nt; ++t)
      pool.emplace_back([&, t] {
        for (size_t i = t; i < chunks.size(); i += nt) fn(i);
      });
    for (auto& th : pool) th.join();
  };
  run([&](size_t i) { count_chunk(chunks[i]); });

  // the sequential reading order decides: values up to the limit, the first
  // bad line before the line that reaches it is an error
  const int64_t cap = max_points > 0 ? max_points * dim : -1;
  int64_t total = 0;
  for (const Chunk& c : chunks) {
    if (cap >= 0 && total + c.count >= cap) {
      total = cap;  // the limit is reached on a good line of this chunk
      break;
    }
    if (c.bad) return DCLUST_E_IO;  // reached before the limit (if any)
    total += c.count;
  }
  if (cap < 0 || total < cap) {
    if (total % dim != 0) return DCLUST_E_IO;  // a partial last point
  }
  if (out && total > 0) {
    std::vector<int64_t> off(chunks.size(), 0);
    int64_t o = 0;
    for (size_t i = 0; i < chunks.size(); ++i) {
      off[i] = o;
      o += chunks[i].count;
    }
    run([&](size_t i) {
      if (off[i] < total)
        write_chunk(chunks[i], out, off[i], std::min(chunks[i].count, total - off[i]));
    });
  }
  *n_points = total / dim;
  return DCLUST_OK;
}

extern "C" int dclust_write_dataset_f64(const char* path, const double* xyz, int64_t n,
                                        int32_t dim) {
  if (!path || n < 0 || (n > 0 && !xyz) || (dim != 2 && dim != 3)) return DCLUST_E_PARAM;
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return DCLUST_E_IO;
  char head[kHeader] = {};
  std::memcpy(head, kMagic, 8);
  std::memcpy(head + 8, &n, 8);
  std::memcpy(head + 16, &dim, 4);
  bool ok = std::fwrite(head, 1, kHeader, fp) == kHeader;
  const size_t cnt = (size_t)n * (size_t)dim;
  ok = ok && (cnt == 0 || std::fwrite(xyz, sizeof(double), cnt, fp) == cnt);
  return (std::fclose(fp) == 0 && ok) ? DCLUST_OK : DCLUST_E_IO;
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
