// Synthetic DBSCAN workloads (host-only).  See include/dclust_synth.h.
//
// The reference has no generator: its benchmarks read external datasets
// (main.cu:469-498 ImportDataset; exp.cu:1101-1314 sweeps).  Everything here
// is this project's own definition of the synthetic benchmark inputs.
#include "dclust_synth.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += kGolden);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// xoshiro256** (Blackman & Vigna), Box-Muller normals.
struct Rng {
  uint64_t s[4];
  bool have = false;
  double spare = 0.0;
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& w : s) w = splitmix64(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
  double uni(double a, double b) { return a + (b - a) * u01(); }
  uint64_t below(uint64_t m) {
    return (uint64_t)(((unsigned __int128)next() * m) >> 64);
  }
  double normal() {
    if (have) {
      have = false;
      return spare;
    }
    double u1 = (double)((next() >> 11) + 1) * 0x1.0p-53;  // (0,1]
    double u2 = u01();
    double r = std::sqrt(-2.0 * std::log(u1));
    double th = 6.283185307179586476925286766559 * u2;
    spare = r * std::sin(th);
    have = true;
    return r * std::cos(th);
  }
};

bool inside_unit(const double* p, int dim) {
  for (int k = 0; k < dim; ++k)
    if (p[k] < 0.0 || p[k] > 1.0) return false;
  return true;
}

// Split `total` over weights: floor shares, remainder one each to the first.
std::vector<int64_t> split_counts(int64_t total, const std::vector<double>& w) {
  double sw = 0;
  for (double x : w) sw += x;
  std::vector<int64_t> c(w.size());
  int64_t used = 0;
  for (size_t j = 0; j < w.size(); ++j) {
    c[j] = (int64_t)std::floor((double)total * w[j] / sw);
    used += c[j];
  }
  for (size_t j = 0; used < total; j = (j + 1) % w.size(), ++used) c[j]++;
  return c;
}

std::vector<double> zipf(int k, double s) {
  std::vector<double> w(k);
  for (int j = 0; j < k; ++j) w[j] = std::pow((double)(j + 1), -s);
  return w;
}

void shuffle_rows(double* xyz, int64_t n, int dim, Rng& rng) {
  for (int64_t i = n - 1; i > 0; --i) {
    int64_t j = (int64_t)rng.below((uint64_t)i + 1);
    if (j != i)
      for (int k = 0; k < dim; ++k) std::swap(xyz[i * dim + k], xyz[j * dim + k]);
  }
}

// ---------------------------------------------------------------- grid ----
struct Grid {
  int dim = 0;
  double mn[3] = {0, 0, 0}, inv[3] = {1, 1, 1};
  int64_t nk[3] = {1, 1, 1};
  std::vector<int64_t> key, start;
  std::vector<int32_t> idx;

  int64_t cell(int k, double x) const {
    double t = std::floor((x - mn[k]) * inv[k]);
    if (!(t >= 0.0)) return 0;
    if (t >= (double)(nk[k] - 1)) return nk[k] - 1;
    return (int64_t)t;
  }
  int64_t lin(const int64_t* c) const {
    int64_t v = c[0];
    for (int k = 1; k < dim; ++k) v = v * nk[k] + c[k];
    return v;
  }
  void build(const double* xyz, int64_t n, int d, double h) {
    dim = d;
    double mx[3];
    for (int k = 0; k < d; ++k) {
      mn[k] = INFINITY;
      mx[k] = -INFINITY;
    }
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < d; ++k) {
        mn[k] = std::min(mn[k], xyz[i * d + k]);
        mx[k] = std::max(mx[k], xyz[i * d + k]);
      }
    for (int k = 0; k < d; ++k) {
      double ext = mx[k] - mn[k];
      double cw = h;
      if (ext / cw > 1048576.0) cw = ext / 1048576.0;
      inv[k] = 1.0 / cw;
      nk[k] = (int64_t)std::floor(ext * inv[k]) + 1;
    }
    std::vector<std::pair<int64_t, int32_t>> kv(n);
    for (int64_t i = 0; i < n; ++i) {
      int64_t c[3] = {0, 0, 0};
      for (int k = 0; k < d; ++k) c[k] = cell(k, xyz[i * d + k]);
      kv[i] = {lin(c), (int32_t)i};
    }
    std::sort(kv.begin(), kv.end());
    idx.resize(n);
    key.clear();
    start.clear();
    for (int64_t i = 0; i < n; ++i) {
      idx[i] = kv[i].second;
      if (i == 0 || kv[i].first != kv[i - 1].first) {
        key.push_back(kv[i].first);
        start.push_back(i);
      }
    }
    start.push_back(n);
  }
  // visit all points in cells overlapping the box [p-r, p+r]
  template <class F>
  void visit(const double* p, double r, F&& f) const {
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) {
      lo[k] = cell(k, p[k] - r);
      hi[k] = cell(k, p[k] + r);
    }
    int last = dim - 1;
    int64_t c[3];
    for (int64_t a = lo[0]; a <= hi[0]; ++a)
      for (int64_t b = (dim == 3 ? lo[1] : 0); b <= (dim == 3 ? hi[1] : 0); ++b) {
        c[0] = a;
        if (dim == 3) c[1] = b;
        c[last] = lo[last];
        int64_t klo = lin(c);
        c[last] = hi[last];
        int64_t khi = lin(c);
        auto it = std::lower_bound(key.begin(), key.end(), klo);
        for (size_t s = it - key.begin(); s < key.size() && key[s] <= khi; ++s)
          for (int64_t j = start[s]; j < start[s + 1]; ++j) f(idx[j]);
      }
  }
};

double dist2(const double* a, const double* b, int dim) {
  double s = 0;
  for (int k = 0; k < dim; ++k) {
    double d = a[k] - b[k];
    s += d * d;
  }
  return s;
}

// ------------------------------------------------------------ configs ----
struct Blob {
  double c[3];
  double sig[3];
  double rot;  // 2-D rotation angle (config 2); 0 otherwise
};

void gen_blob_point(Rng& rng, const Blob& b, int dim, double* out) {
  double z[3];
  for (int k = 0; k < dim; ++k) z[k] = rng.normal() * b.sig[k];
  if (dim == 2 && b.rot != 0.0) {
    double cs = std::cos(b.rot), sn = std::sin(b.rot);
    out[0] = b.c[0] + cs * z[0] - sn * z[1];
    out[1] = b.c[1] + sn * z[0] + cs * z[1];
  } else {
    for (int k = 0; k < dim; ++k) out[k] = b.c[k] + z[k];
  }
}

struct Bezier {
  double p[4][3];
};

void gen_bezier_point(Rng& rng, const Bezier& bz, double jitter, double* out) {
  double t = rng.u01(), u = 1.0 - t;
  double w0 = u * u * u, w1 = 3 * u * u * t, w2 = 3 * u * t * t, w3 = t * t * t;
  for (int k = 0; k < 3; ++k)
    out[k] = w0 * bz.p[0][k] + w1 * bz.p[1][k] + w2 * bz.p[2][k] + w3 * bz.p[3][k] +
             jitter * rng.normal();
}

// eps for config 3: dclust_synth_calibrate_eps on the full 4 M points
// (target median neighbour count 30 over core points, minpts 10, 20 000
// sampled points) returned 0.0014246294...; the config uses it rounded.
constexpr double kCfg3Eps = 0.001425;

int config_info(int32_t config, dclust_synth_info* info) {
  std::memset(info, 0, sizeof(*info));
  switch (config) {
    case 1: *info = {2, 4, 5, 0, 1.0, 10000, 0}; return 0;
    case 2: *info = {2, 4, 0, 0, 0.002, 1000000, 1}; return 0;
    case 3: *info = {3, 10, 0, 0, kCfg3Eps, 4000000, 2}; return 0;
    case 5: *info = {3, 20, 0, 64, 0.001, 32000000, 4}; return 0;
    default: return -1;
  }
}

}  // namespace

extern "C" int dclust_synth_config(int32_t config, int64_t n_override,
                                   uint64_t seed_override, double* xyz,
                                   dclust_synth_info* info) {
  if (!info) return -1;
  if (config_info(config, info)) return -1;
  if (n_override > 0) info->n = n_override;
  if (seed_override != UINT64_MAX) info->seed = seed_override;
  if (!xyz) return 0;
  const int dim = info->dim;
  const int64_t n = info->n;
  Rng rng(info->seed);
  int64_t w = 0;  // rows written
  auto put = [&](const double* p) {
    for (int k = 0; k < dim; ++k) xyz[w * dim + k] = p[k];
    ++w;
  };
  double p[3];
  if (config == 1) {
    // 8 isotropic blobs, centres U[0,100]^2, sigma 1.5, 90 % + 10 % noise
    int64_t n_noise = (int64_t)std::llround(0.1 * (double)n);
    std::vector<Blob> bl(8);
    for (auto& b : bl) {
      b.c[0] = rng.uni(0, 100);
      b.c[1] = rng.uni(0, 100);
      b.sig[0] = b.sig[1] = 1.5;
      b.rot = 0;
    }
    auto cnt = split_counts(n - n_noise, std::vector<double>(8, 1.0));
    for (int j = 0; j < 8; ++j)
      for (int64_t i = 0; i < cnt[j]; ++i) {
        gen_blob_point(rng, bl[j], 2, p);
        put(p);
      }
    for (int64_t i = 0; i < n_noise; ++i) {
      p[0] = rng.uni(0, 100);
      p[1] = rng.uni(0, 100);
      put(p);
    }
  } else if (config == 2 || config == 3) {
    const int nb = config == 2 ? 64 : 128;
    const double zs = config == 2 ? 1.2 : 1.1;
    const double noise_frac = config == 2 ? 0.05 : 0.03;
    const double slo = config == 2 ? 0.002 : 0.005;
    const double shi = config == 2 ? 0.02 : 0.03;
    std::vector<Blob> bl(nb);
    for (auto& b : bl) {
      for (int k = 0; k < dim; ++k) b.c[k] = rng.u01();
      if (config == 2) {
        b.sig[0] = rng.uni(slo, shi);
        b.sig[1] = rng.uni(slo, shi);
        b.rot = rng.uni(0, 3.14159265358979323846);
      } else {
        double s = rng.uni(slo, shi);
        b.sig[0] = b.sig[1] = b.sig[2] = s;
        b.rot = 0;
      }
    }
    int64_t n_noise = (int64_t)std::llround(noise_frac * (double)n);
    auto cnt = split_counts(n - n_noise, zipf(nb, zs));
    for (int j = 0; j < nb; ++j)
      for (int64_t i = 0; i < cnt[j]; ++i) {
        do gen_blob_point(rng, bl[j], dim, p);
        while (!inside_unit(p, dim));
        put(p);
      }
    for (int64_t i = 0; i < n_noise; ++i) {
      for (int k = 0; k < dim; ++k) p[k] = rng.u01();
      put(p);
    }
  } else if (config == 5) {
    // 16 equal blobs sigma 0.002 (centres U[0.1,0.9]^3), 2 filaments of 4 %
    // each (cubic Bezier, control points U[0.1,0.9]^3, jitter 1e-4), 1 % noise
    std::vector<Blob> bl(16);
    for (auto& b : bl) {
      for (int k = 0; k < 3; ++k) b.c[k] = rng.uni(0.1, 0.9);
      b.sig[0] = b.sig[1] = b.sig[2] = 0.002;
      b.rot = 0;
    }
    Bezier bz[2];
    for (auto& z : bz)
      for (auto& q : z.p)
        for (int k = 0; k < 3; ++k) q[k] = rng.uni(0.1, 0.9);
    int64_t n_noise = (int64_t)std::llround(0.01 * (double)n);
    int64_t n_fil = (int64_t)std::llround(0.04 * (double)n);
    auto cnt = split_counts(n - n_noise - 2 * n_fil, std::vector<double>(16, 1.0));
    for (int j = 0; j < 16; ++j)
      for (int64_t i = 0; i < cnt[j]; ++i) {
        do gen_blob_point(rng, bl[j], 3, p);
        while (!inside_unit(p, 3));
        put(p);
      }
    for (int f = 0; f < 2; ++f)
      for (int64_t i = 0; i < n_fil; ++i) {
        do gen_bezier_point(rng, bz[f], 1e-4, p);
        while (!inside_unit(p, 3));
        put(p);
      }
    for (int64_t i = 0; i < n_noise; ++i) {
      for (int k = 0; k < 3; ++k) p[k] = rng.u01();
      put(p);
    }
  }
  if (w != n) return -2;
  shuffle_rows(xyz, n, dim, rng);
  return 0;
}

// ------------------------------------------------------------- batch ----
namespace {
Rng instance_rng(uint64_t seed, int32_t i) {
  uint64_t x = seed ^ (kGolden * (uint64_t)(i + 1));
  return Rng(splitmix64(x));
}
void instance_shape(Rng& rng, int64_t* n, int32_t* dim) {
  *dim = 2 + (int32_t)(rng.next() & 1);
  double ln = rng.uni(std::log(200.0), std::log(200000.0));
  *n = (int64_t)std::llround(std::exp(ln));
}
}  // namespace

extern "C" int dclust_synth_batch_plan(int32_t count, uint64_t seed,
                                       int64_t* n_out, int32_t* dim_out) {
  if (count < 0) return -1;
  for (int32_t i = 0; i < count; ++i) {
    Rng rng = instance_rng(seed, i);
    int64_t n;
    int32_t d;
    instance_shape(rng, &n, &d);
    if (n_out) n_out[i] = n;
    if (dim_out) dim_out[i] = d;
  }
  return 0;
}

extern "C" int dclust_synth_batch_instance(uint64_t seed, int32_t i,
                                           double* xyz,
                                           dclust_synth_info* info) {
  if (!info || i < 0) return -1;
  Rng rng = instance_rng(seed, i);
  int64_t n;
  int32_t dim;
  instance_shape(rng, &n, &dim);
  int nb = 2 + (int)rng.below(31);
  static const int32_t kMinPts[3] = {4, 8, 16};
  int32_t minpts = kMinPts[rng.below(3)];
  double f = rng.uni(0.5, 2.0);
  std::memset(info, 0, sizeof(*info));
  info->dim = dim;
  info->minpts = minpts;
  info->n = n;
  info->seed = seed;
  if (!xyz) return 0;
  std::vector<Blob> bl(nb);
  for (auto& b : bl) {
    for (int k = 0; k < dim; ++k) b.c[k] = rng.u01();
    double s = rng.uni(0.005, 0.05);
    b.sig[0] = b.sig[1] = b.sig[2] = s;
    b.rot = 0;
  }
  int64_t n_noise = (int64_t)std::llround(0.05 * (double)n);
  auto cnt = split_counts(n - n_noise, std::vector<double>(nb, 1.0));
  int64_t w = 0;
  double p[3];
  for (int j = 0; j < nb; ++j)
    for (int64_t t = 0; t < cnt[j]; ++t) {
      do gen_blob_point(rng, bl[j], dim, p);
      while (!inside_unit(p, dim));
      for (int k = 0; k < dim; ++k) xyz[w * dim + k] = p[k];
      ++w;
    }
  for (int64_t t = 0; t < n_noise; ++t) {
    for (int k = 0; k < dim; ++k) xyz[w * dim + k] = rng.u01();
    ++w;
  }
  shuffle_rows(xyz, n, dim, rng);
  double med = dclust_synth_knn_median(xyz, n, dim, 10, 1000, seed + (uint64_t)i);
  info->eps = f * med;
  return 0;
}

extern "C" double dclust_synth_knn_median(const double* xyz, int64_t n,
                                          int32_t dim, int32_t k,
                                          int64_t sample, uint64_t seed) {
  (void)seed;
  if (n <= k) return 0.0;
  // cell width for ~4 points per cell over the bounding box
  double mn[3], mx[3], vol = 1.0;
  for (int d = 0; d < dim; ++d) {
    mn[d] = INFINITY;
    mx[d] = -INFINITY;
  }
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < dim; ++d) {
      mn[d] = std::min(mn[d], xyz[i * dim + d]);
      mx[d] = std::max(mx[d], xyz[i * dim + d]);
    }
  for (int d = 0; d < dim; ++d) vol *= std::max(mx[d] - mn[d], 1e-12);
  double h = std::pow(vol * 4.0 / (double)n, 1.0 / dim);
  Grid g;
  g.build(xyz, n, dim, h);
  int64_t m = std::min(sample, n);
  std::vector<double> kd(m), d2;
  for (int64_t s = 0; s < m; ++s) {
    const double* p = xyz + s * dim;  // rows are already shuffled
    double r = h;
    for (;;) {
      d2.clear();
      g.visit(p, r, [&](int32_t j) { d2.push_back(dist2(p, xyz + (int64_t)j * dim, dim)); });
      bool whole = true;
      for (int d = 0; d < dim; ++d)
        if (p[d] - r > mn[d] || p[d] + r < mx[d]) whole = false;
      if ((int64_t)d2.size() > k) {
        std::nth_element(d2.begin(), d2.begin() + k, d2.end());
        double dk = d2[k];  // index 0 is the point itself (distance 0)
        if (dk <= r * r || whole) {
          kd[s] = std::sqrt(dk);
          break;
        }
      }
      if (whole && (int64_t)d2.size() <= k) {
        kd[s] = 0.0;
        break;
      }
      r *= 2.0;
    }
  }
  std::nth_element(kd.begin(), kd.begin() + m / 2, kd.end());
  return kd[m / 2];
}

extern "C" double dclust_synth_calibrate_eps(const double* xyz, int64_t n,
                                             int32_t dim, int32_t minpts,
                                             double target, int64_t sample,
                                             uint64_t seed) {
  (void)seed;
  int64_t m = std::min(sample, n);
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t i = 0; i < n * dim; ++i) {
    mn = std::min(mn, xyz[i]);
    mx = std::max(mx, xyz[i]);
  }
  double ext = mx - mn;
  double lo = ext * 1e-6, hi = ext * 0.05;
  Grid g;
  double grid_h = 0;
  std::vector<int64_t> cnt(m), core;
  for (int it = 0; it < 40; ++it) {
    double eps = 0.5 * (lo + hi);
    if (grid_h == 0 || eps < 0.5 * grid_h || eps > 2.0 * grid_h) {
      g.build(xyz, n, dim, eps);
      grid_h = eps;
    }
    double e2 = eps * eps;
    core.clear();
    for (int64_t s = 0; s < m; ++s) {
      const double* p = xyz + s * dim;
      int64_t c = 0;
      g.visit(p, eps, [&](int32_t j) { c += dist2(p, xyz + (int64_t)j * dim, dim) <= e2; });
      if (c >= minpts) core.push_back(c);
    }
    double med = 0;
    if (!core.empty()) {
      std::nth_element(core.begin(), core.begin() + core.size() / 2, core.end());
      med = (double)core[core.size() / 2];
    }
    if (med < target) lo = eps; else hi = eps;
  }
  return 0.5 * (lo + hi);
}
