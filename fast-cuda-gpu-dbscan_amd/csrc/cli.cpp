// dclust command-line driver: the counterpart of the reference's
// `main.exe <dataset-file>` (main.cu:35-46), with the common.h parameters
// (common.h:6-25) as flags instead of compile-time macros.
//
//   dclust <dataset-file> [--dim D] [--eps E] [--minpts M] [--r R]
//          [--seeds S] [--blocks B] [--threads T] [--extra X]
//          [--precision fp64|fp32] [--n N] [--out labels.txt] [--device I]
//          [--write-f64 FILE]
//
// Prints the reference's summary (main.cu:422-432: clusters, noise, phase
// times) and, with --out, writes one label per line (0 = noise).  The
// dataset may be a text file or a binary dataset (dclust_write_dataset_f64);
// --write-f64 converts the dataset to the binary format and exits without
// clustering.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dclust.h"

static void usage() {
  std::fprintf(stderr,
               "usage: dclust <dataset-file> [--dim D] [--eps E] [--minpts M] [--r R]\n"
               "              [--seeds S] [--blocks B] [--threads T] [--extra X]\n"
               "              [--precision fp64|fp32] [--n N] [--out FILE] [--device I]\n"
               "              [--write-f64 FILE]\n");
}

int main(int argc, char** argv) {
  if (argc < 2) {
    usage();
    return 2;
  }
  dclust_params p;
  dclust_params_default(&p);
  const char* in = nullptr;
  const char* out = nullptr;
  const char* conv = nullptr;
  int64_t nmax = 0;
  int device = 0;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) {
        usage();
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--dim") p.dim = std::atoi(next());
    else if (a == "--eps") p.eps = std::atof(next());
    else if (a == "--minpts") p.minpts = std::atoi(next());
    else if (a == "--r") p.partitions_r = std::atoi(next());
    else if (a == "--seeds") p.seed_list_size = std::atoi(next());
    else if (a == "--blocks") p.blocks = std::atoi(next());
    else if (a == "--threads") p.block_threads = std::atoi(next());
    else if (a == "--extra") p.extra_collision = std::atoi(next());
    else if (a == "--precision") {
      std::string v = next();
      p.precision = v == "fp32" ? DCLUST_PRECISION_FP32_GUARD : DCLUST_PRECISION_FP64;
    } else if (a == "--n") nmax = std::atoll(next());
    else if (a == "--out") out = next();
    else if (a == "--write-f64") conv = next();
    else if (a == "--device") device = std::atoi(next());
    else if (a[0] == '-') {
      usage();
      return 2;
    } else in = argv[i];
  }
  if (!in) {
    usage();
    return 2;
  }
  std::printf("Using dataset file %s\n", in);
  int64_t n = 0;
  int rc = dclust_import_dataset(in, p.dim, nmax, nullptr, &n);
  if (rc) {
    std::fprintf(stderr, "Error importing the dataset: %s\n", dclust_strerror(rc));
    return 1;
  }
  std::vector<double> xyz((size_t)n * p.dim);
  rc = dclust_import_dataset(in, p.dim, nmax, xyz.data(), &n);
  if (rc) {
    std::fprintf(stderr, "Error importing the dataset: %s\n", dclust_strerror(rc));
    return 1;
  }
  std::printf("Preprocessed %lld data in dataset\n", (long long)n);
  if (conv) {
    rc = dclust_write_dataset_f64(conv, xyz.data(), n, p.dim);
    if (rc) {
      std::fprintf(stderr, "Error writing %s: %s\n", conv, dclust_strerror(rc));
      return 1;
    }
    std::printf("Wrote binary dataset %s\n", conv);
    return 0;
  }
  dclust_handle* h = nullptr;
  rc = dclust_create(&h, device);
  if (rc) {
    std::fprintf(stderr, "dclust_create: %s\n", dclust_strerror(rc));
    return 1;
  }
  std::vector<int32_t> labels((size_t)n);
  dclust_stats st;
  auto t0 = std::chrono::steady_clock::now();
  rc = dclust_run(h, xyz.data(), n, &p, labels.data(), &st);
  auto t1 = std::chrono::steady_clock::now();
  if (rc) {
    std::fprintf(stderr, "dclust_run: %s\n", dclust_strerror(rc));
    dclust_destroy(h);
    return 1;
  }
  std::printf("==============================================\n");
  std::printf("Final cluster after merging: %lld\n", (long long)st.n_clusters);
  std::printf("Number of noises: %lld\n", (long long)st.n_noise);
  std::printf("==============================================\n");
  std::printf("Indexing Time: %.3f ms\n", st.t_index_ms);
  std::printf("Core Time: %.3f ms\n", st.t_core_ms);
  std::printf("Expansion Time: %.3f ms (%lld passes)\n", st.t_expand_ms, (long long)st.n_passes);
  std::printf("Merge Time: %.3f ms\n", st.t_merge_ms);
  std::printf("Border + Relabel Time: %.3f ms\n", st.t_border_ms + st.t_relabel_ms);
  std::printf("Total GPU Time: %.3f ms\n", st.t_total_ms);
  std::printf("Wall Time (incl. transfers): %.3f ms\n",
              std::chrono::duration<double, std::milli>(t1 - t0).count());
  std::printf("==============================================\n");
  if (out) {
    rc = dclust_write_labels(out, labels.data(), n);
    if (rc) {
      std::fprintf(stderr, "writing %s: %s\n", out, dclust_strerror(rc));
      dclust_destroy(h);
      return 1;
    }
  }
  dclust_destroy(h);
  return 0;
}
