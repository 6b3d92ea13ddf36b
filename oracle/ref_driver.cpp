// Reference-oracle build driver (test infrastructure only).
//
// Compiles the reference's own FP64 sequential R-tree DBSCAN
// (cpu-dbscan/dbscan_rtree_exp.cpp with its vendored Rtree.h) IN PLACE from
// the reference tree -- the source is included by path, never copied -- and
// exposes its DBSCAN class through one C entry point so that tests can
// compare labels.  The reference file's own main() (hard-coded dataset
// sweeps) is renamed away; `private` is opened so the per-point `clusters`
// array of DBSCAN::run() (dbscan_rtree_exp.cpp:360-406) can be read.
#include <stdint.h>
#define main dbscan_rtree_exp_unused_main
#define private public
#include "dbscan_rtree_exp.cpp"
#undef private
#undef main

extern "C" int ref_rtree_dbscan(const double* xy, int64_t n, double eps,
                                int minpts, int32_t* labels) {
  if (n <= 0) return 0;
  DATASET_SIZE = (int)n;  // global read by the DBSCAN ctor (:14, :320-341)
  double** rows = (double**)malloc(sizeof(double*) * n);
  for (int64_t i = 0; i < n; ++i) rows[i] = (double*)(xy + 2 * i);
  {
    DBSCAN d(rows, eps, minpts);
    d.run();
    for (int64_t i = 0; i < n; ++i)
      labels[i] = d.clusters[i] < 0 ? 0 : d.clusters[i];  // -1 = noise -> 0
  }
  free(rows);
  return 0;
}
