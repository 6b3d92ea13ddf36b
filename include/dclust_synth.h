/*
 * Synthetic DBSCAN workloads (host-only helper library, libdclust_synth.so).
 *
 * The reference ships no datasets for its benchmarks (test-datasets/2d.txt and
 * 3d.txt are absent; its published runs use external real datasets, see
 * exp.cu:1101-1314).  These generators produce the five synthetic workloads
 * the benchmarks run on, deterministically from a 64-bit seed:
 * PRNG xoshiro256** seeded through splitmix64, Box-Muller normals, Fisher-
 * Yates shuffle, FP64 output in the reference's row-major layout
 * (main.cu:469-498: x0,y0[,z0],x1,y1,...).
 *
 *   config 1  2-D, 10 000 pts: 8 isotropic blobs in [0,100]^2 (sigma 1.5,
 *             90 %) + 10 % uniform noise; eps 1.0, minpts 4, r 5; seed 0.
 *   config 2  2-D, 1 000 000 pts in [0,1]^2: 64 anisotropic rotated blobs,
 *             sigmas U[0.002,0.02], Zipf(1.2) sizes, 95 % + 5 % noise;
 *             eps 0.002, minpts 4; seed 1.
 *   config 3  3-D, 4 000 000 pts in [0,1]^3: 128 blobs, sigma U[0.005,0.03],
 *             Zipf(1.1) sizes, 97 % + 3 % noise; minpts 10; eps calibrated so
 *             the median neighbour count over core points is ~30; seed 2.
 *   config 4  batched: instances of uneven size, see dclust_synth_batch_*.
 *   config 5  3-D, 32 000 000 pts in [0,1]^3: 16 equal blobs sigma 0.002
 *             (91 %), 2 cubic-Bezier filaments (4 % each, jitter 1e-4),
 *             1 % noise; eps 0.001, minpts 20, seed list 64; seed 4.
 */
#ifndef DCLUST_SYNTH_H
#define DCLUST_SYNTH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dclust_synth_info {
  int32_t dim;
  int32_t minpts;
  int32_t partitions_r;   /* 0 = cell width from eps */
  int32_t seed_list_size; /* 0 = library default     */
  double eps;
  int64_t n;
  uint64_t seed;
} dclust_synth_info;

/* Fill *info for config 1,2,3 or 5 and, when xyz_out != NULL, write
 * info->n * info->dim doubles.  n_override > 0 scales every component of the
 * mixture to a total of n_override points (same distribution, same eps).
 * seed_override == UINT64_MAX keeps the config's own seed.  Returns 0 or a
 * negative error. */
int dclust_synth_config(int32_t config, int64_t n_override,
                        uint64_t seed_override, double* xyz_out,
                        dclust_synth_info* info);

/* Batched workload (config 4): `count` instances, instance i of dimension
 * 2 or 3 (p = 0.5) and log-uniform size in [200, 200000].  Fills n_out[i]
 * and dim_out[i] (either may be NULL). */
int dclust_synth_batch_plan(int32_t count, uint64_t seed, int64_t* n_out,
                            int32_t* dim_out);

/* Generate instance i of the batched workload: a mixture of 2..32 Gaussian
 * blobs (sigma U[0.005,0.05]) + 5 % noise in [0,1]^D, minpts in {4,8,16},
 * eps = f * (median distance to the 10th nearest neighbour), f ~ U[0.5,2].
 * xyz_out may be NULL to obtain only n and dim. */
int dclust_synth_batch_instance(uint64_t seed, int32_t i, double* xyz_out,
                                dclust_synth_info* info);

/* Median over up to `sample` points of the distance to the k-th nearest
 * other point (exact, grid search). */
double dclust_synth_knn_median(const double* xyz, int64_t n, int32_t dim,
                               int32_t k, int64_t sample, uint64_t seed);

/* Bisection for eps such that the median of |N_eps(p)| (self included) over
 * sampled points with |N_eps(p)| >= minpts equals target_median. */
double dclust_synth_calibrate_eps(const double* xyz, int64_t n, int32_t dim,
                                  int32_t minpts, double target_median,
                                  int64_t sample, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
