/*
 * CPU oracle for the DBSCAN hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Nothing in the product library links or calls this code.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only
 * as the checker (or as the timed CPU baseline), never as the thing measured.
 *
 * Semantics restated from the reference's FP64 sequential R-tree DBSCAN
 * (cpu-dbscan/dbscan_rtree_exp.cpp:319-438 and the closed-box search of
 * cpu-dbscan/Rtree.h:1573-1587,1605-1643), generalised from 2-D to D in {2,3}:
 *
 *   P(p,q)  = for all k: fl(p_k - eps) <= q_k <= fl(p_k + eps)          (box)
 *             and fl(..fl(fl(dx0*dx0) + fl(dx1*dx1)) [+ fl(dx2*dx2)]) <= fl(eps*eps)
 *             with dx_k = fl(p_k - q_k), summed left to right, no FMA.
 *   |N(p)| counts p itself; p is core iff |N(p)| >= minpts.
 *
 * Labels are written 0 = noise, 1..K = cluster id (the label-file contract of
 * test-results/cluster-check.cpp:120-139).
 */
#ifndef DCLUST_ORACLE_H
#define DCLUST_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Exact neighbour predicate, evaluated from p's side (p = query). */
int oracle_pred(const double* p, const double* q, int dim, double eps);

/* Faithful restatement of DBSCAN::run() (dbscan_rtree_exp.cpp:360-406):
 * sequential, index order, growing neighbour list.  Slow; small inputs. */
int oracle_seq_dbscan(const double* xyz, int64_t n, int dim, double eps,
                      int minpts, int32_t* labels);

/* Fast exact formulation of the same labels (grid + core flags + union-find
 * over core-core edges + min-id border rule).  Also returns core flags (may
 * be NULL) and the number of core-core pairs whose predicate is asymmetric
 * (may be NULL).  nthreads <= 0 uses all OpenMP threads. */
int oracle_fast_dbscan(const double* xyz, int64_t n, int dim, double eps,
                       int minpts, int nthreads, int32_t* labels,
                       uint8_t* core_out, int64_t* asym_pairs_out);

/* The same labels and core flags as oracle_fast_dbscan, with the core-core
 * unions done between clique sub-cells (side 0.99 eps / sqrt(dim)), each
 * pair scanned until its first joined pair: fast on dense inputs, where
 * every core point has thousands of neighbours.  Returns -4 when the extent
 * is too large for the sub-cell grid (more than 2^20 cells per axis). */
int oracle_cells_dbscan(const double* xyz, int64_t n, int dim, double eps,
                        int minpts, int nthreads, int32_t* labels,
                        uint8_t* core_out);

/* Checks given labels / core flags at m sampled points against the
 * definition, enumerating each sampled point's eps-ball with the predicate:
 * a core point shares its label with every core point it is joined to; a
 * non-core point has the smallest label among the core points whose
 * predicate reaches it (0 if none).  out[0] core-label violations, out[1]
 * border violations, out[2] predicate evaluations, out[3] sampled cores. */
int oracle_sample_check(const double* xyz, int64_t n, int dim, double eps,
                        const int32_t* labels, const uint8_t* core,
                        const int64_t* sample, int64_t m, int nthreads,
                        int64_t* out);

/* O(N^2) brute-force neighbour counts (self included), for pinning. */
int oracle_brute_counts(const double* xyz, int64_t n, int dim, double eps,
                        int32_t* counts);

/* Restatement of multicore-cpu/dbscan.cpp:56-209 (FP32, sqrt compare, self
 * excluded from the count, BFS expansion, O(N) OpenMP scan per expanded
 * point).  Timing baseline only.  Returns the number of clusters that
 * program reports (dbscan.cpp:84-101), or a negative value when out of
 * memory.  Its per-point labels are not a clustering: a core point's own row
 * includes itself, so the pass that labels it marks it again (MARKED = 0,
 * dbscan.cpp:187-199), and labels[] (>0 cluster, 0 otherwise) ends all 0. */
int oracle_multicore_dbscan(const double* xyz, int64_t n, int dim, double eps,
                            int minpts, int nthreads, int32_t* labels);

/* Restatement of test-results/cluster-check.cpp:41-199 in O(N) time.
 * out[0] = mismatched noise points, out[1] = agreeing noise points,
 * out[2] = fraction lost due to mismatches, out[3] = final score. */
int oracle_cluster_check(const int32_t* c1, const int32_t* c2, int64_t n,
                         double* out);

#ifdef __cplusplus
}
#endif
#endif
