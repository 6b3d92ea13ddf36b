/*
 * dclust -- MI355X-native CUDA-DClust+ DBSCAN, C ABI (libdclust.so).
 *
 * The reference has no library API: its CUDA-DClust+ path is one program,
 * `main.exe <dataset-file>` (main.cu:35-46), configured by compile-time
 * macros (common.h:6-25), reading a CSV dataset (main.cu:469-498) and
 * running index build (indexing.cu:23-118), the seed-list expansion kernel
 * (dbscan.cu:254-387) and the host merge loop (dbscan.cu:25-187,
 * main.cu:353-403).  This header is the boundary that replaces that program
 * for callers: plain C types, no HIP types, error codes instead of exit().
 *
 * Labels: labels[i] = 0 for noise, 1..K for clusters, with cluster ids
 * numbered by the smallest input index of a core point in each cluster.
 * This is exactly the numbering of the FP64 sequential reference DBSCAN
 * (cpu-dbscan/dbscan_rtree_exp.cpp:360-406, noise -1 written as 0 as
 * test-results/cluster-check.cpp:120-139 reads it).  A border point takes
 * the smallest id among the clusters of the core points it is within eps of.
 *
 * Threading: a handle is used by one host thread at a time; distinct
 * handles may run concurrently.  Every call returns after its results are
 * complete (labels on the host for the host entry points, on the device for
 * the *_device entry points).
 */
#ifndef DCLUST_H
#define DCLUST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes (the reference exits instead: common.h:31-40,        ---- */
/* ---- main.cu:204-207, main.cu:469-474)                                ---- */
#define DCLUST_OK 0
#define DCLUST_E_PARAM -1    /* invalid parameter or non-finite input       */
#define DCLUST_E_BINWIDTH -2 /* cell width (extent / r) < eps, main.cu:204  */
#define DCLUST_E_IO -3       /* dataset / label file error, main.cu:469     */
#define DCLUST_E_HIP -4      /* HIP runtime error (see dclust_stats.hip_error) */
#define DCLUST_E_OOM -5      /* device or host allocation failed            */
#define DCLUST_E_CAPACITY -6 /* a size limit of this build was exceeded     */
#define DCLUST_E_NODEVICE -7 /* no usable GPU                               */

/* precision forms */
#define DCLUST_PRECISION_FP64 0       /* plain FP64 predicate, bit-exact     */
#define DCLUST_PRECISION_FP32_GUARD 1 /* FP32 distances, FP64 recheck in band */

/*
 * Parameters.  Names follow common.h:6-25 one to one; 0 selects the
 * library's tuned value where noted.
 *   dim             DIMENSION (2 or 3)
 *   eps             EPS (> 0)
 *   minpts          MINPTS (>= 1; |N(p)| counts p itself, dbscan.cu:353-369)
 *   partitions_r    PARTITION_SIZE: cells per axis of the shortest extent,
 *                   cell width = min extent / r; must give width >= eps
 *                   (main.cu:192-207).  0 = cell width eps (auto).
 *   seed_list_size  MAX_SEEDS: seed-list entries per chain (0 = auto).
 *   blocks          THREAD_BLOCKS: workgroups of the expansion kernel
 *                   (0 = fill the device at resident occupancy; rounded up
 *                   to a multiple of 8, one share per XCD).
 *   block_threads   THREAD_COUNT: threads per workgroup, 0..1024 (0 = 256).
 *                   Each wavefront of 64 lanes owns one seed list, so the
 *                   expansion runs the value rounded up to whole wavefronts,
 *                   at most 256 threads (common.h:14's 16 runs as 64).
 *   extra_collision EXTRA_COLLISION_SIZE: collision-pair buffer entries per
 *                   workgroup (0 = auto).  Pairs beyond it are merged inline.
 *   precision       DCLUST_PRECISION_*.
 * The launch values never change the labels; the shape a run actually used
 * is reported in dclust_stats.eff_*.
 */
typedef struct dclust_params {
  int32_t dim;
  int32_t minpts;
  double eps;
  int32_t partitions_r;
  int32_t seed_list_size;
  int32_t blocks;
  int32_t block_threads;
  int32_t extra_collision;
  int32_t precision;
} dclust_params;

/* Per-run diagnostics.  Times are HIP-event milliseconds on the handle's
 * stream (the reference's clock() buckets, main.cu:410-432).  A replay of a
 * captured run records only the events of t_total_ms, t_index_ms and
 * t_core_ms (each event inside a graph idles the device for a few
 * microseconds); t_rows_ms, t_adj_ms, t_expand_ms, t_merge_ms, t_far_ms,
 * t_border_ms, t_relabel_ms and diag[4] are then 0, unless the environment
 * variable DCLUST_PHASE_TIMES=1 is set when the call is made. */
typedef struct dclust_stats {
  double t_total_ms;   /* first kernel to labels resident on the device   */
  double t_index_ms;   /* bbox + cell keys + radix sort + CSR + rows        */
  double t_core_ms;    /* neighbour-count (core flag) kernel                */
  double t_adj_ms;     /* near-cell adjacency (core-core pairs within eps)  */
  double t_expand_ms;  /* seed-list expansion passes                        */
  double t_merge_ms;   /* union-find merge of collision pairs               */
  double t_far_ms;     /* far cell pairs across components (+ unions)       */
  double t_border_ms;  /* border assignment                                 */
  double t_relabel_ms; /* canonical relabel + output                        */
  int64_t n_points;
  int64_t n_cells;      /* non-empty grid cells                             */
  int64_t n_core;
  int64_t n_noise;
  int64_t n_clusters;
  int64_t n_chains;     /* chains started by the expansion kernel           */
  int64_t n_collisions; /* collision pairs recorded                         */
  int64_t n_overflow;   /* points pushed past a full seed list (corrected)  */
  int64_t n_passes;     /* expansion passes                                 */
  int64_t cand_total;   /* sum over cells of n_c * (points in the 5^D
                         * stencil); computed only when the environment
                         * variable DCLUST_NBR_STATS=1 is set (extra loads) */
  int64_t nbr_cells_total; /* sum over cells of non-empty stencil cells (same) */
  int64_t band_rechecks;   /* FP32 form: pairs re-decided in FP64           */
  int64_t band_flips;      /* FP32 form: rechecks where FP32 alone was wrong */
  int64_t asym_pairs;      /* core-core pairs with P(p,q) != P(q,p)         */
  int64_t pairs_evaluated; /* distance evaluations in the core kernel       */
  int64_t band_rechecks_core; /* guard-band rechecks inside the core kernel */
  int64_t rec_overflow;       /* non-core points whose neighbour record
                                 * overflowed (their border is found by a scan) */
  int32_t hip_error;       /* last hipError_t, 0 if none                    */
  int32_t key_bits;        /* bits of the sort key                          */
  /* diagnostics: [0] cells expanded, [1] cells resolved as all-core without
   * distance tests, [2] heavy cells (adjacency by one wavefront per cell),
   * [3] adjacent cell pairs (directed), [4] time of expansion pass 0 (us),
   * [5] sum over expansion wavefronts of their lifetime (us), [6] longest
   * lifetime (us), [7] expansion wavefronts, [8] pair tests of the
   * near-adjacency kernel, [9] adjacency bits set without a test, [10]
   * pair tests of the far-pair pass, [11] far pairs that joined two
   * components, [12] most cells popped by one expansion wave in a pass,
   * [13] most chains started by one expansion wave in a pass, [14] 1 if
   * the run was a replay of a captured graph, [15] most pop rounds (up to
   * 64 cells each) of one expansion wavefront in a pass */
  int64_t diag[16];
  /* launch shape the run used, in common.h's terms: THREAD_BLOCKS,
   * THREAD_COUNT (whole wavefronts), MAX_SEEDS, EXTRA_COLLISION_SIZE */
  int32_t eff_blocks;
  int32_t eff_block_threads;
  int32_t eff_seed_list_size;
  int32_t eff_extra_collision;
  double t_rows_ms;       /* the neighbour-line kernel alone (inside t_index) */
  int64_t n_core_cells;   /* cells holding core points                       */
  int64_t n_adj_points;   /* core points of light cells (near-adjacency lanes) */
  int64_t n_lines;        /* possible grid lines (line-table entries)        */
  /* work of the neighbour kernels (the units of their algorithmic bytes):
   * [0] k_core line groups streamed (5 lines each), [1] k_core non-empty
   * lines, [2] k_adj non-empty near lines, [3] k_adj cells on them */
  int64_t work[4];
} dclust_stats;

typedef struct dclust_handle dclust_handle;

/* common.h:13-25 defaults: DIMENSION 2, EPS 1.5, MINPTS 3, PARTITION_SIZE 3,
 * MAX_SEEDS 128, THREAD_BLOCKS 8, THREAD_COUNT 16, EXTRA_COLLISION_SIZE 256,
 * precision FP64.  The reference's 8 x 16 launch runs as 8 workgroups of
 * one wavefront each (exact, but far below what the device holds); set
 * blocks / block_threads / seed_list_size / extra_collision to 0 for the
 * library's tuned shape. */
int dclust_params_default(dclust_params* p);

/* Opens a handle on HIP device `device`.  Replaces the device set-up and
 * allocations of main.cu:72-152 (done lazily and reused across calls). */
int dclust_create(dclust_handle** h, int device);
void dclust_destroy(dclust_handle* h);

/* Cluster n points of dimension p->dim given as host row-major doubles
 * (the ImportDataset layout, main.cu:469-498).  labels_out: n int32 on the
 * host.  Replaces main.cu:155-432 (index, expand/merge loop, result).
 * stats may be NULL. */
int dclust_run(dclust_handle* h, const double* xyz, int64_t n,
               const dclust_params* p, int32_t* labels_out,
               dclust_stats* stats);

/* Same with device-resident input and output (HIP device pointers).
 *
 * Repeated calls: a call whose sizes, parameters, pointers and environment
 * switches all equal the previous call's is captured once into a hipGraph
 * and replayed from then on (the coordinates may change between calls).
 * A replay launches the whole run at once with no host read in between:
 * the radix passes of the previous run's key width and a fixed number of
 * expansion passes, one more than the previous run needed.  When the
 * replay finds that the key or the expansion needed more, the call reruns
 * eagerly and the next identical call captures a new graph.  The
 * environment variable DCLUST_GRAPH=0 turns replay off.  This applies to
 * every entry point that runs a clustering (dclust_run, the batched ones). */
int dclust_run_device(dclust_handle* h, const double* d_xyz, int64_t n,
                      const dclust_params* p, int32_t* d_labels_out,
                      dclust_stats* stats);

/* dclust_run_device without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, labels and statistics as dclust_run_device; every buffer
 * of the run comes from one 64 MiB array per GPU that the library's code
 * object carries, so the call makes no hipMalloc / hipFree.  Limit: the run
 * must fit that array -- about 230,000 3-D points in the FP32 form with
 * minpts <= 10 (about 210,000 in the FP64 form or with minpts >= 16, more
 * in 2-D); larger inputs return DCLUST_E_CAPACITY before any work.  The run is eager (no graph
 * replay).  All calls on one GPU share its array and each leaves its
 * contents there (dclust_last_core_flags reads them): calls that overlap on
 * one GPU, from any handles or threads, race. */
int dclust_run_device_noalloc(dclust_handle* h, const double* d_xyz, int64_t n,
                              const dclust_params* p, int32_t* d_labels_out,
                              dclust_stats* stats);

/* One-step expansion baseline: CUDA-DClust as the reference ships it
 * (cuda-dclust/dbscan_indexing.cu: kernel DBSCAN :529-637, host loop
 * :442-472, MonitorSeedPoints :639-781), on this library's grid index.
 * One seed per chain per launch, a host read of the seed-list lengths
 * after every launch, new seeds only when every list is empty.  It is here
 * to time the design the main path replaces, not as a clustering path of
 * its own.  Parameters: blocks = THREAD_BLOCKS (0: 256), block_threads =
 * THREAD_COUNT (0: 256; a multiple of 64 up to 1024), seed_list_size =
 * MAX_SEEDS (0: 128); the FP64 predicate whatever `precision` says.
 * Labels: core points and noise exactly as dclust_run_device gives them; a
 * border point gets the cluster of some core point within eps (not always
 * the smallest id).  Device pointers, one instance, no graph replay.
 * Stats: n_passes = kernel launches, n_chains = chains started,
 * n_collisions = contacts recorded, t_expand_ms = launches to labels.
 * Contacts go to a device list of max(2n, 2^21) entries that the host folds
 * into its union-find whenever half is used; DCLUST_E_CAPACITY if one
 * launch records more than the other half. */
int dclust_run_onestep_device(dclust_handle* h, const double* d_xyz, int64_t n,
                              const dclust_params* p, int32_t* d_labels_out,
                              dclust_stats* stats);

/* Cluster `count` independent instances in one pass over the device
 * (no counterpart in the reference, which clusters one dataset per process).
 * xyz[i]: n[i] rows of params[i].dim doubles on the host; labels[i]: n[i]
 * int32 on the host, numbered per instance.  inst_err (may be NULL)
 * receives a per-instance error code.  stats aggregates the batch. */
int dclust_run_batched(dclust_handle* h, int32_t count,
                       const double* const* xyz, const int64_t* n,
                       const dclust_params* params, int32_t* const* labels,
                       int32_t* inst_err, dclust_stats* stats);

/* Batched with device-resident data: d_xyz holds the instances' rows back
 * to back (instance i starts at element sum_{j<i} n[j]*dim[j]); d_labels
 * receives sum n[i] labels in the same order.  n/params are host arrays. */
int dclust_run_batched_device(dclust_handle* h, int32_t count,
                              const double* d_xyz, const int64_t* n,
                              const dclust_params* params, int32_t* d_labels,
                              int32_t* inst_err, dclust_stats* stats);

/* Core flags (1 = core) of the last run, in input order, copied to the
 * host; n must equal the last run's total point count. */
int dclust_last_core_flags(dclust_handle* h, uint8_t* core_out, int64_t n);

/* Launch shape for a problem of n points (seed list, blocks, threads,
 * collision buffer): fills the zero fields of *p from the shapes that
 * dclust_autotune_device has measured on this handle for the same dimension,
 * size class (log2 n) and precision, else from the library's built-in table
 * (shapes measured on an MI355X per dimension mix and size class; a run
 * whose seed list, blocks and threads are all 0 uses it too). */
int dclust_autotune(dclust_handle* h, int64_t n, dclust_params* p);

/* Empirical autotuner (no counterpart in the reference, whose launch shape is
 * fixed at compile time, common.h:6-25): clusters the given device-resident
 * instances (layout of dclust_run_batched_device; count = 1 for one dataset)
 * with every candidate shape -- seed-list length {256, 512, 1024, 2048} x
 * workgroup size {128, 256} x workgroups per CU {4, 8, 16} -- keeps the
 * fastest, writes it into params[0..count) and remembers it for
 * dclust_autotune.  Optional: the built-in table holds the shapes this sweep
 * found for representative problems.  The grid does not depend on
 * PARTITION_SIZE (the cell width is fixed by eps; r only validates the bin
 * width), so r is not a tuning dimension: labels and times are the same for
 * every valid r. */
int dclust_autotune_device(dclust_handle* h, int32_t count, const double* d_xyz,
                           const int64_t* n, dclust_params* params);

/* Reads a CSV dataset with the grammar of ImportDataset (main.cu:469-498):
 * lines of comma-separated numbers, all values flattened row-major, dim
 * values per point.  Unlike the reference (fixed DATASET_COUNT, garbage on
 * short files), the point count comes from the file and a field that does
 * not parse is an error.  With out == NULL only *n_points is computed.
 * max_points <= 0 means no limit. */
int dclust_import_dataset(const char* path, int32_t dim, int64_t max_points,
                          double* out, int64_t* n_points);

/* The dataset reader above also reads binary datasets (no counterpart in the
 * reference): a 24-byte header -- the 8 bytes "DCLUSTF8", the point count
 * (int64), the dimension (int32), 4 zero bytes -- then the points as
 * row-major little-endian doubles.  A file that starts with the header is
 * read as binary; its dimension must equal `dim` (else DCLUST_E_PARAM), its
 * size must match the count (else DCLUST_E_IO).  Text files are parsed by
 * up to 16 threads with the sequential result.  This writes the binary
 * format. */
int dclust_write_dataset_f64(const char* path, const double* xyz, int64_t n, int32_t dim);

/* Writes one label per line (0 = noise), the file format read by
 * test-results/cluster-check.cpp:201-224. */
int dclust_write_labels(const char* path, const int32_t* labels, int64_t n);

const char* dclust_strerror(int code);

/* Library version string. */
const char* dclust_version(void);

#ifdef __cplusplus
}
#endif
#endif
