// Internal device/host definitions shared by the HIP translation units.
//
// Data layout in HBM (all arrays live in the handle's arena, sized for the
// largest run seen so far and reused):
//   input   xyz       caller's row-major FP64 rows, instances back to back
//   sort    key[2]    u64 cell keys (ping-pong), val[2] i32 point ids
//   CSR     cell_start[C+1], cell_key[C], cell_inst[C], rows[C][25] (packed
//           runs of the 5^(D-1) neighbour lines, centre line first),
//           cell_ncore[C], cell_label[C] (chain), amask[C] (near adjacency)
//   points  (cell-sorted)  X[3][N] FP64 SoA (FP64 form only), G[3][N] u32 fixed-point
//           coordinates (29 fraction bits per cell, wrapping mod 8 cells),
//           orig[N], cell_of[N], core[N] u8, rec[N][R] + rec_n[N] (the
//           neighbours of each non-core point)
//   merge   parent[] chain union-find, compmin[N], pairs[waves][E] int2
//   output  keyv[N] (component min input index), flag[N] u8, rank[N]
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dclust {

constexpr int kWave = 64;
constexpr int32_t kUnclaimed = -1;
constexpr int32_t kNoKey = 0x7fffffff;
constexpr int kFixBits = 29;  // fixed-point fraction bits per cell width
constexpr int kMaxRows = 25;  // 5^(D-1) neighbour lines for D <= 3
constexpr int kRowCntShift = 28;  // packed row: first cell | cell count << 28
constexpr int32_t kRowCellMask = (1 << kRowCntShift) - 1;
constexpr int64_t kMaxCells = 1LL << kRowCntShift;
constexpr int kHeavy = 32;    // cells with more points are "heavy" (wavefront paths)
constexpr int kRecSlots = 16;  // neighbour records kept per non-core point (at most)

// Stencil lines, nearest first: (offset along axis 0, offset along axis 1)
// of line r for D = 3, offset along axis 0 for D = 2.  The mirror of line r
// is the line with the negated offsets (used to set the adjacency bit of
// the other cell of a pair).
#define DCLUST_ROW_OFF3                                                                     \
  {{0, 0},   {-1, 0}, {1, 0},   {0, -1},  {0, 1},  {-1, -1}, {-1, 1}, {1, -1}, {1, 1},       \
   {-2, 0},  {2, 0},  {0, -2},  {0, 2},   {-2, -1}, {-2, 1}, {2, -1}, {2, 1},  {-1, -2},     \
   {1, -2},  {-1, 2}, {1, 2},   {-2, -2}, {-2, 2}, {2, -2}, {2, 2}}
#define DCLUST_ROW_MIRROR3 \
  {0, 2, 1, 4, 3, 8, 7, 6, 5, 10, 9, 12, 11, 16, 15, 14, 13, 20, 19, 18, 17, 24, 23, 22, 21}
// row of the line with offsets (a - 2, b - 2) along axes 0 and 1 (3-D)
#define DCLUST_ROW_OF_OFF3                                                     \
  {{21, 13, 9, 14, 22}, {17, 5, 1, 6, 19}, {11, 3, 0, 4, 12}, {18, 7, 2, 8, 20}, \
   {23, 15, 10, 16, 24}}
#define DCLUST_ROW_OFF2 {0, -1, 1, -2, 2}
#define DCLUST_ROW_MIRROR2 {0, 2, 1, 4, 3}
static __constant__ int8_t kRowOff3[25][2] = DCLUST_ROW_OFF3;
static __constant__ int8_t kRowOff2[5] = DCLUST_ROW_OFF2;
static __constant__ int8_t kRowMirror3[25] = DCLUST_ROW_MIRROR3;
static __constant__ int8_t kRowMirror2[5] = DCLUST_ROW_MIRROR2;
static __constant__ int8_t kRowOfOff3[5][5] = DCLUST_ROW_OF_OFF3;

namespace tables {
constexpr int8_t off3[25][2] = DCLUST_ROW_OFF3;
constexpr int8_t mir3[25] = DCLUST_ROW_MIRROR3;
constexpr int8_t off2[5] = DCLUST_ROW_OFF2;
constexpr int8_t mir2[5] = DCLUST_ROW_MIRROR2;
constexpr int8_t rof3[5][5] = DCLUST_ROW_OF_OFF3;
constexpr bool mirrors_ok() {
  for (int r = 0; r < 25; ++r)
    if (off3[mir3[r]][0] != -off3[r][0] || off3[mir3[r]][1] != -off3[r][1]) return false;
  for (int a = 0; a < 5; ++a)
    for (int b = 0; b < 5; ++b)
      if (off3[rof3[a][b]][0] != a - 2 || off3[rof3[a][b]][1] != b - 2) return false;
  for (int r = 0; r < 5; ++r)
    if (off2[mir2[r]] != -off2[r]) return false;
  return true;
}
static_assert(mirrors_ok(), "stencil line mirror table");
}  // namespace tables

// Per-instance parameters and derived grid (device table, one per instance).
struct Inst {
  double mn[3];    // grid origin = bounding-box minimum
  double inv_w;    // 1 / cell width: cell(x) = floor(fl(fl(x - mn) * inv_w))
  double w;        // cell width
  double eps;      // EPS
  double eps2;     // fl(eps * eps), the oracle's threshold
  double maxabs;   // max |coordinate| (FP error bounds)
  float d2_in;     // FP32 form: D2 <  d2_in  -> neighbour
  float d2_out;    // FP32 form: D2 >  d2_out -> not a neighbour
  float d2_mid;    // FP32 form: rounded eps^2 in fixed-point units
  int32_t dim;
  int32_t minpts;
  int32_t r;       // PARTITION_SIZE (0 = auto)
  int32_t nk[3];   // cells per axis
  int32_t err;     // per-instance error code (DCLUST_E_*)
  int64_t key_base;            // first key of this instance's cell space
  int64_t line_base;           // first entry of this instance's line table
  int64_t nlines;              // possible lines (product of all but the last axis)
  int64_t pt_begin, pt_end;    // global point-index range
  int64_t xyz_off;             // element offset of the first coordinate
};

// Chunk of points of one instance (host-built work list for per-point
// kernels that need the instance, so no point->instance search is needed).
struct Chunk {
  int32_t inst;
  int32_t pad;
  int64_t begin, end;
};

// Control block (int64 counters in device memory).
enum Ctl : int {
  C_KEYBITS = 0,
  C_NCELL,
  C_NCHAIN,     // chains started (a chain id is its seed cell's index)
  C_NPAIR,
  C_Q0_N,
  C_Q1_N,
  C_Q0_HEAD,
  C_Q1_HEAD,
  C_NCORE,
  C_NNOISE,
  C_NCLUST,
  C_CAND,
  C_NBRC,
  C_BAND,
  C_FLIP,
  C_ASYM,
  C_EVAL,
  C_OVF,
  C_ERR,
  C_SORTPAR,  // number of executed radix scatters (buffer parity)
  C_BAND_CORE,
  C_POPS,       // cells expanded
  C_BUSY_SUM,  // expansion: sum over waves of their lifetime (us)
  C_BUSY_MAX,
  C_SEGS_LEFT,  // expansion waves whose seeding segment is not finished
  C_NLINES,     // total line-table entries
  C_DENSE,      // cells resolved as all-core without distance tests
  C_NHEAVY,     // heavy cells (listed in Dev::heavy)
  C_NADJPTS,    // core points of light cells (listed in Dev::adj_pts)
  C_NCORECELLS, // cells with core points (listed in Dev::core_cells)
  C_EDGES,      // adjacency bits of the expanded cells
  C_RECOVF,     // non-core points whose neighbour record overflowed
  C_ADJ_TESTS,  // pair tests of the adjacency lanes
  C_ADJ_FREE,   // adjacency bits set without a test (cell inside the eps ball)
  C_FAR_TESTS,  // pair tests of the far-pair pass
  C_FAR_JOINS,  // far pairs that joined two components
  C_LASTPASS,   // 1 + the last expansion pass that popped a cell
  C_POPS_MAX,   // most cells popped by one expansion wave in one pass
  C_SEEDS_MAX,  // most chains started by one expansion wave in one pass
  C_ABANDON,    // expansion left work pending in a run with a fixed pass
                // count: the kernels after it return at once
  C_IDLE,       // nothing pending for the current expansion pass
  C_ROUNDS_MAX, // most pop rounds of one expansion wave in one pass
  C_KEYOVF,     // the cell-key space of the whole call exceeds 2^63 keys
  // work counts of the neighbour kernels (their algorithmic bytes)
  C_CORE_GROUPS,  // k_core: groups of 5 stencil lines streamed
  C_CORE_LINES,   // k_core: non-empty stencil lines streamed
  C_ADJ_LINES,    // k_adj: non-empty near lines examined
  C_ADJ_CELLS,    // k_adj: cells of those lines
  C_NROWSFB,      // row work items left to the global-memory pass (Dev::rows_fb)
  C_NBORDERPTS,   // non-core points (listed in Dev::border_pts)
  C_NCTL = 49
};
static_assert(C_NBORDERPTS < C_NCTL, "control block too small");

// Cell-key space: the sort key of a cell is its instance's key base plus its
// linear index, and the radix sort orders keys of up to 63 bits, so the
// cells of all instances of a call must number at most 2^63.  Sums are
// saturating, so an overflowing total stays above the limit.
constexpr unsigned long long kMaxKeySpace = 1ull << 63;
__host__ __device__ __forceinline__ unsigned long long sat_add(unsigned long long a,
                                                                unsigned long long b) {
  const unsigned long long s = a + b;
  return s < a ? ~0ull : s;
}
// Radix-sort key width for a key space of `total` keys (total <= 2^63).
__host__ __device__ __forceinline__ int key_bits_for(unsigned long long total) {
  int bits = 1;
  while (bits < 63 && (1ull << bits) < total) ++bits;
  return bits;
}

// Statistics counters are striped: a wave adds to one of kStripes copies.
constexpr int kStripes = 64;

// A work list filled by many workgroups: kStripes sub-lists (workgroup b
// appends to sub-list b % kStripes, with its own counter), so that the
// appends do not serialise on one address.  Element i of the whole list is
// found through the prefix of the sub-list counts.  The counters live in
// the striped statistics array, so the host reads the total like a counter.
struct SList {
  int32_t* data;             // [kStripes][cap]
  unsigned long long* cnt;   // [kStripes]
  int64_t cap;
};

// Sort-pass bookkeeping: pass p is active when p*8 < key bits; its scatter
// runs unless every key has the same digit (then the data is already in
// order and the buffers are not swapped).
constexpr int kMaxPasses = 8;

// Bounding-box partials: chunk b reduces into stripe b % kBoxStripes of its
// instance (so that the atomics of many chunks do not meet on one address);
// the grid set-up reduces the stripes.
constexpr int kBoxStripes = 32;

// All device pointers of one run, passed to kernels by value.
struct Dev {
  int64_t n;
  int32_t ninst;
  int32_t nchunks;
  const double* xyz;
  Inst* inst;
  const Chunk* chunks;
  unsigned long long* bbox;  // [ninst][kBoxStripes][8] ordered-u64 min0..2 max0..2,
                             // max |x|, non-finite flag
  long long* ctl;
  long long* stat;  // [C_NCTL][kStripes] statistics counters (summed by the host)

  // radix sort
  uint64_t* key[2];
  int32_t* val[2];
  uint32_t* hist;   // [256][nblk] digit-major block histograms
  uint32_t* offs;   // [256][nblk] scanned offsets
  uint32_t* dtot;   // [256] digit totals of the current pass
  int32_t* pass_scatter;  // [kMaxPasses] 1 if the scatter of pass p ran

  // CSR
  int32_t* cell_start;  // [N+1]
  uint64_t* cell_key;   // [N]
  uint64_t* cell_xyz;   // [N] packed integer cell coordinates
  int32_t* cell_z;      // [N] last integer cell coordinate (position in its line)
  int32_t* cell_inst;   // [N]
  int32_t* cell_ncore;  // [N] core points per cell
  int32_t* cell_label;  // [N] chain owning the cell's core points
  unsigned long long* amask;  // [N] near adjacency: bit 5r+k <-> k-th cell of line r < 9
  int32_t* cell_comp;   // [N] component (root chain) of each core cell, -1 otherwise
  SList heavy;          // heavy cells (more than kHeavy points)
  SList adj_pts;        // core points of light cells (near-adjacency lanes)
  SList border_pts;     // non-core points (border-label lanes)
  SList core_cells;     // cells with core points (far-pair lanes)
  int32_t* rows;        // [kMaxRows][rstride] packed neighbour lines (first cell, count),
                        // line-major so that consecutive cells are adjacent
  int64_t rstride;      // >= number of cells
  int32_t* cell_of;     // [N] per sorted point
  int2* lines;          // [line_cap] first / end cell of every possible line
  int64_t line_cap;     // 0: no line table, rows found by binary search
  int32_t nbr_stats;    // 1: k_rows also sums the neighbourhood sizes

  // sorted point data
  double* X[3];         // FP64 form only (null in the FP32 form)
  int64_t* xoff;        // FP32 form: offset of each point's row in xyz
  uint32_t* G[3];
  int32_t* orig;
  uint8_t* core;
  int32_t* rec;         // [N][rec_stride] neighbours of non-core points
  uint8_t* rec_n;       // [N] records written (saturates at 255)
  int32_t rec_stride;   // record slots per point: min(minpts, kRecSlots)
  int32_t* parent;
  int32_t* compmin;
  int2* pairs;          // [waves][pair_per] collision pairs per expansion wave
  int32_t* pair_cnt;    // [waves] pairs recorded (may exceed pair_per)
  int64_t pair_cap;     // waves * pair_per
  int32_t pair_per;
  int32_t* q[2];
  int64_t* seg_pos;     // [waves] seeding cursor of each expansion wave
  int32_t* keyv;
  SList rows_fb;         // row work items (tile * 5 + line group) left to k_rows_global
  uint8_t* flag;
  int32_t* rank;
  int64_t* scan_tmp;  // block sums for scans

  // outputs (input order)
  int32_t* labels_out;
  uint8_t* core_out;
};

// ------------------------------------------------------------ helpers ----
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long ballot(bool p) {
  return __ballot(p);
}

__device__ __forceinline__ int mbcnt(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi(
      (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int rl_i(int v, int l) {
  return __builtin_amdgcn_readlane(v, l);
}
__device__ __forceinline__ uint32_t rl_u(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ float rl_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rl_d(double v, int l) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ int64_t rl_l(int64_t v, int l) {
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ int rfl_i(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// wave barrier + wavefront-scope fence: orders LDS traffic between lanes of
// one wave (seed-list pushes followed by pops).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <class T>
__device__ __forceinline__ T wave_min(T v) {
  for (int o = 32; o > 0; o >>= 1) {
    T u = __shfl_xor(v, o, 64);
    v = u < v ? u : v;
  }
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = 32; o > 0; o >>= 1) {
    T u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  return v;
}
// inclusive prefix sum over the wave
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = lane_id();
  for (int o = 1; o < 64; o <<= 1) {
    T u = __shfl_up(v, o, 64);
    if (l >= o) v += u;
  }
  return v;
}

// Aggregated atomic add of a wave-uniform sum (one atomic per wave).
__device__ __forceinline__ void wave_add(long long* p, long long v) {
  long long s = wave_sum(v);
  if (lane_id() == 0 && s != 0) atomicAdd((unsigned long long*)p, (unsigned long long)s);
}

// Statistics counter update: the wave's sum goes to one of kStripes copies
// of the counter, picked by the wave's position in the launch, so that a
// launch of many waves does not serialise on one address.  The host sums
// the copies.
// chain ids are seed cell indices: the id space is the cell count
__device__ __forceinline__ long long chain_ids(const Dev& d) { return d.ctl[C_NCELL]; }

// set when a run with a fixed number of expansion passes left work pending
__device__ __forceinline__ bool abandoned(const Dev& d) { return d.ctl[C_ABANDON] != 0; }

__device__ __forceinline__ void stat_add(const Dev& d, int ctr, long long v) {
  const long long s = wave_sum(v);
  if (lane_id() == 0 && s != 0) {
    const unsigned w = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kStripes - 1);
    atomicAdd((unsigned long long*)&d.stat[ctr * kStripes + w], (unsigned long long)s);
  }
}

// Appends `value` of the threads with `take` to the workgroup's sub-list of
// L, in thread order, with one atomic per workgroup.  Every thread of the
// workgroup calls it.
__device__ __forceinline__ void block_append(bool take, int32_t value, const SList& L) {
  __shared__ int s_cnt[16];
  __shared__ long long s_base;
  const int wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const unsigned long long m = ballot(take);
  if (lane_id() == 0) s_cnt[wv] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < nw; ++w) tot += s_cnt[w];
    s_base = tot ? (long long)atomicAdd(&L.cnt[blockIdx.x & (kStripes - 1)],
                                        (unsigned long long)tot)
                 : 0;
  }
  __syncthreads();
  long long base = s_base;
  for (int w = 0; w < wv; ++w) base += s_cnt[w];
  if (take) L.data[(int64_t)(blockIdx.x & (kStripes - 1)) * L.cap + base + mbcnt(m)] = value;
  __syncthreads();  // s_cnt / s_base are reused by the next call
}

// Prefix of the sub-list counts of L in LDS (s_pre[kStripes + 1]); returns
// the list length.  Every thread of the workgroup calls it.
__device__ __forceinline__ long long slist_prefix(const SList& L, long long* s_pre) {
  if (threadIdx.x < kStripes) {
    const long long v = (long long)L.cnt[threadIdx.x];
    const long long incl = wave_incl_scan(v);
    s_pre[threadIdx.x + 1] = incl;
    if (threadIdx.x == 0) s_pre[0] = 0;
  }
  __syncthreads();
  return s_pre[kStripes];
}

// Element i (< length) of the list.
__device__ __forceinline__ int32_t slist_at(const SList& L, const long long* s_pre, long long i) {
  int lo = 0, hi = kStripes - 1;  // last sub-list whose range starts at or before i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s_pre[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return L.data[(int64_t)lo * L.cap + (i - s_pre[lo])];
}

// Monotone map double -> u64 (for atomic min/max of doubles).
__device__ __forceinline__ unsigned long long ord_d(double x) {
  unsigned long long u = (unsigned long long)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ULL);
}
__host__ __device__ inline double unord_d(unsigned long long u) {
  unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffULL) : ~u;
  double d;
  __builtin_memcpy(&d, &b, sizeof(d));
  return d;
}

// Blocks of a launch are dealt to the 8 XCDs round-robin (block b runs on
// XCD b % 8).  Remap so that each XCD owns a contiguous range of logical
// blocks: neighbouring work then shares that XCD's L2.  G must be a
// multiple of 8.
__device__ __forceinline__ int xcd_block(int b, int G) {
  int per = G >> 3;
  return (b & 7) * per + (b >> 3);
}

// State of the one-step (CUDA-DClust) baseline, cell-sorted order.
struct OneStep {
  int32_t* label;     // [n] -1 unclassified, -2 noise, >= 0 chain
  int32_t* state;     // [n] 0 not expanded, 1 expanded, 2 expanded core
  int32_t* seeds;     // [blocks * S] per-chain seed lists
  int32_t* len;       // [blocks] seed-list lengths
  int2* contacts;     // [cap] (chain, chain) contacts between core points
  int32_t* ctr;       // [2] contacts recorded since the last host read (may
                      // exceed cap: the rest were not stored), spare
  int32_t blocks, S, cap, pad;
};

// The stencil lines of one cell: rows_of(d, c)[r] is line r (packed).
struct RowsOf {
  const int32_t* base;
  int64_t stride;
  __device__ __forceinline__ int32_t operator[](int r) const { return base[(int64_t)r * stride]; }
};
__device__ __forceinline__ RowsOf rows_of(const Dev& d, int64_t c) {
  return RowsOf{d.rows + c, d.rstride};
}

// Instance of a cell (a run of one instance reads the table entry directly,
// with no per-cell load).
__device__ __forceinline__ const Inst* inst_of_cell(const Dev& d, int64_t c) {
  return d.ninst == 1 ? d.inst : d.inst + d.cell_inst[c];
}

// Cell index along one axis, shared by key computation, gather and every
// search (the reference computes insert and search bins with different
// formulas, indexing.cu:114 vs indexing.cu:96-101; here there is one).
__device__ __forceinline__ double cell_t(double x, double mn, double inv_w) {
  return (x - mn) * inv_w;
}

// ---- chain union-find (lock-free; roots are the minimum id) ------------
// Invariant: par[x] <= x, and a link only ever replaces a root's self-pointer
// (by CAS) with a smaller id, so every value par[x] has held is an ancestor
// of x.  A plain load may therefore return a stale but still valid ancestor:
// uf_find_cached walks with plain (L2-cached) loads, which is enough to
// decide "same set" when two walks meet.  A CAS on a stale "root" fails and
// the retry walks with coherent (agent-scope) loads, which guarantees
// progress.
__device__ __forceinline__ int32_t uf_find(int32_t* par, int32_t x) {
  for (;;) {
    int32_t p = __hip_atomic_load(&par[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p == x) return x;
    int32_t gp = __hip_atomic_load(&par[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (gp != p)
      __hip_atomic_store(&par[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
  }
}

// Root of x by a walk that writes nothing.  The final flatten uses it (no
// unions running): a path-halving write from another thread's walk could
// otherwise replace a root that its owner has just stored with an inner
// ancestor, and a later read of "the root" would then get an inner node.
// The walk does not compress, so it is run on an already compressed forest.
__device__ __forceinline__ int32_t uf_root(const int32_t* par, int32_t x) {
  for (;;) {
    const int32_t p = par[x];
    if (p == x) return x;
    x = p;
  }
}

__device__ __forceinline__ int32_t uf_find_cached(int32_t* par, int32_t x) {
  for (;;) {
    int32_t p = par[x];
    if (p == x) return x;
    int32_t gp = par[p];
    if (gp != p) par[x] = gp;  // path halving: gp is an ancestor of x
    x = p;
  }
}

__device__ __forceinline__ void uf_union(int32_t* par, int32_t a, int32_t b) {
  a = uf_find_cached(par, a);
  b = uf_find_cached(par, b);
  for (;;) {
    if (a == b) return;
    if (a < b) {
      int32_t t = a;
      a = b;
      b = t;
    }
    int32_t old = atomicCAS(&par[a], a, b);
    if (old == a) return;
    a = uf_find(par, old);
    b = uf_find(par, b);
  }
}

}  // namespace dclust
