// Grid-index build on gfx950: the MI355X replacement for the reference's
// tree index (indexing.cu:23-118, main.cu:166-340).
//
//   k_bbox        per-instance bounding box (main.cu:166-181 does it on the
//                 host, initialised to 0 / 999999999; here +-inf)
//   k_grid_setup  cell width, cells per axis, key space per instance
//                 (main.cu:184-207 bin width + eps check, main.cu:209-244 tree
//                 metadata: replaced by a flat grid)
//   k_cellkey     per-point cell key (insertData, indexing.cu:107-118)
//   radix sort    LDS radix sort of (key, id), stable (thrust::sort_by_key,
//                 main.cu:332-333)
//   CSR           wavefront ballot/mbcnt compaction of key changes into cell
//                 ranges (INDEXING_ADJUSTMENT's per-leaf equal_range,
//                 indexing.cu:23-51)
//   k_gather      cell-sorted SoA coordinates (FP64 + fixed-point)
//   k_rows        the 5^(D-1) stencil lines of each cell (cells of side
//                 eps/2, a 5^D stencil) as runs of point ranges
//                 (searchPoints' level-by-level descent, dbscan.cu:390-474)
#include "internal.h"
#include "kernels.h"

#include <algorithm>
#include <type_traits>
#include <cmath>

namespace dclust {

// ------------------------------------------------------------------ fill --
static int fill_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (int)b;
}

// Start of a run, one launch: control counters, statistics stripes, radix
// pass flags and the bounding-box stripes (per instance and stripe: 0..2 min
// as ordered keys starting at the largest key, 3..5 max, 6 max |x|, 7
// non-finite flag, all starting at 0).
__global__ void k_init_head(Dev d) {
  const int64_t nb = (int64_t)d.ninst * kBoxStripes * 8;
  const int64_t ns = (int64_t)C_NCTL * kStripes;
  const int64_t tot = nb + ns + C_NCTL + kMaxPasses + d.ninst;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < tot;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < nb)
      d.bbox[i] = (i & 7) < 3 ? ~0ull : 0ull;
    else if (i < nb + ns)
      d.stat[i - nb] = 0;
    else if (i < nb + ns + C_NCTL)
      d.ctl[i - nb - ns] = 0;
    else if (i < nb + ns + C_NCTL + kMaxPasses)
      d.pass_scatter[i - nb - ns - C_NCTL] = 0;
    else
      d.inst[i - nb - ns - C_NCTL - kMaxPasses].err = 0;  // (the table may be last run's)
  }
}

void launch_init_head(const Dev& d, hipStream_t s) {
  const int64_t tot = (int64_t)d.ninst * kBoxStripes * 8 + (int64_t)C_NCTL * kStripes + C_NCTL +
                      kMaxPasses + d.ninst;
  hipLaunchKernelGGL(k_init_head, dim3(fill_grid(tot)), dim3(256), 0, s, d);
}

// End of a run: the control counters, then their stripes ([C_NCTL], then
// [C_NCTL][kStripes]) into a host buffer (pinned and mapped), one workgroup.
// Inside a captured run this follows the last kernel with no gap, where copy
// nodes would each wait on a separate transfer.
__global__ __launch_bounds__(1024) void k_readback(Dev d, long long* out) {
  constexpr int tot = C_NCTL * (1 + kStripes);
  for (int i = threadIdx.x; i < tot; i += blockDim.x)
    out[i] = i < C_NCTL ? d.ctl[i] : d.stat[i - C_NCTL];
}

void launch_readback(const Dev& d, long long* out, hipStream_t s) {
  hipLaunchKernelGGL(k_readback, dim3(1), dim3(1024), 0, s, d, out);
}

// ------------------------------------------------------------------ bbox --

// One block per chunk (chunks never straddle instances).  Block reduction,
// then one ordered-u64 atomic min/max per axis.  Non-finite input sets the
// instance error.
// Points of a chunk per thread step: the loads of kChunkU points (strided
// by the block) are issued before any is used, so their latencies overlap.
constexpr int kChunkU = 4;

template <int D>
__device__ __forceinline__ void bbox_points(const double* p0, int64_t np, double (&mn)[3],
                                            double (&mx)[3], double& ma, bool& bad) {
  for (int64_t i0 = threadIdx.x; i0 < np; i0 += kChunkU * 256) {
    double v[kChunkU][D];
#pragma unroll
    for (int u = 0; u < kChunkU; ++u) {
      // past the end: the chunk's last point again (min and max unchanged)
      const int64_t i = i0 + u * 256 < np ? i0 + u * 256 : np - 1;
#pragma unroll
      for (int k = 0; k < D; ++k) v[u][k] = p0[i * D + k];
    }
#pragma unroll
    for (int u = 0; u < kChunkU; ++u)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        bad |= !isfinite(v[u][k]);
        mn[k] = fmin(mn[k], v[u][k]);
        mx[k] = fmax(mx[k], v[u][k]);
        ma = fmax(ma, fabs(v[u][k]));
      }
  }
}

__global__ __launch_bounds__(256) void k_bbox(Dev d) {
  const Chunk ch = d.chunks[blockIdx.x];
  const Inst& I = d.inst[ch.inst];
  const int dim = I.dim;
  double mn[3] = {INFINITY, INFINITY, INFINITY};
  double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  double ma = 0.0;
  bool bad = false;
  const double* p0 = d.xyz + I.xyz_off + (ch.begin - I.pt_begin) * dim;
  if (dim == 3)
    bbox_points<3>(p0, ch.end - ch.begin, mn, mx, ma, bad);
  else
    bbox_points<2>(p0, ch.end - ch.begin, mn, mx, ma, bad);
  __shared__ double s_red[4][7];
  __shared__ int s_bad;
  if (threadIdx.x == 0) s_bad = 0;
  for (int k = 0; k < 3; ++k) {
    mn[k] = wave_min(mn[k]);
    mx[k] = wave_max(mx[k]);
  }
  ma = wave_max(ma);
  bool wb = ballot(bad) != 0;
  __syncthreads();
  const int wv = threadIdx.x >> 6;
  if (lane_id() == 0) {
    for (int k = 0; k < 3; ++k) {
      s_red[wv][k] = mn[k];
      s_red[wv][3 + k] = mx[k];
    }
    s_red[wv][6] = ma;
    if (wb) atomicOr(&s_bad, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      for (int k = 0; k < 3; ++k) {
        s_red[0][k] = fmin(s_red[0][k], s_red[w][k]);
        s_red[0][3 + k] = fmax(s_red[0][3 + k], s_red[w][3 + k]);
        s_red[0][6] = fmax(s_red[0][6], s_red[w][6]);
      }
    unsigned long long* bb =
        d.bbox + ((int64_t)ch.inst * kBoxStripes + blockIdx.x % kBoxStripes) * 8;
    for (int k = 0; k < dim; ++k) {
      atomicMin(&bb[k], ord_d(s_red[0][k]));
      atomicMax(&bb[3 + k], ord_d(s_red[0][3 + k]));
    }
    atomicMax(&bb[6], ord_d(s_red[0][6]));
    if (s_bad) atomicOr((unsigned long long*)&bb[7], 1ULL);
  }
}

// ------------------------------------------------------------ grid setup --
// Cells are cubes of width w = (eps/2)*(1+2^-20).  Two consequences:
//  * every cell is a clique: its diagonal sqrt(D)*w <= 0.87*eps, so any two
//    of its points are neighbours with a wide margin (FP64 rounding of the
//    box test and of d^2 is far below that margin);
//  * every eps-neighbour of a point lies in the 5^D cells around its cell
//    (|dx| <= eps < 2w, and the cell function is monotone, so the
//    rounding of floor(fl(fl(x-mn)*inv_w)) -- below 5*2^-53 per cell --
//    cannot move a neighbour further than the 2^-20 margin allows).
// PARTITION_SIZE r is validated as the reference does (min extent / r must
// be >= eps, main.cu:204-207, else DCLUST_E_BINWIDTH); it does not change
// the grid, whose resolution is fixed by eps (DBSCAN labels do not depend on
// the index).  Axes are limited to 2^21 cells (3-D) / 2^31 (2-D), i.e. an
// extent of about 2^20 eps; larger inputs get DCLUST_E_CAPACITY.
//
// FP32 form: distances are formed from fixed-point coordinates in units of
// w*2^-29 (29 fraction bits, wrapping every 8 cells; stencil differences
// stay below 3 cells < 2^31 units).  The FP32 squared distance D2 differs
// from the exact squared distance (same units) by at most 5*2^-24 relative
// (int->float conversion, square, two fused adds) plus the quantisation of
// the coordinates.  Pairs with |D2 - eps^2| <= band are re-decided with the
// exact FP64 predicate; band is 8 ulp = 2^-20 relative unless the bound
// above (plus the box-test slack from coordinate magnitude) requires more,
// in which case it is widened to twice the bound.
//
// Per instance this also sets the line-table base: a line is the set of
// cells sharing all coordinates but the last; the dense line table (one
// entry per possible line) locates a neighbour line in O(1).
__global__ void k_grid_setup(Dev d) {
  __shared__ unsigned long long s_scan[1024];
  __shared__ long long s_scanl[1024];
  __shared__ unsigned long long s_carry;
  __shared__ long long s_carryl;
  __shared__ int s_err;
  if (threadIdx.x == 0) {
    s_carry = 0;
    s_carryl = 0;
    s_err = 0;
  }
  __syncthreads();
  for (int base = 0; base < d.ninst; base += blockDim.x) {
    const int i = base + threadIdx.x;
    unsigned long long cells = 0;  // at most 2^63 (3 axes of at most 2^21 cells)
    long long lines = 0;
    if (i < d.ninst) {
      Inst I = d.inst[i];
      unsigned long long bb[8];
      {
        const unsigned long long* bs = d.bbox + (int64_t)i * kBoxStripes * 8;
#pragma unroll
        for (int k = 0; k < 8; ++k) bb[k] = k < 3 ? ~0ull : 0ull;
#pragma unroll 4
        for (int st = 0; st < kBoxStripes; ++st)
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const unsigned long long v = bs[st * 8 + k];
            bb[k] = k < 3 ? (v < bb[k] ? v : bb[k]) : (v > bb[k] ? v : bb[k]);
          }
      }
      const int dim = I.dim;
      double mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
      double emin = INFINITY;
      bool empty = I.pt_end <= I.pt_begin;
      for (int k = 0; k < dim; ++k) {
        mn[k] = empty ? 0.0 : unord_d(bb[k]);
        mx[k] = empty ? 0.0 : unord_d(bb[3 + k]);
        emin = fmin(emin, mx[k] - mn[k]);
      }
      I.maxabs = empty ? 0.0 : unord_d(bb[6]);
      if (bb[7]) {
        // DCLUST_E_PARAM: non-finite input.  The instance keeps a one-cell
        // grid (every point in cell 0) so that every kernel stays in bounds;
        // its result is discarded.
        I.err = -1;
        for (int k = 0; k < dim; ++k) mn[k] = mx[k] = 0.0;
        I.maxabs = 0.0;
      }
      if (I.r > 0) {
        double wr = emin / (double)I.r;
        if (!(wr >= I.eps) && I.err == 0 && !empty) I.err = -2;  // DCLUST_E_BINWIDTH
      }
      // coordinates so large that FP64 cannot resolve eps
      if (I.maxabs > I.eps * 0x1p40 && I.err == 0) I.err = -1;
      double w = 0.5 * I.eps * (1.0 + 0x1p-20);
      const double cap = dim == 3 ? 2097152.0 : 2147483647.0;
      for (;;) {
        // an oversized input keeps a valid (coarser) grid so that every
        // kernel stays in bounds; its result is discarded (error set)
        double inv = 1.0 / w;
        bool ok = true;
        for (int k = 0; k < dim; ++k)
          if (floor((mx[k] - mn[k]) * inv) + 1.0 > cap) ok = false;
        if (ok) break;
        if (I.err == 0) I.err = -6;  // DCLUST_E_CAPACITY
        w *= 2.0;
      }
      I.w = w;
      I.inv_w = 1.0 / w;
      cells = 1;
      lines = 1;
      for (int k = 0; k < 3; ++k) {
        I.mn[k] = k < dim ? mn[k] : 0.0;
        I.nk[k] = 1;
        if (k < dim) {
          double t = floor(cell_t(mx[k], mn[k], I.inv_w));
          I.nk[k] = (int32_t)t + 1;
          cells *= (unsigned long long)I.nk[k];
          if (k < dim - 1) lines *= (long long)I.nk[k];
        }
      }
      I.nlines = lines;
      // FP32 guard band (units of w * 2^-29)
      const double gs = I.inv_w * 0x1p29;
      const double es = I.eps * gs;
      const double e2 = es * es;
      const double u = 0x1p-24;
      double bound = 5.0 * u * (1.0 + 4.0 * u);
      bound += 2.0 * 1.7320508075688772 * 2.5 / es + 6.25 * 3.0 / e2;  // quantisation
      bound += 8.0 * 0x1p-52 * (I.maxabs / I.eps + 1.0);             // box slack
      double band = 0x1p-20;
      if (bound > 0.5 * band) band = 2.0 * bound;
      float fin = (float)(e2 * (1.0 - band));
      if ((double)fin > e2 * (1.0 - band)) fin = __int_as_float(__float_as_int(fin) - 1);
      float fout = (float)(e2 * (1.0 + band));
      if ((double)fout < e2 * (1.0 + band)) fout = __int_as_float(__float_as_int(fout) + 1);
      I.d2_in = fin;
      I.d2_out = fout;
      I.d2_mid = (float)e2;
      d.inst[i] = I;
      if (I.err) atomicOr(&s_err, 1);
    }
    s_scan[threadIdx.x] = cells;
    s_scanl[threadIdx.x] = lines;
    __syncthreads();
    // inclusive Hillis-Steele scans over the block
    // (the key-space sums saturate: a total beyond 2^63 fails the call below)
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
      unsigned long long v = threadIdx.x >= (unsigned)o ? s_scan[threadIdx.x - o] : 0;
      long long vl = threadIdx.x >= (unsigned)o ? s_scanl[threadIdx.x - o] : 0;
      __syncthreads();
      s_scan[threadIdx.x] = sat_add(s_scan[threadIdx.x], v);
      s_scanl[threadIdx.x] += vl;
      __syncthreads();
    }
    if (i < d.ninst) {
      d.inst[i].key_base = (int64_t)(sat_add(s_carry, s_scan[threadIdx.x]) - cells);
      d.inst[i].line_base = s_carryl + s_scanl[threadIdx.x] - lines;
    }
    __syncthreads();
    if (threadIdx.x == blockDim.x - 1) {
      s_carry = sat_add(s_carry, s_scan[threadIdx.x]);
      s_carryl += s_scanl[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const unsigned long long total = s_carry;
    d.ctl[C_ERR] = s_err;
    d.ctl[C_NLINES] = s_carryl;
    if (total > kMaxKeySpace) {
      // the keys would not fit the sort: the whole call fails (instances of a
      // batch share one key space), and every later kernel returns at once
      d.ctl[C_KEYOVF] = 1;
      d.ctl[C_ERR] = 1;
      d.ctl[C_ABANDON] = 1;
      d.ctl[C_KEYBITS] = 8;
    } else {
      d.ctl[C_KEYBITS] = key_bits_for(total);
    }
  }
}

// --------------------------------------------------------------- cellkey --
// Sort keys are 32-bit when the key space of the call fits (at most four
// 8-bit passes), else 64-bit; the key buffers are viewed as either.  The
// host picks the width from the key bits it knows (the grid set-up's, or
// the previous run's for a captured graph: a replay whose keys are wider
// than its passes cover is abandoned by k_csr_count, so a truncated key is
// never used).
template <bool K32>
using SortKey = typename std::conditional<K32, uint32_t, uint64_t>::type;
template <bool K32>
__device__ __forceinline__ SortKey<K32>* key_buf(const Dev& d, int b) {
  return reinterpret_cast<SortKey<K32>*>(d.key[b]);
}

__device__ __forceinline__ int32_t cell_index(double x, const Inst& I, int k) {
  double t = floor(cell_t(x, I.mn[k], I.inv_w));
  int32_t c = (t >= 0.0) ? (int32_t)fmin(t, (double)(I.nk[k] - 1)) : 0;
  return c;
}

template <int D, bool K32>
__device__ __forceinline__ void cellkey_points(const Dev& d, const Inst& I, const Chunk& ch) {
  const double* p0 = d.xyz + I.xyz_off + (ch.begin - I.pt_begin) * D;
  const int64_t np = ch.end - ch.begin;
  for (int64_t i0 = threadIdx.x; i0 < np; i0 += kChunkU * 256) {
    double v[kChunkU][D];
#pragma unroll
    for (int u = 0; u < kChunkU; ++u) {
      const int64_t i = i0 + u * 256 < np ? i0 + u * 256 : np - 1;
#pragma unroll
      for (int k = 0; k < D; ++k) v[u][k] = p0[i * D + k];
    }
#pragma unroll
    for (int u = 0; u < kChunkU; ++u) {
      const int64_t i = i0 + u * 256;
      if (i < np) {
        int64_t lin = cell_index(v[u][0], I, 0);
#pragma unroll
        for (int k = 1; k < D; ++k) lin = lin * I.nk[k] + cell_index(v[u][k], I, k);
        key_buf<K32>(d, 0)[ch.begin + i] = (SortKey<K32>)(I.key_base + lin);
        d.val[0][ch.begin + i] = (int32_t)(ch.begin + i);
      }
    }
  }
}

template <bool K32>
__global__ __launch_bounds__(256) void k_cellkey(Dev d) {
  const Chunk ch = d.chunks[blockIdx.x];
  // a copy: the key stores could alias the instance table as far as the
  // compiler knows, which would reload its fields after every store
  const Inst I = d.inst[ch.inst];
  if (I.dim == 3)
    cellkey_points<3, K32>(d, I, ch);
  else
    cellkey_points<2, K32>(d, I, ch);
}

// ------------------------------------------------------------ radix sort --
// LSD radix sort, 8-bit digits, 2048-key tiles (256 threads x 8 keys; with
// 64-bit keys the scatter's 31 KiB of LDS lets five workgroups share a CU,
// and 4096-key tiles were slower, 1024-key tiles too), or 4096-key tiles for
// 32-bit keys where they need fewer rounds of workgroups (sort_items).
// Buffer parity: the number of scatters executed so far.
constexpr int RS_T = 256;
constexpr int RS_I = 8;
constexpr int RS_TILE = RS_T * RS_I;
// Keys per thread: 8 (2048-key tiles), or 16 (4096-key tiles) for 32-bit
// keys when that takes fewer rounds of resident workgroups (the histogram
// buffers are sized for the smaller tiles; see sort_items).

__device__ __forceinline__ int sort_parity(const Dev& d, int pass) {
  int p = 0;
  for (int q = 0; q < pass; ++q) p += d.pass_scatter[q];
  return p & 1;
}

// peers of this lane with the same 8-bit digit among `valid` lanes
__device__ __forceinline__ unsigned long long match8(uint32_t dg, bool valid) {
  unsigned long long peers = ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    bool bit = (dg >> b) & 1;
    unsigned long long m = ballot(bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

__device__ __forceinline__ bool pass_active(const Dev& d, int pass) {
  return pass * 8 < (int)d.ctl[C_KEYBITS];
}

template <bool K32, int NI>
__global__ __launch_bounds__(RS_T) void k_rs_hist(Dev d, int pass, int nblk) {
  if (!pass_active(d, pass)) return;
  const SortKey<K32>* keys = key_buf<K32>(d, sort_parity(d, pass));
  const int shift = pass * 8;
  const int64_t t0 = (int64_t)blockIdx.x * (RS_T * NI);
  // all of the thread's keys are loaded at once (an index past the end reads
  // key 0 and is not counted), so their latencies overlap
  uint32_t dg[NI];
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    const int64_t i = t0 + r * RS_T + threadIdx.x;
    dg[r] = (uint32_t)(keys[i < d.n ? i : 0] >> shift) & 255u;
  }
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    const bool v = t0 + r * RS_T + threadIdx.x < d.n;
    const unsigned long long peers = match8(dg[r], v);
    if (v && mbcnt(peers) == 0) atomicAdd(&h[dg[r]], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  d.hist[(int64_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
  // the pass runs unless the row scan finds one digit holding every key
  // (no other block of this launch reads the flag of this pass)
  if (blockIdx.x == 0 && threadIdx.x == 0) d.pass_scatter[pass] = 1;
}

// Per-digit row scans, one block per digit: offs[d][b] = keys of digit d in
// the blocks before b; the block also publishes its digit's total (the
// scatter derives the digit bases from the 256 totals).  A pass in which one
// digit holds every key leaves the data in order: that digit's block unmarks
// the pass, so its scatter is skipped and the buffers are not swapped.
__global__ __launch_bounds__(256) void k_rs_rowscan(Dev d, int pass, int nblk) {
  if (!pass_active(d, pass)) return;
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  const uint32_t* row = d.hist + (int64_t)blockIdx.x * nblk;
  uint32_t* out = d.offs + (int64_t)blockIdx.x * nblk;
  // eight chunks' loads in flight per thread, then the chunks scanned in turn
  constexpr int RU = 8;
  uint32_t xv[RU];
  for (int b0 = 0; b0 < nblk; b0 += blockDim.x) {
    const int u = (b0 / 256) % RU;
    if (u == 0) {
#pragma unroll
      for (int q = 0; q < RU; ++q) {
        const int bq = b0 + q * 256 + threadIdx.x;
        xv[q] = row[bq < nblk ? bq : nblk - 1];
      }
    }
    int b = b0 + threadIdx.x;
    uint32_t x = 0;
#pragma unroll
    for (int q = 0; q < RU; ++q) x = q == u ? xv[q] : x;
    x = b < nblk ? x : 0u;
    uint32_t inc = wave_incl_scan(x);
    if (lane_id() == 63) s_w[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t pre = s_carry;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) pre += s_w[w];
    if (b < nblk) out[b] = pre + inc - x;
    __syncthreads();
    if (threadIdx.x == 0) s_carry += s_w[0] + s_w[1] + s_w[2] + s_w[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    d.dtot[blockIdx.x] = s_carry;
    if ((int64_t)s_carry == d.n) d.pass_scatter[pass] = 0;  // every key has this digit
  }
}

// Stable scatter.  Each wave owns a contiguous run of 64 * NI keys of the
// tile and ranks them round by round (64 consecutive keys per round, lanes in
// order) against counters of its own in LDS, with no barrier between rounds
// (a wave's LDS accesses complete in order); one barrier later, each digit's
// counts of the waves before it and the tile's local starts turn the ranks
// into positions, so equal digits keep their input order.  The tile is
// staged in LDS in local digit order and written out coalesced per digit run.
template <bool K32, int NI>
__global__ __launch_bounds__(RS_T) void k_rs_scatter(Dev d, int pass, int nblk) {
  using KT = SortKey<K32>;
  constexpr int TILE = RS_T * NI;
  if (!pass_active(d, pass) || !d.pass_scatter[pass]) return;
  const int par = sort_parity(d, pass);
  const KT* kin = key_buf<K32>(d, par);
  const int32_t* vin = d.val[par];
  KT* kout = key_buf<K32>(d, par ^ 1);
  int32_t* vout = d.val[par ^ 1];
  const int shift = pass * 8;
  __shared__ KT s_k[TILE];
  __shared__ int32_t s_v[TILE];
  __shared__ uint32_t s_lstart[256], s_goff[256];
  __shared__ uint32_t s_wc[4][256];
  __shared__ uint32_t s_ws[4], s_wt[4];
  const int tid = threadIdx.x, wv = tid >> 6;
  const int64_t t0 = (int64_t)blockIdx.x * TILE;
  const int64_t w0 = t0 + (int64_t)wv * (64 * NI) + lane_id();
  // the wave's keys and values are loaded at once, before the digit
  // offsets, so that their latencies overlap (an index past the end reads
  // element 0, which is not ranked)
  KT kr[NI];
  int32_t vr[NI];
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    const int64_t i = w0 + r * 64;
    const int64_t ic = i < d.n ? i : 0;
    kr[r] = kin[ic];
    vr[r] = vin[ic];
  }
  const uint32_t c = d.hist[(int64_t)tid * nblk + blockIdx.x];
  const uint32_t off = d.offs[(int64_t)tid * nblk + blockIdx.x];
  const uint32_t tot = d.dtot[tid];
#pragma unroll
  for (int w = 0; w < 4; ++w) s_wc[w][tid] = 0;
  // the tile's local digit starts (an exclusive scan of its digit counts)
  // and the digit bases (an exclusive scan of the pass's digit totals)
  const uint32_t inc = wave_incl_scan(c);
  const uint32_t inct = wave_incl_scan(tot);
  if (lane_id() == 63) {
    s_ws[wv] = inc;
    s_wt[wv] = inct;
  }
  __syncthreads();
  uint32_t lst = inc - c, base = inct - tot;
  for (int w = 0; w < wv; ++w) {
    lst += s_ws[w];
    base += s_wt[w];
  }
  s_lstart[tid] = lst;
  s_goff[tid] = base + off;
  uint32_t lr[NI];
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    const bool v = w0 + r * 64 < d.n;
    const uint32_t dg = (uint32_t)(kr[r] >> shift) & 255u;
    const unsigned long long peers = match8(dg, v);
    const uint32_t rank = mbcnt(peers);
    const uint32_t base = s_wc[wv][dg];
    if (v && rank == 0) s_wc[wv][dg] = base + (uint32_t)__popcll(peers);
    lr[r] = base + rank;
  }
  __syncthreads();
  {
    // position of wave w's first key with digit tid
    const uint32_t a0 = s_wc[0][tid], a1 = s_wc[1][tid], a2 = s_wc[2][tid];
    s_wc[0][tid] = lst;
    s_wc[1][tid] = lst + a0;
    s_wc[2][tid] = lst + a0 + a1;
    s_wc[3][tid] = lst + a0 + a1 + a2;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    if (w0 + r * 64 < d.n) {
      const uint32_t dg = (uint32_t)(kr[r] >> shift) & 255u;
      const uint32_t lp = s_wc[wv][dg] + lr[r];
      s_k[lp] = kr[r];
      s_v[lp] = vr[r];
    }
  }
  __syncthreads();
  // write-out: every slot is read (past the tile's end the reads are
  // unused), the stores are guarded
  const int64_t tn = (d.n - t0) < TILE ? (d.n - t0) : TILE;
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    const int l = r * RS_T + tid;
    kr[r] = s_k[l];
    vr[r] = s_v[l];
  }
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    const int l = r * RS_T + tid;
    const uint32_t dg = (uint32_t)(kr[r] >> shift) & 255u;
    const int64_t g = (int64_t)s_goff[dg] + (l - (int)s_lstart[dg]);
    if (l < tn) {
      kout[g] = kr[r];
      vout[g] = vr[r];
    }
  }
}

// ------------------------------------------------------- CSR compaction --
// Flag = sorted key differs from its predecessor.  Pass 1 counts flags per
// 2048-key tile with ballot/popcount; a single block scans the tile counts;
// pass 2 ranks flags inside the tile with mbcnt and writes the cell ranges.
constexpr int CS_T = 256;
constexpr int CS_I = 8;
constexpr int CS_TILE = CS_T * CS_I;

template <bool K32>
__device__ __forceinline__ const SortKey<K32>* sorted_keys(const Dev& d) {
  return key_buf<K32>(d, sort_parity(d, kMaxPasses));
}

// Sorted keys i and i - 1 of each of a thread's rounds, loaded before any
// is used (indices clamped into the array) so that the loads are in flight
// together; key i starts a cell when i < n and it differs from key i - 1.
template <int NI, typename KT>
__device__ __forceinline__ void load_key_pairs(const Dev& d, const KT* key, int64_t t0,
                                               uint64_t (&k)[NI], uint64_t (&kp)[NI]) {
#pragma unroll
  for (int r = 0; r < NI; ++r) {
    const int64_t i = t0 + r * CS_T + threadIdx.x;
    const int64_t ic = i < d.n ? i : 0;
    k[r] = key[ic];
    kp[r] = key[ic > 0 ? ic - 1 : 0];
  }
}

// (also, after a fixed number of radix passes: a key wider than they cover
// leaves the keys unsorted, so the run is abandoned: every later kernel
// returns at once and the host runs the call again; counting unsorted keys
// here is harmless)
template <bool K32>
__global__ __launch_bounds__(CS_T) void k_csr_count(Dev d, int passes) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && passes < kMaxPasses &&
      d.ctl[C_KEYBITS] > 8 * passes)
    d.ctl[C_ABANDON] = 1;
  if (abandoned(d)) return;
  const auto* key = sorted_keys<K32>(d);
  const int64_t t0 = (int64_t)blockIdx.x * CS_TILE;
  uint64_t kr[CS_I], kp[CS_I];
  load_key_pairs<CS_I>(d, key, t0, kr, kp);
  int c = 0;
#pragma unroll
  for (int r = 0; r < CS_I; ++r) {
    const int64_t i = t0 + r * CS_T + threadIdx.x;
    c += __popcll(ballot(i < d.n && (i == 0 || kr[r] != kp[r])));
  }
  __shared__ int s[4];
  if (lane_id() == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) d.scan_tmp[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// exclusive scan of nb int64 values in place by one block; writes the total
// to *total (may be null)
__global__ __launch_bounds__(1024) void k_scan_top(int64_t* v, int nb, long long* total) {
  __shared__ int64_t s[1024];
  const int per = (nb + 1023) / 1024;
  const int b0 = threadIdx.x * per;
  int64_t sum = 0;
  for (int j = 0; j < per; ++j)
    if (b0 + j < nb) sum += v[b0 + j];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int64_t u = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += u;
    __syncthreads();
  }
  int64_t run = s[threadIdx.x] - sum;
  for (int j = 0; j < per; ++j)
    if (b0 + j < nb) {
      int64_t x = v[b0 + j];
      v[b0 + j] = run;
      run += x;
    }
  if (threadIdx.x == 1023 && total) *total = s[1023];
}

// instance owning a key: last i with key_base <= key
__device__ __forceinline__ int32_t inst_of_key(const Dev& d, uint64_t key) {
  int lo = 0, hi = d.ninst - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if ((uint64_t)d.inst[mid].key_base <= key) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Integer cell coordinates of a key, decoded once per cell: 21 bits per
// axis in 3-D (at most 2^21 cells per axis), 32 bits per axis in 2-D.
__device__ __forceinline__ uint64_t pack_cell(const Inst& I, uint64_t key) {
  int64_t lin = (int64_t)(key - (uint64_t)I.key_base);
  uint64_t kc[3] = {0, 0, 0};
  for (int k = I.dim - 1; k >= 0; --k) {
    kc[k] = (uint64_t)(lin % I.nk[k]);
    lin /= I.nk[k];
  }
  return I.dim == 3 ? (kc[0] | (kc[1] << 21) | (kc[2] << 42)) : (kc[0] | (kc[1] << 32));
}

__device__ __forceinline__ void unpack_cell(int dim, uint64_t pc, int32_t* kc) {
  if (dim == 3) {
    kc[0] = (int32_t)(pc & 0x1fffff);
    kc[1] = (int32_t)((pc >> 21) & 0x1fffff);
    kc[2] = (int32_t)(pc >> 42);
  } else {
    kc[0] = (int32_t)(pc & 0xffffffffu);
    kc[1] = (int32_t)(pc >> 32);
    kc[2] = 0;
  }
}

template <bool K32>
__global__ __launch_bounds__(CS_T) void k_csr_write(Dev d) {
  if (abandoned(d)) return;
  const auto* key = sorted_keys<K32>(d);
  const int64_t t0 = (int64_t)blockIdx.x * CS_TILE;
  __shared__ int s_w[4];
  __shared__ int64_t s_run;
  const int wv = threadIdx.x >> 6;
  uint64_t kr[CS_I], kp[CS_I];
  load_key_pairs<CS_I>(d, key, t0, kr, kp);
  // every round's instance and packed cell coordinates, read from the
  // instance table before the first store (the stores could alias it as far
  // as the compiler knows, which would hold each round's reads behind the
  // previous round's stores)
  int32_t ir[CS_I], zr[CS_I];
  uint64_t pcr[CS_I];
#pragma unroll
  for (int r = 0; r < CS_I; ++r) {
    ir[r] = 0;
    zr[r] = 0;
    pcr[r] = 0;
    const int64_t i = t0 + r * CS_T + threadIdx.x;
    if (i < d.n && (i == 0 || kr[r] != kp[r])) {
      ir[r] = inst_of_key(d, kr[r]);
      const Inst& I = d.inst[ir[r]];
      pcr[r] = pack_cell(I, kr[r]);
      zr[r] = (int32_t)(I.dim == 3 ? pcr[r] >> 42 : pcr[r] >> 32);
    }
  }
  if (threadIdx.x == 0) s_run = d.scan_tmp[blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < CS_I; ++r) {
    int64_t i = t0 + r * CS_T + threadIdx.x;
    bool inr = i < d.n;
    const uint64_t k = kr[r];
    const bool f = inr && (i == 0 || k != kp[r]);
    unsigned long long m = ballot(f);
    if (lane_id() == 0) s_w[wv] = __popcll(m);
    __syncthreads();
    int64_t pre = s_run;
    for (int w = 0; w < wv; ++w) pre += s_w[w];
    int64_t cid = pre + mbcnt(m) + (f ? 1 : 0) - 1;  // inclusive - 1
    if (inr) {
      d.cell_of[i] = (int32_t)cid;
      if (f) {
        const uint64_t pc = pcr[r];
        d.cell_start[cid] = (int32_t)i;
        d.cell_key[cid] = k;
        d.cell_inst[cid] = ir[r];
        d.cell_xyz[cid] = pc;
        d.cell_z[cid] = zr[r];
      }
      if (i == d.n - 1) {
        d.cell_start[cid + 1] = (int32_t)d.n;
        d.ctl[C_NCELL] = cid + 1;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) s_run += s_w[0] + s_w[1] + s_w[2] + s_w[3];
    __syncthreads();
  }
}

// ---------------------------------------------------------------- gather --
// One point: its whole row is loaded before anything is stored (a store
// could alias the row as far as the compiler knows, which would otherwise
// hold each coordinate's random-access load until the previous store).
template <int D>
__device__ __forceinline__ void gather_point(const Dev& d, const Inst& I, int64_t i, int32_t src) {
  const int64_t off = I.xyz_off + ((int64_t)src - I.pt_begin) * D;
  const double* p = d.xyz + off;
  double x[D], mn[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    x[k] = p[k];
    mn[k] = I.mn[k];
  }
  const double inv_w = I.inv_w;
  if (!d.X[0]) d.xoff[i] = off;  // the FP32 form's way back to FP64 rows
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (d.X[0]) d.X[k][i] = x[k];  // the FP64 form's sorted copy
    const double t = cell_t(x[k], mn[k], inv_w);  // same rounding as the key
    const double f = floor(t * 0x1p29);           // exact scaling by 2^29
    d.G[k][i] = (uint32_t)(unsigned long long)(long long)f;  // wraps mod 8 cells
  }
  d.orig[i] = src;
}

__global__ __launch_bounds__(256) void k_gather(Dev d) {
  if (abandoned(d)) return;
  const int32_t* vsrc = d.val[sort_parity(d, kMaxPasses)];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < d.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t src = vsrc[i];
    const Inst& I = *inst_of_cell(d, d.ninst == 1 ? 0 : d.cell_of[i]);
    if (I.dim == 3)
      gather_point<3>(d, I, i, src);
    else
      gather_point<2>(d, I, i, src);
  }
}

// ------------------------------------------------------------------ rows --
// line of a cell: all coordinates but the last; global line-table index
__device__ __forceinline__ int64_t cell_line(const Dev& d, int32_t c) {
  const Inst& I = *inst_of_cell(d, c);
  int32_t kc[3];
  unpack_cell(I.dim, d.cell_xyz[c], kc);
  return I.line_base + (I.dim == 3 ? (int64_t)kc[0] * I.nk[1] + kc[1] : (int64_t)kc[0]);
}

// Line table: for every possible line, the range of its non-empty cells in
// the sorted cell array (-1 when empty).
__global__ __launch_bounds__(256) void k_line_clear(Dev d) {
  if (abandoned(d)) return;
  const long long P = d.ctl[C_NLINES];
  if (P > d.line_cap) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < P;
       i += (int64_t)gridDim.x * blockDim.x)
    d.lines[i] = make_int2(-1, -1);
}

__global__ __launch_bounds__(256) void k_line_mark(Dev d) {
  if (abandoned(d)) return;
  const int32_t C = (int32_t)d.ctl[C_NCELL];
  if (d.ctl[C_NLINES] > d.line_cap) return;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < C;
       c += (int64_t)gridDim.x * blockDim.x) {
    // the three lines are read before either store (which could alias the
    // cell arrays as far as the compiler knows)
    const int64_t L = cell_line(d, (int32_t)c);
    const int64_t Lp = c > 0 ? cell_line(d, (int32_t)c - 1) : -1;
    const int64_t Ln = c < C - 1 ? cell_line(d, (int32_t)c + 1) : -1;
    if (Lp != L) d.lines[L].x = (int32_t)c;
    if (Ln != L) d.lines[L].y = (int32_t)c + 1;
  }
}

// Five lower bounds lo[j] = first index in [lo[j], hi[j]) with a[i] >= v[j],
// in lockstep (the loads of the five searches overlap).
__device__ __forceinline__ void lower_bound5(const uint64_t* a, int32_t (&lo)[5],
                                             int32_t (&hi)[5], const uint64_t (&v)[5]) {
  for (;;) {
    uint64_t km[5];
    bool any = false;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      km[j] = lo[j] < hi[j] ? a[(lo[j] + hi[j]) >> 1] : 0;
      any |= lo[j] < hi[j];
    }
    if (!any) return;
    for (int j = 0; j < 5; ++j)
      if (lo[j] < hi[j]) {
        const int32_t mid = (lo[j] + hi[j]) >> 1;
        if (km[j] < v[j])
          lo[j] = mid + 1;
        else
          hi[j] = mid;
      }
  }
}

// Per-lane view of one cell and one line group: the cell's coordinates,
// its window [zlo, zhi] along the last axis, its own line (global
// line-table index), the shift to the neighbour lines of the group, and
// which of the five neighbour lines exist.  3-D: the lines with offset
// g - 2 along axis 0 and -2..2 along axis 1 (consecutive line-table
// entries); 2-D: the five lines (g = 0).
struct RowLane {
  const Inst* I;
  int32_t iid;
  int dim;
  int32_t zlo, zhi;
  int64_t own, shift;
  bool ok[5];
};

__device__ __forceinline__ void row_lane(const Dev& d, int32_t c, bool act, int g, int32_t C,
                                         RowLane& L) {
  const int32_t cc = act ? c : C - 1;
  L.iid = d.ninst == 1 ? 0 : d.cell_inst[cc];
  L.I = d.inst + L.iid;
  const Inst& I = *L.I;
  const int dim = I.dim;
  L.dim = dim;
  const int last = dim - 1;
  int32_t kc[3];
  unpack_cell(dim, d.cell_xyz[cc], kc);
  L.zlo = kc[last] >= 2 ? kc[last] - 2 : 0;
  L.zhi = kc[last] + 2 < I.nk[last] ? kc[last] + 2 : I.nk[last] - 1;
  const int a = dim == 3 ? g - 2 : 0;
  L.own = I.line_base + (dim == 3 ? (int64_t)kc[0] * I.nk[1] + kc[1] : (int64_t)kc[0]);
  L.shift = dim == 3 ? (int64_t)a * I.nk[1] : 0;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    int o0 = a, o1 = j - 2;
    if (dim != 3) {
      o0 = kRowOff2[j];
      o1 = 0;
    }
    const int32_t r0 = kc[0] + o0, r1 = kc[1] + o1;
    bool ok = act && (dim == 3 || g == 0) && r0 >= 0 && r0 < I.nk[0];
    if (dim == 3) ok = ok && r1 >= 0 && r1 < I.nk[1];
    L.ok[j] = ok;
  }
}

__device__ __forceinline__ int row_off(int dim, int j) { return dim == 3 ? j - 2 : kRowOff2[j]; }

__device__ __forceinline__ void write_rows(const Dev& d, int32_t c, bool act, int dim, int g,
                                           const int32_t* packed, long long& cand,
                                           long long& nbrc) {
  if (d.nbr_stats) {
#pragma unroll
    for (int j = 0; j < 5; ++j)
      if (packed[j]) {
        const int32_t a = packed[j] & kRowCellMask, b = a + (packed[j] >> kRowCntShift);
        nbrc += b - a;
        cand += (long long)(d.cell_start[c + 1] - d.cell_start[c]) *
                (d.cell_start[b] - d.cell_start[a]);
      }
  }
  if (act) {
    // the five row indices are read before the first store
    int r[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      // (2-D: groups g > 0 write the unused rows 5g .. 5g+4, empty)
      r[j] = dim == 3 ? kRowOfOff3[g][j] : 5 * g + j;
    }
#pragma unroll
    for (int j = 0; j < 5; ++j) d.rows[(int64_t)r[j] * d.rstride + c] = packed[j];
  }
}

// For each (cell, stencil line): the run of non-empty cells of that line
// whose last coordinate is within +-2 of the cell's, packed as
// first | count << 28 (count <= 5).  One wavefront per (tile of 64
// consecutive cells, line group g).  The tile's cells lie in a few
// consecutive lines (cells are sorted, so the first and last lanes bound
// them), and so do their neighbour lines: the line table gives the range
// of cells those lines hold, and the wave copies the cells' sort keys,
// relative to the span's first line (below 2^29), into LDS with coalesced
// loads, padded with ~0 to a larger power of two.  A window is then two
// searches in that sorted array: the first key >= the key of (line, zlo),
// by a branchless binary search of a depth uniform over the wave, and the
// last key <= the key of (line, zhi), at most five further.  A tile whose
// cells do not fit the LDS slots (one that crosses an axis-0 boundary,
// long or sparse lines, two instances of a batch, no line table, 2-D) is
// listed for k_rows_global.
constexpr int RW_LINES = 512;  // line-table entries of a tile's span (at most)
constexpr int RW_CELLS = 512;  // cell keys staged per wavefront (a power of two)

// Rows of one (tile, line group) from LDS; false when the tile does not fit
// (then nothing is written).  3-D tiles of one instance only, so the
// instance's fields are wave-uniform (scalar loads).
__device__ __forceinline__ bool rows_staged(const Dev& d, int64_t tile, int g, int32_t C,
                                            int64_t NL, uint32_t* ck) {
  const int lane = lane_id();
  const int32_t c = (int32_t)(tile * 64 + lane);
  const bool act = c < C;
  const int32_t cc = act ? c : C - 1;
  const int nact = __popcll(ballot(act));
  int32_t iid = 0;
  if (d.ninst != 1) {
    const int32_t v = d.cell_inst[cc];
    iid = rfl_i(v);
    if (iid != rl_i(v, nact - 1)) return false;
  }
  const Inst& I = d.inst[iid];
  if (I.dim != 3) {
    // 2-D: groups g > 0 hold no lines (their rows are written empty); the
    // five lines of g = 0 are long and rarely fit the slots
    if (g == 0) return false;
    if (act) {
#pragma unroll
      for (int j = 0; j < 5; ++j) d.rows[(int64_t)(5 * g + j) * d.rstride + c] = 0;
    }
    return true;
  }
  const int32_t nk0 = I.nk[0], nk1 = I.nk[1], nk2 = I.nk[2];
  const uint64_t pc = d.cell_xyz[cc];
  const int32_t x = (int32_t)(pc & 0x1fffff), y = (int32_t)((pc >> 21) & 0x1fffff);
  const int32_t z = (int32_t)(pc >> 42);
  const int a = g - 2;
  const bool okx = act & (x + a >= 0) & (x + a < nk0);
  // (the lines of a tile's first and last lanes bound its lines: cells are
  // sorted)
  const int64_t own = (int64_t)x * nk1 + y;  // within the instance
  const int64_t own0 = rl_l(own, 0), own1 = rl_l(own, nact - 1);
  const int64_t nl = (int64_t)nk0 * nk1;
  int64_t lmin = own0 + (int64_t)a * nk1 - 2, lmax = own1 + (int64_t)a * nk1 + 2;
  lmin = lmin < 0 ? 0 : lmin;
  lmax = lmax >= nl ? nl - 1 : lmax;
  if (lmax - lmin >= RW_LINES) return false;
  const int64_t gl = I.line_base + lmin;  // first line-table entry of the span
  // the cells of the span: [first start, last end) of its non-empty lines
  // (the lines are in cell order: the first and last non-empty entries)
  const int span = lmax >= lmin ? (int)(lmax - lmin + 1) : 0;
  int32_t cmin = -1, cmax = -1;
  // the entries' loads are all issued before the first is used
  int2 ev[RW_LINES / 64];
#pragma unroll
  for (int k = 0; k < RW_LINES / 64; ++k) {
    const int i = lane + 64 * k;
    ev[k] = make_int2(-1, -1);
    if (i < span && gl + i < NL) ev[k] = d.lines[gl + i];
  }
#pragma unroll
  for (int k = 0; k < RW_LINES / 64; ++k) {
    if (64 * k < span) {
      const int2 e = ev[k];
      const unsigned long long m = ballot(e.x >= 0 && e.y > e.x);
      if (m) {
        if (cmin < 0) cmin = rl_i(e.x, __ffsll((long long)m) - 1);
        cmax = rl_i(e.y, 63 - __clzll((long long)m));
      }
    }
  }
  const int ncz = cmin >= 0 ? cmax - cmin : 0;
  if (ncz >= RW_CELLS) return false;
  cmin = cmin < 0 ? 0 : cmin;
  // a power of two > ncz: the last slot is always a ~0 sentinel, so the
  // search (which reaches index pad - 2 at most) is exact
  const int pad = 1 << (32 - __clz(ncz));
  // keys relative to kbase fit 32 bits, so the low halves of the keys are
  // enough (little-endian), all loaded before the first is stored
  const uint32_t kbase = (uint32_t)(I.key_base + lmin * nk2);
  const uint32_t* key_lo = reinterpret_cast<const uint32_t*>(d.cell_key);
  uint32_t kv[RW_CELLS / 64];
#pragma unroll
  for (int k = 0; k < RW_CELLS / 64; ++k) {
    const int i = lane + 64 * k;
    kv[k] = 0u;
    if (i < ncz) kv[k] = key_lo[2 * (int64_t)(cmin + i)];
  }
#pragma unroll
  for (int k = 0; k < RW_CELLS / 64; ++k) {
    const int i = lane + 64 * k;
    if (i < pad) ck[i] = i < ncz ? kv[k] - kbase : 0xffffffffu;
  }
  wave_sync();
  // keys of (line of j, zlo), relative to the span's first line
  const int32_t zlo = z >= 2 ? z - 2 : 0, zhi = z + 2 < nk2 ? z + 2 : nk2 - 1;
  const uint32_t k0 = (uint32_t)((own + (int64_t)a * nk1 - 2 - lmin) * nk2) + (uint32_t)zlo;
  uint32_t klo[5];
  int32_t pos[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    klo[j] = k0 + (uint32_t)(j * nk2);
    pos[j] = -1;  // last index with key < klo
  }
  for (int h = pad >> 1; h > 0; h >>= 1) {
    uint32_t v[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) v[j] = ck[pos[j] + h];
#pragma unroll
    for (int j = 0; j < 5; ++j) pos[j] = v[j] < klo[j] ? pos[j] + h : pos[j];
  }
  const uint32_t dz = (uint32_t)(zhi - zlo);
  int rrow[5];  // read before the first store
#pragma unroll
  for (int j = 0; j < 5; ++j) rrow[j] = kRowOfOff3[g][j];
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    // the run: keys up to (line, zhi), at most five after pos
    int32_t e = pos[j];
#pragma unroll
    for (int h = 4; h > 0; h >>= 1) {
      const int32_t q = e + h;
      e = (q < pad && ck[q < pad ? q : 0] <= klo[j] + dz) ? q : e;
    }
    const int32_t cnt = e - pos[j];
    const bool ok = okx & (y + j - 2 >= 0) & (y + j - 2 < nk1);
    const int32_t packed = (ok && cnt > 0) ? ((cmin + pos[j] + 1) | (cnt << kRowCntShift)) : 0;
    if (act) d.rows[(int64_t)rrow[j] * d.rstride + c] = packed;
  }
  return true;
}

__global__ __launch_bounds__(256, 8) void k_rows(Dev d) {
  if (abandoned(d)) return;
  __shared__ uint32_t s_ck[4][RW_CELLS];
  const int32_t C = (int32_t)d.ctl[C_NCELL];
  const int64_t NL = d.ctl[C_NLINES];
  const bool table = NL <= d.line_cap;
  const int wv = threadIdx.x >> 6;
  // work item: tile * 5 + line group; the four waves of a workgroup take
  // four consecutive items per step (the steps are uniform over the
  // workgroup, as the list append wants)
  const int64_t items = (((int64_t)C + 63) >> 6) * 5;
  for (int64_t base = (int64_t)blockIdx.x * 4; base < items; base += (int64_t)gridDim.x * 4) {
    const int64_t it = base + wv;
    bool left = false;
    if (it < items) {
      const int64_t tile = it / 5;
      // (the neighbourhood statistics run takes the global-memory pass)
      left = !(table && !d.nbr_stats &&
               rows_staged(d, tile, (int)(it - tile * 5), C, NL, s_ck[wv]));
    }
    block_append(lane_id() == 0 && left, (int32_t)it, d.rows_fb);
  }
}

// The (tile, line group) pairs k_rows left: one lane per cell, the line's
// cell range from the line table (or by two lower bounds over all cell
// keys when there is none), then the first cell with z >= zlo by binary
// search over cell_z in global memory, the five lines' searches in lockstep
// so that their loads overlap.
__global__ __launch_bounds__(256) void k_rows_global(Dev d) {
  if (abandoned(d)) return;
  __shared__ long long s_pre[kStripes + 1];
  const long long nit = slist_prefix(d.rows_fb, s_pre);
  const int32_t C = (int32_t)d.ctl[C_NCELL];
  const bool table = d.ctl[C_NLINES] <= d.line_cap;
  const int wv = threadIdx.x >> 6, lane = lane_id();
  long long cand = 0, nbrc = 0;
  for (long long w = (long long)blockIdx.x * 4 + wv; w < nit; w += (long long)gridDim.x * 4) {
    const int64_t it = slist_at(d.rows_fb, s_pre, w);
    const int64_t tile = it / 5;
    const int g = (int)(it - tile * 5);
    const int32_t c = (int32_t)(tile * 64 + lane);
    const bool act = c < C;
    RowLane L;
    row_lane(d, c, act, g, C, L);
    const Inst& I = *L.I;
    const int last = L.dim - 1;
    int32_t ls[5], le[5];
    uint64_t k0[5];  // first key of each neighbour line (no line table)
    bool need = false;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const int64_t line = L.own + L.shift + row_off(L.dim, j) - I.line_base;
      ls[j] = -1;
      le[j] = -1;
      k0[j] = (uint64_t)(I.key_base + line * I.nk[last]);
      if (L.ok[j] && table) {
        const int2 lr = d.lines[I.line_base + line];
        ls[j] = lr.x;
        le[j] = lr.y;
      } else if (L.ok[j]) {
        ls[j] = 0;  // searched below
        le[j] = C;
        need = true;
      }
    }
    if (need) {
      int32_t hi[5];
#pragma unroll
      for (int j = 0; j < 5; ++j) hi[j] = le[j];
      lower_bound5(d.cell_key, ls, hi, k0);
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        hi[j] = le[j];
        le[j] = ls[j];
        k0[j] += (uint64_t)I.nk[last];
      }
      lower_bound5(d.cell_key, le, hi, k0);
#pragma unroll
      for (int j = 0; j < 5; ++j)
        if (le[j] <= ls[j]) ls[j] = le[j] = -1;  // empty line
    }
    // lines of more than 4 cells: first cell with z >= zlo
    int32_t lo[5], hi[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const bool srch = ls[j] >= 0 && le[j] - ls[j] > 4;
      lo[j] = srch ? ls[j] : 0;
      hi[j] = srch ? le[j] : 0;
    }
    for (;;) {
      int32_t zm[5];
      bool any = false;
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        zm[j] = lo[j] < hi[j] ? d.cell_z[(lo[j] + hi[j]) >> 1] : 0;
        any |= lo[j] < hi[j];
      }
      if (!any) break;
#pragma unroll
      for (int j = 0; j < 5; ++j)
        if (lo[j] < hi[j]) {
          const int32_t mid = (lo[j] + hi[j]) >> 1;
          if (zm[j] < L.zlo)
            lo[j] = mid + 1;
          else
            hi[j] = mid;
        }
    }
    // from the first candidate (line start of a short line, search result
    // of a long one) the next up to 5 cells decide the run
    int32_t packed[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const bool live = ls[j] >= 0 && le[j] > ls[j];
      const bool shortl = le[j] - ls[j] <= 4;
      const int32_t a0 = shortl ? ls[j] : lo[j];
      int32_t a = a0, b = a0;
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        const int32_t z = (live && a0 + i < le[j] && (i < 4 || !shortl)) ? d.cell_z[a0 + i]
                                                                         : 0x7fffffff;
        a += z < L.zlo;
        b += z <= L.zhi;
      }
      packed[j] = (ls[j] >= 0 && b > a) ? (a | ((b - a) << kRowCntShift)) : 0;
    }
    write_rows(d, c, act, L.dim, g, packed, cand, nbrc);
  }
  stat_add(d, C_CAND, cand);
  stat_add(d, C_NBRC, nbrc);
}


// ------------------------------------------------------------- launchers --
void launch_index_head(const Dev& d, int nchunks, hipStream_t s) {
  hipLaunchKernelGGL(k_bbox, dim3(nchunks), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(1024), 0, s, d);
}

// Keys per thread of the radix passes over n keys.  A scatter workgroup of
// 2048 32-bit keys (23 KiB of LDS) shares a CU with five others, one of
// 4096 keys (39 KiB) with three; a 4096-key tile takes about 1.6x as long.
// The larger tiles are taken when their rounds of resident workgroups cost
// less: config 3's 4 M keys run as 977 tiles in one round instead of 1954
// tiles in two; 1 M keys (one partial round either way) and 32 M keys stay
// with the smaller tiles.
int sort_items(bool k32, int64_t n, int cus) {
  if (!k32 || cus <= 0) return RS_I;
  const int64_t r2 = (n + 2048 * 6LL * cus - 1) / (2048 * 6LL * cus);
  const int64_t r4 = (n + 4096 * 4LL * cus - 1) / (4096 * 4LL * cus);
  return 16 * r4 < 10 * r2 ? 16 : RS_I;
}

template <bool K32, int NI>
static void index_sort(const Dev& d, int64_t n, int nchunks, int passes, hipStream_t s) {
  hipLaunchKernelGGL(k_cellkey<K32>, dim3(nchunks), dim3(256), 0, s, d);
  const int nblk = (int)((n + RS_T * NI - 1) / (RS_T * NI));
  for (int p = 0; p < passes; ++p) {
    hipLaunchKernelGGL((k_rs_hist<K32, NI>), dim3(nblk), dim3(RS_T), 0, s, d, p, nblk);
    hipLaunchKernelGGL(k_rs_rowscan, dim3(256), dim3(256), 0, s, d, p, nblk);
    hipLaunchKernelGGL((k_rs_scatter<K32, NI>), dim3(nblk), dim3(RS_T), 0, s, d, p, nblk);
  }
  const int ncs = (int)((n + CS_TILE - 1) / CS_TILE);
  hipLaunchKernelGGL(k_csr_count<K32>, dim3(ncs), dim3(CS_T), 0, s, d, passes);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, d.scan_tmp, ncs, (long long*)nullptr);
  hipLaunchKernelGGL(k_csr_write<K32>, dim3(ncs), dim3(CS_T), 0, s, d);
}

// `passes` radix passes: keys of at most 32 bits (four passes) are sorted as
// 32-bit keys
void launch_index_sort(const Dev& d, int64_t n, int nchunks, int passes, int cus, hipStream_t s) {
  if (passes > 4)
    index_sort<false, RS_I>(d, n, nchunks, passes, s);
  else if (sort_items(true, n, cus) == 16)
    index_sort<true, 16>(d, n, nchunks, passes, s);
  else
    index_sort<true, RS_I>(d, n, nchunks, passes, s);
}

void launch_gather(const Dev& d, int64_t n, hipStream_t s) {
  // random-access gather: more blocks in flight than a fill needs
  const int64_t gb = (n + 255) / 256;
  hipLaunchKernelGGL(k_gather, dim3((int)std::max<int64_t>(1, std::min<int64_t>(16384, gb))),
                     dim3(256), 0, s, d);
}

void launch_lines(const Dev& d, int grid_cells, hipStream_t s) {
  hipLaunchKernelGGL(k_line_clear, dim3(grid_cells), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_line_mark, dim3(grid_cells), dim3(256), 0, s, d);
}

// Row pass grid: one wavefront per (tile of 64 cells, line group) for up
// to n cells, at most kRowGrid workgroups (the steps loop beyond).
constexpr int64_t kRowGrid = 16384;
int64_t row_blocks(int64_t n) {
  return std::max<int64_t>(1, std::min<int64_t>(kRowGrid, (((n + 63) / 64) * 5 + 3) / 4));
}
int64_t row_list_cap(int64_t n) {
  const int64_t items = ((n + 63) / 64) * 5;
  const int64_t g = row_blocks(n);
  const int64_t steps = (items + 4 * g - 1) / (4 * g);
  return (g + kStripes - 1) / kStripes * 4 * steps;  // appends per stripe, at most
}

void launch_rows(const Dev& d, int64_t n, int grid_cells, hipStream_t s) {
  hipLaunchKernelGGL(k_rows, dim3((unsigned)row_blocks(n)), dim3(256), 0, s, d);
  // the tiles left over, one wavefront each: enough waves for the list
  const int64_t gb = std::max<int64_t>(1, std::min<int64_t>(4LL * grid_cells, (n + 255) / 256));
  hipLaunchKernelGGL(k_rows_global, dim3((unsigned)gb), dim3(256), 0, s, d);
}

void launch_scan_top(int64_t* v, int nb, hipStream_t s) {
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, v, nb, (long long*)nullptr);
}

int64_t sort_hist_entries(int64_t n) { return 256 * ((n + RS_TILE - 1) / RS_TILE); }
int64_t scan_tmp_entries(int64_t n) { return (n + CS_TILE - 1) / CS_TILE + 16; }

}  // namespace dclust
