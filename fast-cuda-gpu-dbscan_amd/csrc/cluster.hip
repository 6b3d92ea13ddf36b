// Clustering kernels on gfx950: the MI355X replacement for the reference's
// expansion kernel and host merge loop.
//
// Grid: cells of side (eps/2)(1+2^-20), so every cell is a clique (all its
// points are mutual neighbours) and every neighbour of a point lies in the
// 5^D stencil of cells around its cell.  The stencil is stored per cell as
// 5^(D-1) packed runs ("lines") of at most 5 consecutive non-empty cells
// along the last axis; line r, cell k of a cell's stencil is its stencil
// slot 5r+k.
//
//   k_core       one lane per point: neighbour count with early exit at
//                minpts -> core flag.  (dbscan.cu:331-369 counts while it
//                expands; here the count is its own pass, so every later
//                kernel knows which points are core.)  A cell holding
//                >= minpts points is all core without a distance test.  A
//                point that stays below minpts has seen all its neighbours
//                and records them for the border pass.
//   k_adj        cell adjacency: bit 5r+k of a core cell's 128-bit mask is
//                set when stencil slot 5r+k holds a core point within eps of
//                one of the cell's core points.  Light cells: one lane per
//                core point.  Heavy cells (> kHeavy points): one wavefront
//                per cell (k_adj_heavy), which also sets the mirrored bit of
//                its light neighbours.
//   k_expand     persistent seed-list expansion (dbscan.cu:254-387) over
//                cells: each wavefront owns a seed list of cells in LDS and
//                pops up to 64 of them at once, one per lane; each lane
//                claims the adjacent unclaimed cells of its cell with one CAS
//                each (MarkAsCandidate, dbscan.cu:189-222) and pushes them;
//                cells already claimed by another chain become collision
//                pairs in the wave's own buffer (extraCollision,
//                dbscan.cu:202-211).  Seed-list overflow and lists left when
//                a wave's budget runs out go to a correction queue that the
//                next pass expands, instead of being dropped
//                (dbscan.cu:196-199).
//   k_merge_lds / k_union_pairs / k_flatten   GPU atomic union-find over
//                chain ids (MonitorSeedPoints' CPU BFS over the collision
//                matrix and thrust::replace relabels, dbscan.cu:60-125).
//   k_cell_comp / k_flatten_all / k_keys / k_border / rank scan / k_output
//                canonical labels:
//                cluster id = rank of the component's minimum core input
//                index; border = minimum id over adjacent cores; noise = 0
//                (GetDbscanResult, dbscan.cu:225-251, writes sorted labels
//                instead; this writes per-point labels).
//
// Neighbour predicate P(p,q), FP64 form (exactly the reference oracle's,
// cpu-dbscan/dbscan_rtree_exp.cpp:351-358,420-438 and Rtree.h:1573-1587):
// closed box fl(p-eps) <= q <= fl(p+eps) on every axis, and
// fl(fl(fl(dx0^2) + fl(dx1^2)) [+ fl(dx2^2)]) <= fl(eps*eps).  This file is
// compiled with -ffp-contract=off so no FMA is formed.  d^2 is symmetric in
// p and q (fl(p-q) = -fl(q-p)); only the box test can differ between the two
// directions.  Two core points are joined when d^2 <= eps^2 and the box test
// holds from at least one side (the oracle's union runs from every core
// point, cpu-dbscan/dbscan_rtree_exp.cpp:379-404).
#include <climits>

#include "internal.h"
#include "kernels.h"
#include "stencil.h"

namespace dclust {

constexpr int kQCache = 256;  // query points a heavy-cell wavefront keeps in LDS

constexpr int kCoreBatch = 4;  // neighbour-count candidates loaded per lane step

struct Tally {  // per-lane event counts (summed per wave by stat_add)
  int band = 0, flip = 0, asym = 0, eval = 0;
  int groups = 0, lines = 0, cells = 0;  // stencil work (algorithmic bytes)
};

// FP32 squared distance from fixed-point coordinates (units of w*2^-29).
// Differences are exact in wrapped int32 arithmetic because two stencil
// points are less than three cells apart on every axis.
template <int D>
__device__ __forceinline__ float d2fix(const uint32_t* a, const uint32_t* b) {
  float f = (float)(int32_t)(b[0] - a[0]);
  float s = f * f;
#pragma unroll
  for (int k = 1; k < D; ++k) {
    float g = (float)(int32_t)(b[k] - a[k]);
    s = __builtin_fmaf(g, g, s);
  }
  return s;
}

// FP32 squared distance (fixed point) from point g to the box [lo, hi] of
// points of one cell (no wrap inside a cell).  Never above d2fix(g, q) for
// any q in the box, so d2box > d2_out rules out every pair.
template <int D>
__device__ __forceinline__ float d2box(const uint32_t* g, const uint32_t* lo, const uint32_t* hi) {
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    int32_t a = (int32_t)(lo[k] - g[k]);
    int32_t b = (int32_t)(g[k] - hi[k]);
    int32_t gap = a > 0 ? a : (b > 0 ? b : 0);
    float f = (float)gap;
    s = __builtin_fmaf(f, f, s);
  }
  return s;
}

// Scalar parameters of one instance, read once per lane.
struct Par {
  double eps, eps2;
  float d2_in, d2_out, d2_mid;
  int minpts;
  __device__ __forceinline__ explicit Par(const Inst& I)
      : eps(I.eps), eps2(I.eps2), d2_in(I.d2_in), d2_out(I.d2_out), d2_mid(I.d2_mid),
        minpts(I.minpts) {}
};

// Core-core relation of the pair (a, b): d^2 <= eps^2 and the box test from
// at least one side.  FP32 form: decided by the fixed-point distance outside
// the guard band, exactly in FP64 inside it.
template <int D, bool F32>
__device__ __forceinline__ bool edge_pair(const Dev& d, const Par& P, int32_t a, int32_t b,
                                          const uint32_t* ag, const uint32_t* bg,
                                          const double* ax, const double* bx, Tally& t) {
  double xa[D], xb[D];
  if (F32) {
    const float s = d2fix<D>(ag, bg);
    if (s < P.d2_in) return true;
    if (s > P.d2_out) return false;
    t.band++;
    load_x<D, F32>(d, a, xa);
    load_x<D, F32>(d, b, xb);
  } else {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      xa[k] = ax[k];
      xb[k] = bx[k];
    }
  }
  const double dd = d2_64<D>(xa, xb);
  if (!(dd <= P.eps2)) return false;
  const bool f = box_64<D>(xa, xb, P.eps), r = box_64<D>(xb, xa, P.eps);
  t.asym += (f != r) && a < b;
  return f || r;
}

__device__ __forceinline__ bool mask_has(unsigned long long m0, unsigned long long m1, int b) {
  return ((b < 64 ? (m0 >> b) : (m1 >> (b - 64))) & 1ull) != 0;
}
__device__ __forceinline__ void mask_set(unsigned long long& m0, unsigned long long& m1, int b) {
  if (b < 64)
    m0 |= 1ull << b;
  else
    m1 |= 1ull << (b - 64);
}

// Per-lane dispatch on the instance dimension (a wavefront can straddle two
// instances of a batch).
#define DCLUST_BY_DIM(act, dim, CALL3, CALL2)      \
  do {                                             \
    if (ballot((act) && (dim) == 3)) {             \
      if ((act) && (dim) == 3) { CALL3; }          \
    }                                              \
    if (ballot((act) && (dim) == 2)) {             \
      if ((act) && (dim) == 2) { CALL2; }          \
    }                                              \
  } while (0)


// ------------------------------------------------------------------ core --
// Neighbour count of point p with early exit at minpts.  Every candidate
// pair with d^2 <= eps^2 (either box direction) is recorded while the count
// is below minpts; a point that ends non-core thereby holds all its
// neighbours outside its own cell (at most minpts - 1 plus box-asymmetric
// pairs).
template <int D, bool F32>
__device__ void core_lane(const Dev& d, const Par& P, int32_t p, int32_t c, int32_t qs,
                          int32_t qe, int R, int32_t* rk, Tally& t, bool& isc) {
  constexpr int NR = D == 3 ? 25 : 5;
  constexpr int B = kCoreBatch;
  int cnt = qe - qs;  // the own cell is a clique, p itself included
  if (cnt >= P.minpts) {
    isc = true;
    return;
  }
  uint32_t pg[D];
  double px[D], lo[D], hi[D];
  if (F32) {
    load_g<D>(d, p, pg);
  } else {
    load_x<D, F32>(d, p, px);
#pragma unroll
    for (int k = 0; k < D; ++k) {
      lo[k] = px[k] - P.eps;
      hi[k] = px[k] + P.eps;
    }
  }
  // records stay in this lane's LDS slots until the point is known to be
  // non-core
  int nrec = 0;
  int ev = 0;
  const RowsOf rows = rows_of(d, c);
  // the next group's packed rows are loaded while this group streams
  int32_t pkn[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) pkn[j] = rows[j];
  for (int g = 0; g < NR && cnt < P.minpts; g += 5) {
    int32_t aa[5], ll[5];
    line_ranges(d, pkn, aa, ll);
    if (g + 5 < NR) {
#pragma unroll
      for (int j = 0; j < 5; ++j) pkn[j] = rows[g + 5 + j];
    }
    ++t.groups;
#pragma unroll
    for (int j = 0; j < 5; ++j) t.lines += ll[j] > 0;
    Stream5 s;
    s.set(aa, ll);
    for (int32_t t0 = 0; t0 < s.tot && cnt < P.minpts; t0 += B) {
      int32_t qi[B];
      uint32_t cg[B][D];
      double cx[B][D];
#pragma unroll
      for (int j = 0; j < B; ++j) {
        int jr;
        qi[j] = t0 + j < s.tot ? s.at(t0 + j, jr) : p;
        if (F32)
          load_g<D>(d, qi[j], cg[j]);
        else
          load_x<D, F32>(d, qi[j], cx[j]);
      }
#pragma unroll
      for (int j = 0; j < B; ++j) {
        const int32_t q = qi[j];
        if (t0 + j < s.tot && (q < qs || q >= qe)) {
          bool fwd, sym;
          if (F32) {
            const float sd = d2fix<D>(pg, cg[j]);
            if (sd < P.d2_in) {
              fwd = sym = true;
            } else if (sd > P.d2_out) {
              fwd = sym = false;
            } else {
              double ax[D], bx[D];
              load_x<D, F32>(d, p, ax);
              load_x<D, F32>(d, q, bx);
              const double dd = d2_64<D>(ax, bx);
              sym = dd <= P.eps2;
              fwd = sym && box_64<D>(ax, bx, P.eps);
              t.band++;
              t.flip += fwd != (sd <= P.d2_mid);
            }
          } else {
            const double dd = d2_64<D>(px, cx[j]);
            sym = dd <= P.eps2;
            bool in = true;
#pragma unroll
            for (int k = 0; k < D; ++k) in = in && lo[k] <= cx[j][k] && cx[j][k] <= hi[k];
            fwd = sym && in;
          }
          cnt += fwd;
          if (sym) {
            if (nrec < kRecSlots) rk[nrec] = q;
            ++nrec;
          }
          ++ev;
        }
      }
    }
  }
  isc = cnt >= P.minpts;
  if (!isc) {
    int32_t* rec = d.rec + (int64_t)p * R;
    for (int i = 0; i < nrec && i < R; ++i) rec[i] = rk[i];
    d.rec_n[p] = (uint8_t)(nrec < 255 ? nrec : 255);
  }
  t.eval += ev;
}

// 6 waves per SIMD for the FP32 form (80 VGPRs with a batch of four
// candidates; latency bound, so residency pays more than a longer batch), 5
// for the FP64 form, whose coordinates need more registers
template <bool F32>
__global__ __launch_bounds__(256, F32 ? 6 : 5) void k_core(Dev d) {
  if (abandoned(d)) return;
  const int64_t pl = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const bool act = pl < d.n;
  const int32_t p = act ? (int32_t)pl : 0;
  const int lane = lane_id();
  int32_t c = 0, qs = 0, qe = 0;
  int dim = 0;
  const Inst* I = d.inst;
  if (act) {
    c = d.cell_of[p];
    qs = d.cell_start[c];
    qe = d.cell_start[c + 1];
    I = inst_of_cell(d, c);
    dim = I->dim;
  }
  Tally t;
  bool isc = false;
  const int R = d.rec_stride;
  const Par P(*I);
  // kRecSlots record slots per lane (stride kRecSlots + 1: no bank conflicts)
  __shared__ int32_t s_rec[256 * (kRecSlots + 1)];
  int32_t* rk = s_rec + threadIdx.x * (kRecSlots + 1);
  DCLUST_BY_DIM(act, dim, (core_lane<3, F32>(d, P, p, c, qs, qe, R, rk, t, isc)),
                (core_lane<2, F32>(d, P, p, c, qs, qe, R, rk, t, isc)));
  if (act) d.core[p] = isc;
  // core points of light cells: the lanes of the near-adjacency pass
  block_append(isc && qe - qs <= kHeavy, p, d.adj_pts);
  // non-core points: the lanes of the border-label pass
  block_append(act && !isc, p, d.border_pts);
  // per-cell core counts: the lanes of one cell form a contiguous run
  const bool head = act && (p == qs || lane == 0);
  const unsigned long long mh = ballot(head), mc = ballot(isc);
  long long dense = 0;
  if (head) {
    const unsigned long long above = lane == 63 ? 0ull : (mh & (~0ull << (lane + 1)));
    const int nxt = above ? __ffsll((long long)above) - 1 : 64;
    const unsigned long long seg = (nxt == 64 ? ~0ull : ((1ull << nxt) - 1ull)) & (~0ull << lane);
    const int k = __popcll(mc & seg);
    if (k) atomicAdd(&d.cell_ncore[c], k);
    if (p == qs) {
      // first point of its cell: clear the adjacency mask
      d.amask[c] = 0ull;
      dense = (qe - qs) >= P.minpts;
    }
  }
  // heavy cells, listed once (by their first point)
  block_append(act && p == qs && qe - qs > kHeavy, c, d.heavy);
  stat_add(d, C_NCORE, isc ? 1 : 0);
  stat_add(d, C_DENSE, dense);
  stat_add(d, C_EVAL, t.eval);
  stat_add(d, C_BAND_CORE, t.band);
  stat_add(d, C_BAND, t.band);
  stat_add(d, C_FLIP, t.flip);
  stat_add(d, C_CORE_GROUPS, t.groups);
  stat_add(d, C_CORE_LINES, t.lines);
}

// ------------------------------------------------------------- adjacency --
// Distance (fixed-point units) from a point at fraction f of its cell to
// the cell o cells away along one axis (0 for o = 0).
__device__ __forceinline__ float cell_gap(int o, int32_t f) {
  const int32_t g = o > 0 ? (o << kFixBits) - f : (o < 0 ? f + ((-o - 1) << kFixBits) : 0);
  return (float)g;
}

// Distance (fixed-point units) from a point at fraction f of its cell to
// the far side of the cell o cells away along one axis.
__device__ __forceinline__ float cell_far(int o, int32_t f) {
  constexpr int32_t W = 1 << kFixBits;
  const int32_t g = o > 0 ? (o + 1) * W - f : (o < 0 ? f - o * W : (f > W - f ? f : W - f));
  return (float)g;
}

// Near and far stencil cells.  A stencil cell is "near" when it is at most
// one cell away on every axis: the cells of lines 0..NN-1 (the lines whose
// offsets are at most 1) with a last-axis offset of at most 1.  The other
// stencil cells are "far".  Adjacency over near cells is computed for every
// core cell and drives the seed-list expansion; a far pair is decided only
// when its two cells are still in different components after the
// expansion and merge (k_far), which is rare: in the interior of a cluster
// the near cells already connect everything.
template <int D>
struct Near {
  static constexpr int lines = D == 3 ? 9 : 3;
};

// Is some core point of A (points [qs, qe)) joined to some core point of B
// (points [b0, b1), holding at least one core point)?  off = offset of B's
// cell from A's.  A's core points are tried in turn: skipped when B's cell
// box lies beyond eps of the point, accepted without a test when B's whole
// box lies inside the guard band around it (every point of B is then
// joined to it), else tested against B's core points.  The box distances
// use the same fixed-point arithmetic as the distance, so they bound it.
template <int D, bool F32>
__device__ bool cells_joined(const Dev& d, const Par& P, int32_t qs, int32_t qe, bool allc,
                             int32_t b0, int32_t b1, int off0, int off1, int offl, Tally& t,
                             long long& freeb) {
  constexpr uint32_t kFrac = (1u << kFixBits) - 1u;
  const int off[3] = {off0, D == 3 ? off1 : offl, offl};
  for (int32_t a = qs; a < qe; ++a) {
    // the flag and the coordinates are loaded together
    uint32_t ag[D];
    load_g<D>(d, a, ag);
    if (!allc && !d.core[a]) continue;
    float g2 = 0.f, h2 = 0.f;
#pragma unroll
    for (int x = 0; x < D; ++x) {
      const int32_t f = (int32_t)(ag[x] & kFrac);
      const float e = cell_gap(off[x], f), fe = cell_far(off[x], f);
      g2 = __builtin_fmaf(e, e, g2);
      h2 = __builtin_fmaf(fe, fe, h2);
    }
    if (g2 > P.d2_out) continue;
    if (h2 < P.d2_in) {
      ++freeb;
      return true;
    }
    double ax[D];
#pragma unroll
    for (int x = 0; x < D; ++x) ax[x] = F32 ? 0.0 : d.X[x][a];
    for (int32_t q0 = b0; q0 < b1; q0 += 4) {
      bool cf[4];
      uint32_t qg[4][D];
      double qx[4][D];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool in = q0 + j < b1;
        const int32_t q = in ? q0 + j : b0;
        // (coordinates loaded with the flag, not after it)
        cf[j] = in && d.core[q];
#pragma unroll
        for (int x = 0; x < D; ++x) {
          qg[j][x] = F32 ? d.G[x][q] : 0u;
          qx[j][x] = !F32 ? d.X[x][q] : 0.0;
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (cf[j]) {
          ++t.eval;
          if (edge_pair<D, F32>(d, P, a, q0 + j, ag, qg[j], ax, qx[j], t)) return true;
        }
      }
    }
  }
  return false;
}

// Offset along the last axis of cell B (in stencil line r of cell A).
// (B lies in a neighbour line of A, so the difference of their last
// coordinates is the offset.)
template <int D>
__device__ __forceinline__ int cell_offset_last(const Dev& d, const Inst& I, int32_t A, int32_t B,
                                                int r) {
  (void)I;
  (void)r;
  return d.cell_z[B] - d.cell_z[A];
}

// Near lines above the cell (every cell of them has a larger index): line
// (0,0) (cells one step up the line only), (1,0), (0,1), (1,-1), (1,1) in
// 3-D; (0) and (1) in 2-D.  Nibble i of the constant is the i-th such row.
template <int D>
struct Upper {
  static constexpr int lines = D == 3 ? 5 : 2;
  static constexpr unsigned long long rows = D == 3 ? 0x87420ull : 0x20ull;
  // nibble r: the mirror of near row r (the line with negated offsets)
  static constexpr unsigned long long mirror = D == 3 ? 0x567834120ull : 0x120ull;
};

// Near adjacency of light cells: one lane per core point p of a light cell
// (compacted list written by k_core).  Adjacency is symmetric (edge_pair
// is), so a pair of light cells is decided once, by the lower cell: its
// lanes examine only the near lines above it (Upper) and the bits they find
// are mirrored into the upper cell's mask at the end.  The lanes of one
// cell examine those lines in a rotated order (lane rank r0 within the
// cell's run of lanes starts at line r0) and share what they found through
// LDS after each step, so a near cell found adjacent by one lane is not
// tested again by the others.  A near cell is skipped when it is heavy (the
// heavy-cell pass sets the bits of both cells), is already known adjacent,
// or its box lies beyond eps of p; it is adjacent without a test when its
// whole box lies inside the guard band around p and it holds a core point;
// otherwise its core points are tested until one is joined to p.  The box
// distances use the same fixed-point arithmetic as the distance, so they
// bound it.
template <int D, bool F32>
__device__ void adj_lane(const Dev& d, const Inst& I, const Par& P, int32_t p, int32_t c,
                         bool act, int head, int rank, int nrun, unsigned long long* xm,
                         unsigned long long& m0, Tally& t, long long& freeb) {
  constexpr int NN = Upper<D>::lines;
  constexpr int B = 4;  // near cells need few tests: a short batch keeps registers low
  constexpr uint32_t kFrac = (1u << kFixBits) - 1u;
  uint32_t pg[D];
  double px[D];
  int32_t fr[D];
#pragma unroll
  for (int k = 0; k < D; ++k) pg[k] = act ? d.G[k][p] : 0u;
#pragma unroll
  for (int k = 0; k < D; ++k) px[k] = (!F32 && act) ? d.X[k][p] : 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) fr[k] = (int32_t)(pg[k] & kFrac);
  const int32_t zA = act ? d.cell_z[c] : 0;
  const RowsOf rows = rows_of(d, c);
  // the next line's packed row is loaded while this line is examined
  auto row_at = [](int i) { return (int)((Upper<D>::rows >> (4 * i)) & 15ull); };
  int32_t pk_next = act ? rows[row_at(rank % NN)] : 0;
  for (int step = 0; step < NN; ++step) {
    const int r = row_at((step + rank) % NN);
    const int o0 = D == 3 ? kRowOff3[r][0] : kRowOff2[r];
    const int o1 = D == 3 ? kRowOff3[r][1] : 0;
    const int32_t pk = pk_next;
    if (step + 1 < NN) pk_next = act ? rows[row_at((step + 1 + rank) % NN)] : 0;
    // own line: only the cell one step up; other lines: offsets -1..1
    const int ol_lo = r == 0 ? 1 : -1;
    const int32_t len = pk >> kRowCntShift;
    if (len > 0) {
      ++t.lines;
      t.cells += len;
      const int32_t st = pk & kRowCellMask;
      int32_t bnd[6], nco[5], zb[5];
      // unconditional loads at clamped indices (st + len <= C), so that all
      // sixteen are in flight together; entries past the run are unused
#pragma unroll
      for (int k = 0; k < 6; ++k) bnd[k] = d.cell_start[k <= len ? st + k : st];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        zb[k] = d.cell_z[k < len ? st + k : st];
        nco[k] = d.cell_ncore[k < len ? st + k : st];
      }
      float g2 = 0.f, h2 = 0.f;  // over the line's axes
      {
        const float a = cell_gap(o0, fr[0]), fa = cell_far(o0, fr[0]);
        g2 = a * a;
        h2 = fa * fa;
        if (D == 3) {
          const float b = cell_gap(o1, fr[1]), fb = cell_far(o1, fr[1]);
          g2 = __builtin_fmaf(b, b, g2);
          h2 = __builtin_fmaf(fb, fb, h2);
        }
      }
      int32_t aa[5], ll[5];
      const int base = 5 * r;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int32_t n = k < len ? bnd[k + 1] - bnd[k] : 0;
        const int ol = zb[k] - zA;  // offset along the last axis
        const float e = cell_gap(ol, fr[D - 1]), fe = cell_far(ol, fr[D - 1]);
        const bool reach = !(__builtin_fmaf(e, e, g2) > P.d2_out);
        const bool inside = __builtin_fmaf(fe, fe, h2) < P.d2_in;
        bool use = k < len && ol >= ol_lo && ol <= 1 && nco[k] > 0 && n <= kHeavy && reach &&
                   !((m0 >> (base + k)) & 1ull);
        if (use && inside) {
          // every point of the cell is joined to p, and one of them is core
          m0 |= 1ull << (base + k);
          ++freeb;
          use = false;
        }
        aa[k] = bnd[k];
        ll[k] = use ? n : 0;
      }
      Stream5 s;
      s.set(aa, ll);
      for (int32_t t0 = 0; t0 < s.tot; t0 += B) {
        int32_t qi[B];
        int kk[B];
        bool cf[B], need[B];
        uint8_t cb[B];
        uint32_t cg[B][D];
        double cx[B][D];
        // flags and coordinates of the batch loaded together, unconditionally
        // (qi is a point either way)
#pragma unroll
        for (int j = 0; j < B; ++j) {
          int jr = 0;
          const bool in = t0 + j < s.tot;
          qi[j] = in ? s.at(t0 + j, jr) : p;
          kk[j] = jr;
          need[j] = in && !((m0 >> (base + jr)) & 1ull);
          cb[j] = d.core[qi[j]];
#pragma unroll
          for (int k = 0; k < D; ++k) {
            if (F32)
              cg[j][k] = d.G[k][qi[j]];
            else
              cx[j][k] = d.X[k][qi[j]];
          }
        }
#pragma unroll
        for (int j = 0; j < B; ++j) cf[j] = need[j] && cb[j];
#pragma unroll
        for (int j = 0; j < B; ++j) {
          if (cf[j] && !((m0 >> (base + kk[j])) & 1ull)) {
            ++t.eval;
            if (edge_pair<D, F32>(d, P, p, qi[j], pg, cg[j], px, cx[j], t))
              m0 |= 1ull << (base + kk[j]);
          }
        }
      }
    }
    // share this step's finds with the other lanes of the cell
    if (act && m0) atomicOr(&xm[head], m0);
    wave_sync();
    if (act) m0 |= xm[head];
  }
  // m0 is now the cell's whole find (this wave's run of its lanes).  Each
  // bit (line r, cell k of its run) names an upper cell B; A's bit in B's
  // mask is in B's mirrored line rm, at A's place in that line's run.  The
  // bits are dealt out over the run's lanes, four at a time per lane so
  // that their loads overlap.
  if (!act) return;
  unsigned long long u = m0;
  int j = 0;
  while (u) {
    int32_t tb[4], rm[4], pos[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      while (u && j % nrun != rank) {
        u &= u - 1;
        ++j;
      }
      const int b = u ? __ffsll((long long)u) - 1 : -1;
      if (u) {
        u &= u - 1;
        ++j;
      }
      const int r = b >= 0 ? (b * 205) >> 10 : 0;  // b / 5 for 0 <= b < 128
      rm[q] = (int)((Upper<D>::mirror >> (4 * r)) & 15ull);
      tb[q] = b >= 0 ? (rows[r] & kRowCellMask) + (b - 5 * r) : -1;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) pos[q] = tb[q] >= 0 ? rows_of(d, tb[q])[rm[q]] & kRowCellMask : 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (tb[q] >= 0) atomicOr(&d.amask[tb[q]], 1ull << (5 * rm[q] + (c - pos[q])));
  }
}

template <bool F32>
// 5 waves per SIMD (96 VGPRs, a few bytes of spill): the lanes are latency
// bound, so residency pays more than the registers
__global__ __launch_bounds__(256, 5) void k_adj(Dev d) {
  if (abandoned(d)) return;
  __shared__ unsigned long long s_m[4][64];
  __shared__ long long s_pre[kStripes + 1];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const bool act = i < slist_prefix(d.adj_pts, s_pre);
  const int32_t p = act ? slist_at(d.adj_pts, s_pre, i) : 0;
  const int32_t c = act ? d.cell_of[p] : -1;
  // the lanes of one cell form a contiguous run; head = its first lane here
  const int32_t cprev = __shfl_up(c, 1, 64);
  const unsigned long long mh = ballot(act && (lane == 0 || cprev != c));
  const unsigned long long below = lane == 63 ? mh : (mh & ((2ull << lane) - 1ull));
  const int head = below ? 63 - __clzll((long long)below) : 0;
  const int rank = lane - head;
  // run length: up to the next head or the end of the active lanes (a prefix)
  const unsigned long long above = head == 63 ? 0ull : (mh & (~0ull << (head + 1)));
  const int nact = __popcll(ballot(act));
  const int nrun = (above ? __ffsll((long long)above) - 1 : nact) - head;
  s_m[wv][lane] = 0ull;
  wave_sync();
  int dim = 0;
  const Inst* I = d.inst;
  if (act) {
    I = inst_of_cell(d, c);
    dim = I->dim;
  }
  Tally t;
  long long freeb = 0;
  unsigned long long m0 = 0;
  const Par P(*I);
  // every lane runs the line loop (inactive lanes only take part in the
  // exchange), so the exchange points are uniform over the wave
  const bool any3 = ballot(act && dim == 3) != 0, any2 = ballot(act && dim == 2) != 0;
  if (any3)
    adj_lane<3, F32>(d, *I, P, p, c, act && dim == 3, head, rank, nrun, s_m[wv], m0, t, freeb);
  if (any2)
    adj_lane<2, F32>(d, *I, P, p, c, act && dim == 2, head, rank, nrun, s_m[wv], m0, t, freeb);
  // one atomic per cell run: the exchange slot of its head lane holds the
  // union of the run's finds
  if (act && lane == head) {
    const unsigned long long u = s_m[wv][lane];
    if (u) atomicOr(&d.amask[c], u);
  }
  stat_add(d, C_BAND, t.band);
  stat_add(d, C_ASYM, t.asym);
  stat_add(d, C_ADJ_TESTS, t.eval);
  stat_add(d, C_ADJ_FREE, freeb);
  stat_add(d, C_ADJ_LINES, t.lines);
  stat_add(d, C_ADJ_CELLS, t.cells);
}

// Heavy cell A (more than kHeavy points), one wavefront, near cells only:
// its core points are the queries (the first kQCache cached in LDS).  Each
// near cell with core points is decided in turn: its core points within
// eps of the queries' bounding box are the candidates (one per lane), and
// the queries near the candidates' bounding box are tried until some lane
// finds a joined pair.  Light neighbours never test a heavy cell, so A
// decides every light neighbour; a pair of heavy cells is decided once, by
// the lower cell.  A joined neighbour also gets the mirrored bit (A lies in
// its mirrored stencil line), and A's own bits are OR-ed into its mask
// (lower heavy neighbours set bits there too).
template <int D, bool F32>
__device__ void adj_heavy(const Dev& d, const Inst& I, const Par& P, int32_t A, int32_t* s_qi,
                          uint32_t* s_qg, double* s_qx, Tally& t) {
  constexpr int NN = Near<D>::lines;
  const int lane = lane_id();
  const int32_t qs = d.cell_start[A], qe = d.cell_start[A + 1];
  const int32_t ncA = d.cell_ncore[A];
  if (ncA == 0) return;  // no core point: the mask stays clear
  const bool allc = ncA == qe - qs;
  uint32_t blo[D], bhi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    blo[k] = 0xffffffffu;
    bhi[k] = 0u;
  }
  int nq = 0;
  for (int32_t i0 = qs; i0 < qe; i0 += 64) {
    const int32_t i = i0 + lane;
    const bool v = i < qe && (allc || d.core[i]);
    uint32_t g[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      g[k] = v ? d.G[k][i] : 0u;
      blo[k] = v && g[k] < blo[k] ? g[k] : blo[k];
      bhi[k] = v && g[k] > bhi[k] ? g[k] : bhi[k];
    }
    const unsigned long long m = ballot(v);
    const int pos = nq + mbcnt(m);
    if (v && pos < kQCache) {
      s_qi[pos] = i;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        s_qg[pos * 3 + k] = g[k];
        if (!F32) s_qx[pos * 3 + k] = d.X[k][i];
      }
    }
    nq += __popcll(m);
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    blo[k] = wave_min(blo[k]);
    bhi[k] = wave_max(bhi[k]);
  }
  wave_sync();
  const bool cached = nq <= kQCache;
  const int NQ = cached ? nq : (qe - qs);
  const int32_t pkl = lane < NN ? rows_of(d, A)[lane] : 0;
  unsigned long long m0 = 0;
  for (int r = 0; r < NN; ++r) {
    const int32_t pk = rl_i(pkl, r);
    const int32_t len = pk >> kRowCntShift, st = pk & kRowCellMask;
    for (int k = 0; k < len; ++k) {
      const int32_t Bc = st + k;
      if (Bc == A) continue;
      const int ol = cell_offset_last<D>(d, I, A, Bc, r);
      if (ol < -1 || ol > 1) continue;  // a far cell
      const int32_t b0 = d.cell_start[Bc], b1 = d.cell_start[Bc + 1];
      if (d.cell_ncore[Bc] == 0) continue;
      if (Bc < A && b1 - b0 > kHeavy) continue;  // decided by the lower heavy cell
      bool edge = false;
      for (int32_t c0 = b0; c0 < b1 && !edge; c0 += 64) {
        const int32_t q = c0 + lane;
        bool cand = q < b1 && d.core[q];
        uint32_t cg[D];
        double cx[D];
#pragma unroll
        for (int kk = 0; kk < D; ++kk) cg[kk] = cand ? d.G[kk][q] : 0u;
        cand = cand && !(d2box<D>(cg, blo, bhi) > P.d2_out);
        if (!ballot(cand)) continue;
#pragma unroll
        for (int kk = 0; kk < D; ++kk) cx[kk] = (!F32 && cand) ? d.X[kk][q] : 0.0;
        uint32_t clo[D], chi[D];
#pragma unroll
        for (int kk = 0; kk < D; ++kk) {
          clo[kk] = wave_min(cand ? cg[kk] : 0xffffffffu);
          chi[kk] = wave_max(cand ? cg[kk] : 0u);
        }
        for (int i = 0; i < NQ && !edge; ++i) {
          int32_t ui;
          uint32_t ug[D];
          double ux[D];
          if (cached) {
            ui = s_qi[i];
#pragma unroll
            for (int kk = 0; kk < D; ++kk) {
              ug[kk] = s_qg[i * 3 + kk];
              ux[kk] = F32 ? 0.0 : s_qx[i * 3 + kk];
            }
          } else {
            ui = qs + i;
            if (!allc && !d.core[ui]) continue;
#pragma unroll
            for (int kk = 0; kk < D; ++kk) {
              ug[kk] = d.G[kk][ui];
              ux[kk] = F32 ? 0.0 : d.X[kk][ui];
            }
          }
          if (d2box<D>(ug, clo, chi) > P.d2_out) continue;
          const bool hit = cand && edge_pair<D, F32>(d, P, ui, q, ug, cg, ux, cx, t);
          edge = ballot(hit) != 0;
        }
      }
      if (edge) {
        m0 |= 1ull << (5 * r + k);
        const int rm = D == 3 ? kRowMirror3[r] : kRowMirror2[r];
        const int32_t pm = rows_of(d, Bc)[rm];
        const int bm = 5 * rm + (A - (pm & kRowCellMask));
        if (lane == 0) atomicOr(&d.amask[Bc], 1ull << bm);
      }
    }
  }
  if (lane == 0 && m0) atomicOr(&d.amask[A], m0);
}

template <bool F32>
__global__ __launch_bounds__(64) void k_adj_heavy(Dev d) {
  if (abandoned(d)) return;
  __shared__ int32_t s_qi[kQCache];
  __shared__ uint32_t s_qg[kQCache * 3];
  __shared__ double s_qx[F32 ? 1 : kQCache * 3];
  __shared__ long long s_pre[kStripes + 1];
  const long long nh = slist_prefix(d.heavy, s_pre);
  Tally t;
  for (long long h = blockIdx.x; h < nh; h += gridDim.x) {
    const int32_t A = slist_at(d.heavy, s_pre, h);
    const Inst& I = *inst_of_cell(d, A);
    const Par P(I);
    if (I.dim == 3)
      adj_heavy<3, F32>(d, I, P, A, s_qi, s_qg, s_qx, t);
    else
      adj_heavy<2, F32>(d, I, P, A, s_qi, s_qg, s_qx, t);
    wave_sync();  // the query cache is reused
  }
  stat_add(d, C_BAND, t.band);
  stat_add(d, C_ASYM, t.asym);
}

// ------------------------------------------------------------ far pairs --
// Component of every core cell after the expansion and merge (the root of
// its chain), and the compacted list of core cells.
__global__ __launch_bounds__(1024) void k_cell_comp(Dev d) {
  if (abandoned(d)) return;
  const int64_t C = d.ctl[C_NCELL];
  {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = c < C;
    const int64_t cc = in ? c : 0;
    // every load before the store (which could alias them as far as the
    // compiler knows), each level of dependent loads issued together
    const int32_t ncore = d.cell_ncore[cc], q0 = d.cell_start[cc], q1 = d.cell_start[cc + 1];
    const int32_t lab = d.cell_label[cc];
    const bool hc = in && ncore > 0;
    int32_t chain = -1, cmin = kNoKey, comp = -1;
    if (hc) {
      chain = lab;
      comp = d.parent[chain];
      // the sort is stable, so input indices increase within a cell: the
      // cell's smallest core index is that of its first core point
      int32_t q = q0;
      if (ncore != q1 - q0)
        while (!d.core[q]) ++q;
      cmin = d.orig[q];
    }
    if (in) d.cell_comp[c] = comp;
    // chain minimum: one atomic per distinct chain of the wave
    bool pend = hc;
    unsigned long long m = ballot(pend);
    while (m) {
      const int32_t lead = rl_i(chain, __ffsll((long long)m) - 1);
      const bool mine = pend && chain == lead;
      const int32_t v = wave_min(mine ? cmin : kNoKey);
      if (lane_id() == 0) atomicMin(&d.compmin[lead], v);
      pend = pend && !mine;
      m = ballot(pend);
    }
    block_append(hc, (int32_t)c, d.core_cells);
  }
}

// Core cell A, one lane: every far stencil cell above A (each unordered pair
// is taken by its lower cell) with core points in another component is
// decided; a joined pair unions the two components (lock-free union-find on
// the chain forest).
// Stencil lines whose cells can lie above a cell (offsets (a, b) >= (0, 0)
// in the key order, the own line included), in row order; -1 pads.
#define DCLUST_UPPER_ROWS3 {0, 2, 4, 7, 8, 10, 12, 15, 16, 18, 20, 23, 24, -1, -1}
#define DCLUST_UPPER_ROWS2 {0, 2, 4, -1, -1}
static __constant__ int8_t kUpperRows3[15] = DCLUST_UPPER_ROWS3;
static __constant__ int8_t kUpperRows2[5] = DCLUST_UPPER_ROWS2;
namespace tables {
constexpr int8_t up3[15] = DCLUST_UPPER_ROWS3;
constexpr int8_t up2[5] = DCLUST_UPPER_ROWS2;
constexpr bool above(int a, int b) { return a > 0 || (a == 0 && b >= 0); }
constexpr bool upper_ok() {
  int i = 0;
  for (int r = 0; r < 25; ++r)
    if (above(off3[r][0], off3[r][1]) && up3[i++] != r) return false;
  if (i != 13 || up3[13] != -1 || up3[14] != -1) return false;
  i = 0;
  for (int r = 0; r < 5; ++r)
    if (off2[r] >= 0 && up2[i++] != r) return false;
  if (i != 3) return false;
  // near rows of the adjacency pass (nibbles) and their mirrors
  for (int k = 0; k < 5; ++k) {
    const int r = (int)((Upper<3>::rows >> (4 * k)) & 15ull);
    if (r > 8 || !above(off3[r][0], off3[r][1]) || up3[k] != r) return false;
  }
  for (int r = 0; r < 9; ++r)
    if ((int)((Upper<3>::mirror >> (4 * r)) & 15ull) != mir3[r]) return false;
  for (int k = 0; k < 2; ++k)
    if ((int)((Upper<2>::rows >> (4 * k)) & 15ull) != up2[k]) return false;
  for (int r = 0; r < 3; ++r)
    if ((int)((Upper<2>::mirror >> (4 * r)) & 15ull) != mir2[r]) return false;
  return true;
}
static_assert(upper_ok(), "upper stencil lines");
}  // namespace tables

template <int D, bool F32>
__device__ long long far_cell(const Dev& d, const Inst& I, const Par& P, int32_t A, Tally& t,
                              long long& freeb) {
  constexpr int NR = D == 3 ? 13 : 3;  // upper lines
  const int32_t qs = d.cell_start[A], qe = d.cell_start[A + 1];
  const bool allc = d.cell_ncore[A] == qe - qs;
  const int32_t compA = d.cell_comp[A];
  const int32_t zA = d.cell_z[A];
  const RowsOf rows = rows_of(d, A);
  long long joins = 0;
  // five lines at a time: their packed rows in one round of loads, then the
  // components of all their cells above A in the next; bit 5j+k of `todo`
  // marks cell k of line j as above A in another component (cell_comp is
  // -1 for a cell without core points)
  for (int g = 0; g < NR; g += 5) {
    int32_t pk[5];
    int rr[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      rr[j] = D == 3 ? kUpperRows3[g + j] : kUpperRows2[g + j];
      pk[j] = rr[j] >= 0 ? rows[rr[j]] : 0;
    }
    unsigned todo = 0u;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const int32_t len = pk[j] >> kRowCntShift, st = pk[j] & kRowCellMask;
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int32_t comp = (k < len && st + k > A) ? d.cell_comp[st + k] : compA;
        if (comp >= 0 && comp != compA) todo |= 1u << (5 * j + k);
      }
    }
    while (todo) {
      const int b = __ffs(todo) - 1;
      todo &= todo - 1;
      const int j = (b * 205) >> 10;  // b / 5 for b < 25
      const int k = b - 5 * j;
      int32_t pkj = pk[0];
      int r = rr[0];
#pragma unroll
      for (int q = 1; q < 5; ++q) {
        pkj = j == q ? pk[q] : pkj;
        r = j == q ? rr[q] : r;
      }
      const int32_t B = (pkj & kRowCellMask) + k;
      const int o0 = D == 3 ? kRowOff3[r][0] : kRowOff2[r];
      const int o1 = D == 3 ? kRowOff3[r][1] : 0;
      const bool far_l = o0 == 2 || o0 == -2 || o1 == 2 || o1 == -2;
      // B's place along the line and its point range, loaded together
      const int32_t zB = d.cell_z[B], b0 = d.cell_start[B], b1 = d.cell_start[B + 1];
      const int ol = zB - zA;  // B lies in a neighbour line of A
      // in a near line only the cells two steps away along the line are far
      if (!far_l && ol != 2 && ol != -2) continue;
      if (cells_joined<D, F32>(d, P, qs, qe, allc, b0, b1, o0, o1, ol, t, freeb)) {
        uf_union(d.parent, compA, d.cell_comp[B]);
        ++joins;
      }
    }
  }
  (void)I;
  return joins;
}

template <bool F32>
__global__ __launch_bounds__(256, 5) void k_far(Dev d) {
  if (abandoned(d)) return;
  __shared__ long long s_pre[kStripes + 1];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < slist_prefix(d.core_cells, s_pre);
  const int32_t A = act ? slist_at(d.core_cells, s_pre, i) : 0;
  int dim = 0;
  const Inst* I = d.inst;
  if (act) {
    I = inst_of_cell(d, A);
    dim = I->dim;
  }
  Tally t;
  long long freeb = 0, joins = 0;
  const Par P(*I);
  DCLUST_BY_DIM(act, dim, (joins = far_cell<3, F32>(d, *I, P, A, t, freeb)),
                (joins = far_cell<2, F32>(d, *I, P, A, t, freeb)));
  stat_add(d, C_BAND, t.band);
  stat_add(d, C_ASYM, t.asym);
  stat_add(d, C_FAR_TESTS, t.eval);
  stat_add(d, C_FAR_JOINS, joins);
}

// Roots of every chain after the far-pair unions, in two passes.  The first
// compresses the forest with path-halving walks; a concurrent walk may
// leave an entry at an inner ancestor instead of the root, so the second
// pass walks the (now shallow) forest without writing anything but each
// chain's own root, and moves each chain's minimum core input index to its
// root: every entry is then exactly its root.
__global__ __launch_bounds__(256) void k_flatten_all(Dev d) {
  if (abandoned(d)) return;
  const long long nc = chain_ids(d);
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc;
       c += (long long)gridDim.x * blockDim.x)
    d.parent[c] = uf_find(d.parent, (int32_t)c);
}

__global__ __launch_bounds__(256) void k_roots_final(Dev d) {
  if (abandoned(d)) return;
  const long long nc = chain_ids(d);
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc;
       c += (long long)gridDim.x * blockDim.x) {
    const int32_t m = d.compmin[c];
    const int32_t r = uf_root(d.parent, (int32_t)c);
    d.parent[c] = r;
    if (m != kNoKey && r != c) atomicMin(&d.compmin[r], m);
  }
}

// ---------------------------------------------------------------- expand --
// Persistent expansion pass.  Wave w scans its own contiguous segment of
// the cells for unclaimed cells with core points and starts a new chain at
// each (device-side seeding; the reference pops host-side seeds,
// dbscan.cu:142-160); it also drains the correction queue left by the
// previous pass.  A wave pops up to 64 cells of its list at once, one per
// lane; lane l walks the adjacency bits of its cell: an unclaimed neighbour
// is claimed for the cell's chain (CAS) and pushed, a neighbour of another
// chain is a collision pair.  After `budget` cells a wave hands its list to
// the next pass and exits, so one long chain cannot hold a pass; a wave also
// exits when its list is empty and no seed source is left.  Every wave
// reaches the end of the kernel: there is no inter-wave waiting.
constexpr int kSeen = 256;   // per-wave table of recently recorded collision pairs

// The cells one wave popped in a round (one per lane).
struct PopStage {
  unsigned long long mask[64];  // near adjacency bits
  int32_t rows[9][64];          // near lines
  int32_t own[64];              // chain
  int32_t pre[64];              // exclusive prefix of the mask bit counts
};

__global__ __launch_bounds__(256) void k_expand(Dev d, int pass, int S, int budget) {
  extern __shared__ int32_t s_lists[];
  __shared__ unsigned long long s_seen[4][kSeen];
  __shared__ uint8_t s_tag[4][kSeen];  // per-wave dedup slots of one step
  __shared__ PopStage s_pop[4];         // per-wave popped cells of a round
  if (d.ctl[C_IDLE] || abandoned(d)) return;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const int nwb = blockDim.x >> 6;
  const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * nwb + wv;
  const int64_t W = (int64_t)gridDim.x * nwb;
  const int32_t C = (int32_t)d.ctl[C_NCELL];
  int32_t* list = s_lists + wv * S;
  unsigned long long* seen = s_seen[wv];
  uint8_t* tag = s_tag[wv];
  PopStage& pst = s_pop[wv];
  int32_t* qout = d.q[(pass + 1) & 1];
  unsigned long long* qout_n = (unsigned long long*)&d.ctl[C_Q0_N + ((pass + 1) & 1)];
  const int32_t* qin = d.q[pass & 1];
  unsigned long long* qin_head = (unsigned long long*)&d.ctl[C_Q0_HEAD + (pass & 1)];
  const long long qn = pass > 0 ? d.ctl[C_Q0_N + (pass & 1)] : 0;
  bool qlive = qn > 0;
  const int pcap = d.pair_per;
  int2* pbuf = d.pairs + gw * (int64_t)pcap;
  const int pc0 = rfl_i(d.pair_cnt[gw]);  // pairs recorded by earlier passes
  int pc = pc0;
  for (int i = lane; i < kSeen; i += 64) seen[i] = ~0ull;
  wave_sync();
  const unsigned long long t_begin = __builtin_amdgcn_s_memrealtime();
  // the seeding cursor over this wave's segment of cells persists across passes
  // (cell indices, below 2^31; the per-pass counters are bounded by the
  // pop budget: 32-bit, which keeps the wave-uniform state in fewer
  // scalar registers)
  int32_t seg = (int32_t)d.seg_pos[gw];
  const int32_t seg_end = (int32_t)((int64_t)C * (gw + 1) / W);
  int left = budget;
  int sl = 0;  // seed-list length (wave-uniform)
  int pops = 0, ovf = 0, edges = 0, seeds = 0, rounds = 0;
  int32_t cur = -1;    // chain of the cells in the list when not mixed
  bool mixed = false;  // the list holds queue cells (chains read per cell)
  for (;;) {
    if (left <= 0) {
      if (sl > 0) {
        long long base = 0;
        if (lane == 0) base = (long long)atomicAdd(qout_n, (unsigned long long)sl);
        base = __shfl(base, 0, 64);
        for (int i = lane; i < sl; i += 64) qout[base + i] = list[i];
      }
      break;
    }
    if (sl == 0) {
      bool got = false;
      if (qlive) {
        const int take = S < 64 ? S : 64;  // a refill fits the seed list
        long long base = 0;
        if (lane == 0) base = (long long)atomicAdd(qin_head, (unsigned long long)take);
        base = __shfl(base, 0, 64);
        if (base < qn) {
          const int cnt = (int)((qn - base) < take ? (qn - base) : take);
          if (lane < cnt) list[lane] = qin[base + lane];
          sl = cnt;
          got = true;
          mixed = true;  // queue cells: chains of other waves
        } else {
          qlive = false;
        }
      }
      while (!got && seg < seg_end) {
        const int32_t i = seg + lane;
        // both loads together (seg < seg_end, so seg is a cell)
        const int32_t ic = i < seg_end ? i : seg;
        const int32_t nco = d.cell_ncore[ic], lb = d.cell_label[ic];
        const bool cand = i < seg_end && nco > 0 && lb == kUnclaimed;
        unsigned long long m = ballot(cand);
        while (m) {
          const int l = __ffsll((long long)m) - 1;
          int ok = 0;
          if (lane == 0) {
            // a chain is named by its seed cell: unique, below the cell
            // count, and no shared counter to contend on
            const int32_t id = (int32_t)(seg + l);
            ok = atomicCAS(&d.cell_label[seg + l], kUnclaimed, id) == kUnclaimed;
          }
          ok = __shfl(ok, 0, 64);
          if (ok) {
            if (lane == 0) list[0] = (int32_t)(seg + l);
            ++seeds;
            sl = 1;
            got = true;
            cur = (int32_t)(seg + l);
            mixed = false;
            seg = seg + l + 1;
            break;
          }
          m &= m - 1;
        }
        if (!got) seg += 64;
      }
      if (!got) break;
      wave_sync();
    }
    // pop the top P cells, one per lane
    const int P = sl < 64 ? (sl < left ? sl : left) : (64 < left ? 64 : left);
    const bool act = lane < P;
    const int32_t A = act ? list[sl - P + lane] : 0;
    wave_sync();  // popped before the pushes below overwrite the slots
    sl -= P;
    left -= P;
    pops += P;
    ++rounds;
    int32_t own = 0;
    unsigned long long m0 = 0;
    if (act) {
      // A list filled only by this wave's seeding and pushes holds cells of
      // its current chain alone (a new chain is seeded only into an empty
      // list).  After a refill from the queue the chain is read: coherently,
      // since the claim may have been made moments ago on another XCD, whose
      // L2 is not coherent with this one for plain loads.
      own = mixed ? __hip_atomic_load(&d.cell_label[A], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT)
                  : cur;
      m0 = d.amask[A];
    }
    // the popped cells' chains, masks and near lines go to LDS (the lines
    // loaded together with the label and mask), and their adjacency bits are
    // dealt out evenly: item t of the round is bit k of the mask of lane
    // src, where src is the last lane whose bit prefix is at most t; a round
    // then takes ceil(bits / (64 T)) steps instead of the most bits of one
    // lane / T
    {
      // all nine loads in flight before the first LDS write
      const RowsOf rows = rows_of(d, A);
      int32_t rv[9];
#pragma unroll
      for (int r = 0; r < 9; ++r) rv[r] = act ? rows[r] : 0;
#pragma unroll
      for (int r = 0; r < 9; ++r) pst.rows[r][lane] = rv[r];
    }
    const int nb = __popcll(m0);
    edges += nb;
    const int incl = wave_incl_scan(nb);
    const int total = rfl_i(__shfl(incl, 63, 64));
    pst.own[lane] = own;
    pst.mask[lane] = m0;
    pst.pre[lane] = incl - nb;
    wave_sync();
    constexpr int T = 8;  // items per lane step
    for (int base = 0; base < total; base += 64 * T) {
      // sub-steps of 64 items that hold any (wave-uniform: a small round,
      // the usual case late in a chain, skips the empty ones)
      const int nsub = (total - base + 63) >> 6 < T ? (total - base + 63) >> 6 : T;
      int32_t tg[T], ow[T];
      bool v[T];
#pragma unroll
      for (int j = 0; j < T; ++j) {
        v[j] = false;
        tg[j] = 0;
        ow[j] = 0;
        if (j >= nsub) continue;
        const int t = base + j * 64 + lane;
        v[j] = t < total;
        int src = 0;
#pragma unroll
        for (int st = 32; st >= 1; st >>= 1)
          if (pst.pre[src + st] <= t) src += st;
        // bit k of the source mask: its k-th set bit
        int k = t - pst.pre[src];
        unsigned long long m = pst.mask[src];
        int b = 0;
#pragma unroll
        for (int w = 32; w >= 1; w >>= 1) {
          const int c = __popcll(m & ((1ull << w) - 1ull));
          if (k >= c) {
            k -= c;
            m >>= w;
            b += w;
          }
        }
        const int r = (b * 205) >> 10;  // b / 5 for 0 <= b < 128
        tg[j] = v[j] ? (pst.rows[r < 9 ? r : 0][src] & kRowCellMask) + (b - 5 * r) : 0;
        ow[j] = pst.own[src];
      }
      // a plain load can only be stale as kUnclaimed (a label is written
      // once), so only those cells pay for a CAS
      int32_t lab[T];
#pragma unroll
      for (int j = 0; j < T; ++j) lab[j] = (j < nsub && v[j]) ? d.cell_label[tg[j]] : ow[j];
      // the claims are all issued before any result is used, so that they
      // are in flight together
      bool cl[T], need[T];
      int32_t old[T];
#pragma unroll
      for (int j = 0; j < T; ++j) {
        need[j] = j < nsub && v[j] && lab[j] == kUnclaimed;
        old[j] = lab[j];
      }
#pragma unroll
      for (int j = 0; j < T; ++j)
        if (need[j]) old[j] = atomicCAS(&d.cell_label[tg[j]], kUnclaimed, ow[j]);
#pragma unroll
      for (int j = 0; j < T; ++j) {
        cl[j] = need[j] && old[j] == kUnclaimed;
        lab[j] = cl[j] ? ow[j] : old[j];
      }
#pragma unroll
      for (int j = 0; j < T; ++j) {
        if (j >= nsub) break;
        // pushes: one slot range per step
        const unsigned long long mc = ballot(cl[j]);
        if (mc) {
          const int pos = sl + mbcnt(mc);
          if (cl[j]) {
            if (pos < S) {
              list[pos] = tg[j];
            } else {
              const unsigned long long qi = atomicAdd(qout_n, 1ull);
              qout[qi] = tg[j];
              ++ovf;
            }
          }
          sl += __popcll(mc);
          sl = sl < S ? sl : S;
        }
        // collisions, recorded by all lanes at once: a pair already in the
        // wave's table of recent pairs is skipped; among the lanes of this
        // step with the same pair, the one whose id survives in the tag slot
        // of the pair's hash records it (a lane whose hash slot was won by a
        // different pair records its own; a rare duplicate is harmless)
        const bool coll = v[j] && lab[j] != ow[j];
        if (ballot(coll)) {
          const unsigned long long key =
              ((unsigned long long)(uint32_t)ow[j] << 32) | (uint32_t)lab[j];
          const int h =
              (int)(((uint32_t)ow[j] * 0x9E3779B1u ^ (uint32_t)lab[j] * 0x85EBCA6Bu) >> 24);
          bool nw = coll && seen[h] != key;
          if (nw) tag[h] = (uint8_t)lane;
          wave_sync();
          const int w = nw ? (int)tag[h] : lane;
          const uint32_t wlo = (uint32_t)__shfl((int)(uint32_t)key, w, 64);
          const uint32_t whi = (uint32_t)__shfl((int)(uint32_t)(key >> 32), w, 64);
          if (nw && w != lane && (((unsigned long long)whi << 32) | wlo) == key) nw = false;
          const unsigned long long mn = ballot(nw);
          if (nw) {
            seen[h] = key;
            const int slot = pc + (int)mbcnt(mn);
            if (slot < pcap)
              pbuf[slot] = make_int2(ow[j], lab[j]);
            else
              uf_union(d.parent, ow[j], lab[j]);  // buffer full: merge inline
          }
          pc += __popcll(mn);
        }
      }
    }
    wave_sync();
  }
  if (lane == 0) {
    // 100 MHz constant clock -> microseconds of this wave's life
    const unsigned long long us = (__builtin_amdgcn_s_memrealtime() - t_begin) / 100ull;
    d.pair_cnt[gw] = pc;
    d.seg_pos[gw] = seg;
    if (seg < seg_end) atomicAdd((unsigned long long*)&d.ctl[C_SEGS_LEFT], 1ull);
    if (pc > pc0) atomicAdd((unsigned long long*)&d.ctl[C_NPAIR], (unsigned long long)(pc - pc0));
    atomicAdd((unsigned long long*)&d.ctl[C_BUSY_SUM], us);
    atomicMax((unsigned long long*)&d.ctl[C_BUSY_MAX], us);
    atomicAdd((unsigned long long*)&d.ctl[C_POPS], (unsigned long long)pops);
    atomicMax((unsigned long long*)&d.ctl[C_POPS_MAX], (unsigned long long)pops);
    atomicMax((unsigned long long*)&d.ctl[C_SEEDS_MAX], (unsigned long long)seeds);
    atomicMax((unsigned long long*)&d.ctl[C_ROUNDS_MAX], (unsigned long long)rounds);
    if (pops) atomicMax((unsigned long long*)&d.ctl[C_LASTPASS], (unsigned long long)(pass + 1));
  }
  stat_add(d, C_OVF, ovf);
  stat_add(d, C_EDGES, edges);
  stat_add(d, C_NCHAIN, lane == 0 ? seeds : 0);  // seeds is wave-uniform
}



// Resets the counters of the queue that pass `pass` fills and of the head
// of the queue it drains, and marks the pass idle when nothing is pending.
__global__ void k_pass_prep(Dev d, int pass) {
  // nothing left by the previous pass: the waves of this one return at once
  d.ctl[C_IDLE] = pass > 0 && d.ctl[C_Q0_N + (pass & 1)] == 0 && d.ctl[C_SEGS_LEFT] == 0;
  d.ctl[C_Q0_N + ((pass + 1) & 1)] = 0;
  d.ctl[C_Q0_HEAD + (pass & 1)] = 0;
  d.ctl[C_SEGS_LEFT] = 0;
}

// After a pass: queue entries no wave took (every wave spent its budget)
// are appended to the queue of the next pass.  One block.  With `check` =
// the number of passes (after the last pass of a captured run), work still
// pending afterwards (queue or seeding segments) marks the run as abandoned,
// and the kernels after the expansion return at once (the host runs the call
// again).
__global__ __launch_bounds__(256) void k_pass_carry(Dev d, int pass, int check) {
  __shared__ long long s_base;
  long long qn = 0, head = 0;
  if (!abandoned(d) && pass > 0) {
    qn = d.ctl[C_Q0_N + (pass & 1)];
    head = d.ctl[C_Q0_HEAD + (pass & 1)];
  }
  if (head < qn) {  // block-uniform: every thread read the same two counters
    if (threadIdx.x == 0)
      s_base = (long long)atomicAdd((unsigned long long*)&d.ctl[C_Q0_N + ((pass + 1) & 1)],
                                    (unsigned long long)(qn - head));
    __syncthreads();
    const int32_t* qin = d.q[pass & 1];
    int32_t* qout = d.q[(pass + 1) & 1];
    for (long long i = head + threadIdx.x; i < qn; i += blockDim.x)
      qout[s_base + (i - head)] = qin[i];
  }
  if (check > 0 && threadIdx.x == 0) {
    // coherent loads: the add above went to L2, past this CU's cache
    const long long left = __hip_atomic_load(&d.ctl[C_Q0_N + (check & 1)], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    const long long segs = __hip_atomic_load(&d.ctl[C_SEGS_LEFT], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
    if (left != 0 || segs != 0) d.ctl[C_ABANDON] = 1;
  }
}

// ----------------------------------------------------------------- merge --
// Chain ids are seed cell indices.  When the id space fits in LDS (up to
// 32768 cells), one workgroup merges every recorded pair on an LDS copy of
// the chain forest with LDS atomics and writes the flattened roots back.
// Otherwise the pairs are merged in global memory by the whole grid (the
// ids are spread over the cell range, so the unions rarely share a cache
// line).
constexpr int kLdsChains = 32768;  // 128 KiB of LDS (gfx950 allows 160 KiB per workgroup)
constexpr int kMergeChunk = 4096;  // pair buffers scanned at a time (16 KiB of LDS)

__device__ __forceinline__ int32_t lds_find(int32_t* par, int32_t x) {
  for (;;) {
    int32_t p = par[x];
    if (p == x) return x;
    int32_t gp = par[p];
    if (gp != p) par[x] = gp;
    x = p;
  }
}

// number of expansion waves whose pair buffers are in use
__device__ __forceinline__ int64_t pair_waves(const Dev& d) { return d.pair_cap / d.pair_per; }

__global__ __launch_bounds__(1024) void k_merge_lds(Dev d) {
  if (abandoned(d)) return;
  __shared__ int32_t par[kLdsChains];
  __shared__ int32_t s_pre[kMergeChunk + 1];
  __shared__ int32_t s_wsum[16];
  const long long nc = chain_ids(d);
  if (nc > kLdsChains) return;
  // parent may already hold inline merges (pairs past a full buffer)
  for (int i = threadIdx.x; i < nc; i += blockDim.x) par[i] = d.parent[i];
  // the pair buffers of kMergeChunk expansion waves at a time: their counts
  // are scanned in LDS, then every thread takes pairs by flat index, so the
  // buffers are read in parallel
  const int64_t nw = pair_waves(d);
  const int per = d.pair_per;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  for (int64_t w0 = 0; w0 < nw; w0 += kMergeChunk) {
    const int nb = (int)((nw - w0) < kMergeChunk ? (nw - w0) : kMergeChunk);
    // per thread: kMergeChunk / 1024 consecutive buffers
    constexpr int kPer = kMergeChunk / 1024;
    int cnt[kPer];
    int sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const int b = threadIdx.x * kPer + j;
      cnt[j] = b < nb ? min(d.pair_cnt[w0 + b], per) : 0;
      sum += cnt[j];
    }
    const int incl = wave_incl_scan(sum);
    if (lane == 63) s_wsum[wv] = incl;
    __syncthreads();
    int off = incl - sum;
    for (int w = 0; w < wv; ++w) off += s_wsum[w];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      s_pre[threadIdx.x * kPer + j] = off;
      off += cnt[j];
    }
    if (threadIdx.x == blockDim.x - 1) s_pre[kMergeChunk] = off;
    __syncthreads();
    const int tot = s_pre[kMergeChunk];
    // kBatch pairs per thread at a time: fixed-step searches for their
    // buffers, then their loads, all independent, before the unions
    constexpr int kBatch = 8;
    for (int f0 = 0; f0 < tot; f0 += (int)blockDim.x * kBatch) {
      int2 pr[kBatch];
#pragma unroll
      for (int j = 0; j < kBatch; ++j) {
        const int f = f0 + j * (int)blockDim.x + (int)threadIdx.x;
        // last buffer whose range starts at or before f (s_pre[0] = 0)
        int lo = 0;
#pragma unroll
        for (int step = kMergeChunk / 2; step >= 1; step >>= 1)
          if (s_pre[lo + step] <= f) lo += step;
        pr[j] = f < tot ? d.pairs[(w0 + lo) * per + (f - s_pre[lo])] : make_int2(0, 0);
      }
#pragma unroll
      for (int j = 0; j < kBatch; ++j) {
        int32_t a = lds_find(par, pr[j].x), b = lds_find(par, pr[j].y);
        while (a != b) {
          if (a < b) {
            int32_t tt = a;
            a = b;
            b = tt;
          }
          int32_t old = atomicCAS(&par[a], a, b);
          if (old == a) break;
          a = lds_find(par, old);
          b = lds_find(par, b);
        }
      }
    }
    __syncthreads();  // s_pre is rewritten by the next chunk
  }
  for (int i = threadIdx.x; i < nc; i += blockDim.x) {
    int32_t r = i;
    while (par[r] != r) r = par[r];
    d.parent[i] = r;
  }
}

__global__ __launch_bounds__(256) void k_union_pairs(Dev d) {
  if (abandoned(d)) return;
  if (chain_ids(d) <= kLdsChains) return;
  const int64_t nw = pair_waves(d);
  const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nwv = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t w = wid; w < nw; w += nwv) {
    const int cnt = min(d.pair_cnt[w], d.pair_per);
    for (int k = lane_id(); k < cnt; k += 64) {
      int2 p = d.pairs[w * d.pair_per + k];
      uf_union(d.parent, p.x, p.y);
    }
  }
}

__global__ __launch_bounds__(256) void k_flatten(Dev d) {
  if (abandoned(d)) return;
  const long long nc = chain_ids(d);
  if (nc <= kLdsChains) return;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc;
       c += (long long)gridDim.x * blockDim.x)
    d.parent[c] = uf_find(d.parent, (int32_t)c);
}

// chain of a core point: the chain that claimed its cell
__device__ __forceinline__ int32_t core_chain(const Dev& d, int64_t i) {
  return d.core[i] ? d.cell_label[d.cell_of[i]] : -1;
}

// Points per thread step of the per-point relabel passes: each level of
// their dependent loads is issued for all of them before the next level,
// and nothing is stored until every load is done, so the latencies overlap.
constexpr int kRelU = 4;

__global__ __launch_bounds__(256) void k_keys(Dev d) {
  if (abandoned(d)) return;
  for (int64_t i0 = blockIdx.x * (int64_t)(kRelU * 256); i0 < d.n;
       i0 += (int64_t)gridDim.x * (kRelU * 256)) {
    int64_t ic[kRelU];
    bool cr[kRelU];
    int32_t c[kRelU], o[kRelU], li[kRelU], pr[kRelU], k[kRelU];
#pragma unroll
    for (int u = 0; u < kRelU; ++u) {
      const int64_t i = i0 + u * 256 + threadIdx.x;
      ic[u] = i < d.n ? i : 0;
      cr[u] = i < d.n && d.core[ic[u]];
      c[u] = d.cell_of[ic[u]];
      o[u] = d.orig[ic[u]];
    }
    // (a core point's cell is claimed by a chain; other lanes read index 0)
#pragma unroll
    for (int u = 0; u < kRelU; ++u) li[u] = cr[u] ? d.cell_label[c[u]] : 0;
#pragma unroll
    for (int u = 0; u < kRelU; ++u) pr[u] = d.parent[li[u]];
#pragma unroll
    for (int u = 0; u < kRelU; ++u) k[u] = d.compmin[pr[u]];
#pragma unroll
    for (int u = 0; u < kRelU; ++u) {
      const int64_t i = i0 + u * 256 + threadIdx.x;
      if (i < d.n) {
        d.keyv[i] = cr[u] ? k[u] : kNoKey;
        if (cr[u] && o[u] == k[u]) d.flag[k[u]] = 1;
      }
    }
  }
}
// ---------------------------------------------------------------- border --
// P(q -> p) for a core point q and a non-core point p, evaluated from the
// core point's side as its expansion in the reference oracle would.
template <int D, bool F32>
__device__ __forceinline__ bool pred_to(const Dev& d, const Par& P, int32_t q, int32_t p,
                                        const uint32_t* pg, const double* px, Tally& t) {
  double qx[D];
  if (F32) {
    uint32_t qg[D];
    load_g<D>(d, q, qg);
    const float s = d2fix<D>(qg, pg);
    if (s < P.d2_in) return true;
    if (s > P.d2_out) return false;
    t.band++;
    double ax[D];
    load_x<D, F32>(d, q, qx);
    load_x<D, F32>(d, p, ax);
    const bool ex = box_64<D>(qx, ax, P.eps) && d2_64<D>(qx, ax) <= P.eps2;
    t.flip += ex != (s <= P.d2_mid);
    return ex;
  }
  load_x<D, F32>(d, q, qx);
  return box_64<D>(qx, px, P.eps) && d2_64<D>(qx, px) <= P.eps2;
}

// Smallest component key over the core points q with P(q -> p): the own
// cell (a clique), then the recorded neighbours, or the whole stencil when
// the record overflowed or is absent.
template <int D, bool F32>
__device__ int32_t border_lane(const Dev& d, const Par& P, int32_t p, int32_t c, int32_t qs,
                               int32_t qe, int R, Tally& t, long long& ovf) {
  constexpr int NR = D == 3 ? 25 : 5;
  int32_t best = kNoKey;
  for (int32_t q = qs; q < qe; ++q)
    if (d.core[q]) {
      const int32_t k = d.keyv[q];
      best = k < best ? k : best;
    }
  uint32_t pg[D];
  double px[D];
  if (F32)
    load_g<D>(d, p, pg);
  else
    load_x<D, F32>(d, p, px);
  const int n = R > 0 ? (int)d.rec_n[p] : 256;
  if (n <= R) {
    const int32_t* rec = d.rec + (int64_t)p * R;
    // four records at a time: their indices, then their core flags and
    // keys, each level of loads issued together
    for (int j0 = 0; j0 < n; j0 += 4) {
      int32_t q[4], k[4];
      bool cq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) q[u] = j0 + u < n ? rec[j0 + u] : p;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        cq[u] = j0 + u < n && d.core[q[u]];
        k[u] = d.keyv[q[u]];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (cq[u] && k[u] < best && pred_to<D, F32>(d, P, q[u], p, pg, px, t)) best = k[u];
    }
    return best;
  }
  if (R > 0) ++ovf;
  const RowsOf rows = rows_of(d, c);
  for (int g = 0; g < NR; g += 5) {
    int32_t aa[5], ll[5];
    line_group(d, rows, g, aa, ll);
    Stream5 s;
    s.set(aa, ll);
    for (int32_t i = 0; i < s.tot; ++i) {
      int jr;
      const int32_t q = s.at(i, jr);
      if ((q >= qs && q < qe) || !d.core[q]) continue;
      const int32_t k = d.keyv[q];
      if (k < best && pred_to<D, F32>(d, P, q, p, pg, px, t)) best = k;
    }
  }
  return best;
}

// One lane per non-core point (the list k_core fills; a wave's points are
// consecutive ones of one workgroup there, in cell order).
template <bool F32>
__global__ __launch_bounds__(256) void k_border(Dev d) {
  if (abandoned(d)) return;
  __shared__ long long s_pre[kStripes + 1];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < slist_prefix(d.border_pts, s_pre);
  const int32_t p = act ? slist_at(d.border_pts, s_pre, i) : 0;
  int32_t c = 0, qs = 0, qe = 0;
  int dim = 0;
  const Inst* I = d.inst;
  if (act) {
    c = d.cell_of[p];
    qs = d.cell_start[c];
    qe = d.cell_start[c + 1];
    I = inst_of_cell(d, c);
    dim = I->dim;
  }
  Tally t;
  long long ovf = 0;
  int32_t best = kNoKey;
  const int R = d.rec_stride;
  const Par P(*I);
  DCLUST_BY_DIM(act, dim, (best = border_lane<3, F32>(d, P, p, c, qs, qe, R, t, ovf)),
                (best = border_lane<2, F32>(d, P, p, c, qs, qe, R, t, ovf)));
  if (act) d.keyv[p] = best;
  stat_add(d, C_BAND, t.band);
  stat_add(d, C_FLIP, t.flip);
  stat_add(d, C_RECOVF, ovf);
}

// ------------------------------------------------------------ rank + out --
// exclusive scan of the component-minimum flags (indexed by input index):
// rank[k] = number of clusters whose minimum core index is below k.
constexpr int RK_T = 256;
constexpr int RK_I = 16;
constexpr int RK_TILE = RK_T * RK_I;

// component-minimum flag of input index i (false past the end); the load is
// unconditional so that a thread's loads for all of its rounds overlap
__device__ __forceinline__ bool rank_flag(const Dev& d, int64_t i) {
  const uint8_t f = d.flag[i < d.n ? i : 0];
  return i < d.n && f;
}

__global__ __launch_bounds__(RK_T) void k_rank_count(Dev d) {
  if (abandoned(d)) return;
  const int64_t t0 = (int64_t)blockIdx.x * RK_TILE;
  bool f[RK_I];
#pragma unroll
  for (int r = 0; r < RK_I; ++r) f[r] = rank_flag(d, t0 + r * RK_T + threadIdx.x);
  int c = 0;
#pragma unroll
  for (int r = 0; r < RK_I; ++r) c += __popcll(ballot(f[r]));
  __shared__ int s[4];
  if (lane_id() == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) d.scan_tmp[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ __launch_bounds__(RK_T) void k_rank_write(Dev d) {
  if (abandoned(d)) return;
  const int64_t t0 = (int64_t)blockIdx.x * RK_TILE;
  __shared__ int s_w[4];
  __shared__ int64_t s_run;
  const int wv = threadIdx.x >> 6;
  bool fr[RK_I];
#pragma unroll
  for (int r = 0; r < RK_I; ++r) fr[r] = rank_flag(d, t0 + r * RK_T + threadIdx.x);
  if (threadIdx.x == 0) s_run = d.scan_tmp[blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RK_I; ++r) {
    int64_t i = t0 + r * RK_T + threadIdx.x;
    bool f = fr[r];
    unsigned long long m = ballot(f);
    if (lane_id() == 0) s_w[wv] = __popcll(m);
    __syncthreads();
    int64_t pre = s_run;
    for (int w = 0; w < wv; ++w) pre += s_w[w];
    if (i < d.n) d.rank[i] = (int32_t)(pre + mbcnt(m));
    if (i == d.n - 1) {
      d.rank[d.n] = (int32_t)(pre + mbcnt(m) + (f ? 1 : 0));
      d.ctl[C_NCLUST] = pre + mbcnt(m) + (f ? 1 : 0);
    }
    __syncthreads();
    if (threadIdx.x == 0) s_run += s_w[0] + s_w[1] + s_w[2] + s_w[3];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_output(Dev d) {
  if (abandoned(d)) return;
  long long noise = 0;
  for (int64_t i0 = blockIdx.x * (int64_t)(kRelU * 256); i0 < d.n;
       i0 += (int64_t)gridDim.x * (kRelU * 256)) {
    int32_t o[kRelU], k[kRelU], rk[kRelU], rb[kRelU];
    int64_t pb[kRelU];
#pragma unroll
    for (int u = 0; u < kRelU; ++u) {
      const int64_t i = i0 + u * 256 + threadIdx.x;
      const int64_t ic = i < d.n ? i : 0;
      o[u] = d.orig[ic];
      k[u] = d.keyv[ic];
      pb[u] = inst_of_cell(d, d.ninst == 1 ? 0 : d.cell_of[ic])->pt_begin;
    }
#pragma unroll
    for (int u = 0; u < kRelU; ++u) {
      rk[u] = d.rank[k[u] == kNoKey ? 0 : k[u]];
      rb[u] = d.rank[pb[u]];
    }
#pragma unroll
    for (int u = 0; u < kRelU; ++u) {
      const int64_t i = i0 + u * 256 + threadIdx.x;
      if (i < d.n) {
        const int32_t lab = k[u] == kNoKey ? 0 : rk[u] - rb[u] + 1;
        d.labels_out[o[u]] = lab;
        noise += lab == 0;
      }
    }
  }
  stat_add(d, C_NNOISE, noise);
}

// Per-run state of the clustering phases, one launch after the index:
// chain labels of the cells (unclaimed), the chain forest (every id its own
// root), component minima, component flags, per-cell core counts (all over
// the n entries that bound the cells and chain ids), and per expansion
// wave its pair count and the start of its seeding segment of the cells.
__global__ void k_init_run(Dev d, int64_t waves) {
  const int64_t m = d.n > waves ? d.n : waves;
  const int64_t C = d.ctl[C_NCELL];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i < d.n) {
      d.cell_label[i] = kUnclaimed;
      d.parent[i] = (int32_t)i;
      d.compmin[i] = kNoKey;
      d.flag[i] = 0;
      d.cell_ncore[i] = 0;
    }
    if (i < waves) {
      d.pair_cnt[i] = 0;
      d.seg_pos[i] = C * i / waves;
    }
  }
}

// ------------------------------------------------------------- launchers --
void launch_init_run(const Dev& d, int64_t waves, hipStream_t s) {
  const int64_t m = d.n > waves ? d.n : waves;
  int64_t b = (m + 255) / 256;
  b = b > 4096 ? 4096 : (b < 1 ? 1 : b);
  hipLaunchKernelGGL(k_init_run, dim3((int)b), dim3(256), 0, s, d, waves);
}

// one lane per point, grid a multiple of 8 (XCD remap)
static int point_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  b = (b + 7) / 8 * 8;
  return (int)(b < 8 ? 8 : b);
}

int64_t point_blocks(int64_t n) { return point_grid(n); }
int64_t cell_comp_blocks(int64_t n) { return (n + 1023) / 1024 > 0 ? (n + 1023) / 1024 : 1; }

void launch_core(const Dev& d, bool f32, hipStream_t s) {
  if (f32)
    hipLaunchKernelGGL(k_core<true>, dim3(point_grid(d.n)), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(k_core<false>, dim3(point_grid(d.n)), dim3(256), 0, s, d);
}

void launch_adj_light(const Dev& d, bool f32, hipStream_t s) {
  if (f32)
    hipLaunchKernelGGL(k_adj<true>, dim3(point_grid(d.n)), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(k_adj<false>, dim3(point_grid(d.n)), dim3(256), 0, s, d);
}

void launch_adj_heavy(const Dev& d, bool f32, int heavy_grid, hipStream_t s) {
  if (f32)
    hipLaunchKernelGGL(k_adj_heavy<true>, dim3(heavy_grid), dim3(64), 0, s, d);
  else
    hipLaunchKernelGGL(k_adj_heavy<false>, dim3(heavy_grid), dim3(64), 0, s, d);
}

void launch_adj(const Dev& d, bool f32, int heavy_grid, hipStream_t s) {
  launch_adj_light(d, f32, s);
  launch_adj_heavy(d, f32, heavy_grid, s);
}

void launch_pass_prep(const Dev& d, int pass, hipStream_t s) {
  hipLaunchKernelGGL(k_pass_prep, dim3(1), dim3(1), 0, s, d, pass);
}

void launch_pass_carry(const Dev& d, int pass, int check, hipStream_t s) {
  hipLaunchKernelGGL(k_pass_carry, dim3(1), dim3(256), 0, s, d, pass, check);
}

void launch_expand(const Dev& d, int pass, int grid, int seeds, int threads, int budget,
                   hipStream_t s) {
  size_t lds = (size_t)(threads / 64) * seeds * sizeof(int32_t);
  hipLaunchKernelGGL(k_expand, dim3(grid), dim3(threads), lds, s, d, pass, seeds, budget);
}

int expand_blocks_per_cu(int seeds, int threads) {
  int nb = 0;
  size_t lds = (size_t)(threads / 64) * seeds * sizeof(int32_t);
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_expand, threads, lds);
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

void launch_merge(const Dev& d, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_merge_lds, dim3(1), dim3(1024), 0, s, d);
  hipLaunchKernelGGL(k_union_pairs, dim3(grid), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_flatten, dim3(grid), dim3(256), 0, s, d);
}

void launch_far(const Dev& d, bool f32, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_cell_comp, dim3(cell_comp_blocks(d.n)), dim3(1024), 0, s, d);
  if (f32)
    hipLaunchKernelGGL(k_far<true>, dim3(point_grid(d.n)), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(k_far<false>, dim3(point_grid(d.n)), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_flatten_all, dim3(grid), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_roots_final, dim3(grid), dim3(256), 0, s, d);
}

void launch_border(const Dev& d, bool f32, hipStream_t s) {
  int g2 = (int)((d.n + 255) / 256);
  if (g2 > 4096) g2 = 4096;
  if (g2 < 1) g2 = 1;
  hipLaunchKernelGGL(k_keys, dim3(g2), dim3(256), 0, s, d);
  if (f32)
    hipLaunchKernelGGL(k_border<true>, dim3(point_grid(d.n)), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(k_border<false>, dim3(point_grid(d.n)), dim3(256), 0, s, d);
}

// core flags to input order (on request, after a run)
__global__ __launch_bounds__(256) void k_core_out(const uint8_t* core, const int32_t* orig,
                                                  int64_t n, uint8_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[orig[i]] = core[i];
}

void launch_core_out(const uint8_t* core, const int32_t* orig, int64_t n, uint8_t* out,
                     hipStream_t s) {
  int g = (int)((n + 255) / 256);
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_core_out, dim3(g), dim3(256), 0, s, core, orig, n, out);
}

void launch_relabel(const Dev& d, int64_t n, hipStream_t s) {
  const int nb = (int)((n + RK_TILE - 1) / RK_TILE);
  hipLaunchKernelGGL(k_rank_count, dim3(nb), dim3(RK_T), 0, s, d);
  launch_scan_top(d.scan_tmp, nb, s);
  hipLaunchKernelGGL(k_rank_write, dim3(nb), dim3(RK_T), 0, s, d);
  int g2 = (int)((n + 255) / 256);
  if (g2 > 4096) g2 = 4096;
  if (g2 < 1) g2 = 1;
  hipLaunchKernelGGL(k_output, dim3(g2), dim3(256), 0, s, d);
}

}  // namespace dclust
