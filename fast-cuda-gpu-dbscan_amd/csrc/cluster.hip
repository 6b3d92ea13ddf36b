// Clustering kernels on gfx950: the MI355X replacement for the reference's
// expansion kernel and host merge loop.
//
// Grid: cells of side (eps/2)(1+2^-20), so every cell is a clique (all its
// points are mutual neighbours) and every neighbour lies in the 5^D stencil
// of cells around a point's cell, stored per cell as 5^(D-1) packed runs of
// consecutive cells along the last axis ("lines").
//
//   k_core     neighbour counts -> core flags.  A cell holding >= minpts
//              points is all core without a distance test; otherwise the
//              count starts at the cell's own size and streams the other
//              stencil cells with early exit at minpts.  (dbscan.cu:331-364
//              counts neighbours while expanding; here the count is a
//              separate pass so every later kernel knows which candidates
//              are core.)
//   k_expand   persistent seed-list expansion (dbscan.cu:254-387) over
//              cells: each wavefront owns a chain and a seed list of cells in
//              LDS, pops a cell, looks up the labels of its stencil cells,
//              tests only cells that hold core points and are not yet its
//              own, claims adjacent unclaimed cells with one CAS each
//              (MarkAsCandidate, dbscan.cu:189-222), pushes them with
//              ballot/mbcnt and records collisions with other chains as
//              pairs in its own buffer (extraCollision, dbscan.cu:202-211).
//              Seed-list overflow and lists left when a wave's pop budget
//              runs out go to a correction queue that the next pass expands,
//              instead of being dropped (dbscan.cu:196-199).
//   k_merge_lds / k_union_pairs / k_flatten   GPU atomic union-find over
//              chain ids (MonitorSeedPoints' CPU BFS over the collision
//              matrix and thrust::replace relabels, dbscan.cu:60-125).
//   k_compmin / k_keys / k_border / rank scan / k_output   canonical labels:
//              cluster id = rank of the component's minimum core input
//              index; border = minimum id over adjacent cores; noise = 0
//              (GetDbscanResult, dbscan.cu:225-251, writes sorted labels
//              instead; this writes per-point labels).
//
// Neighbour predicate P(p,q), FP64 form (exactly the reference oracle's,
// cpu-dbscan/dbscan_rtree_exp.cpp:351-358,420-438 and Rtree.h:1573-1587):
// closed box fl(p-eps) <= q <= fl(p+eps) on every axis, and
// fl(fl(fl(dx0^2) + fl(dx1^2)) [+ fl(dx2^2)]) <= fl(eps*eps).  This file is
// compiled with -ffp-contract=off so no FMA is formed.
#include "internal.h"
#include "kernels.h"

namespace dclust {

constexpr int kSegMax = 128;  // segments of one stream (per-wave LDS table)

// ------------------------------------------------------- segment streams --
// A stream is the concatenation of up to kSegMax contiguous index ranges
// (runs of points, or runs of cells).  pref[k] = stream position of the
// first element of segment k; off[k] = segment start - pref[k].
struct Seg {
  int32_t* pref;
  int32_t* off;
  int K;
  int32_t M;
};

// Each lane contributes one range (start, len); empty ranges are dropped
// and the order of lanes is kept.
__device__ __forceinline__ void seg_build(int32_t start, int32_t len, int32_t* pref,
                                          int32_t* off, Seg& s) {
  const bool has = len > 0;
  const unsigned long long m = ballot(has);
  const int32_t incl = wave_incl_scan(len);
  const int32_t excl = incl - len;
  if (has) {
    const int k = mbcnt(m);
    pref[k] = excl;
    off[k] = start - excl;
  }
  s.pref = pref;
  s.off = off;
  s.K = __popcll(m);
  s.M = rl_i(incl, 63);
  wave_sync();
}

// element at stream position m (m < M); k receives its segment index
__device__ __forceinline__ int32_t seg_at(const Seg& s, int32_t m, int& k) {
  int lo = 0, hi = s.K - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s.pref[mid] <= m) lo = mid; else hi = mid - 1;
  }
  k = lo;
  return s.off[lo] + m;
}

// The stencil of cell c as runs of cells: lane r < NR holds line r, lane NR
// the part of the centre line after c.  With skip_self the cell itself is
// left out (the centre line is split around it).
template <int NR>
__device__ __forceinline__ void stencil_cells(const Dev& d, int32_t c, bool skip_self,
                                              int32_t& st, int32_t& len) {
  const int lane = lane_id();
  st = 0;
  len = 0;
  if (lane < NR) {
    const int32_t pk = d.rows[(int64_t)c * kMaxRows + lane];
    st = pk & kRowCellMask;
    len = pk >> kRowCntShift;
  }
  if (skip_self) {
    const int32_t s0 = rl_i(st, 0), e0 = s0 + rl_i(len, 0);  // centre line, contains c
    if (lane == 0) len = c - s0;
    if (lane == NR) {
      st = c + 1;
      len = e0 - (c + 1);
    }
  }
}

// The stencil of cell c as runs of points.
template <int NR>
__device__ __forceinline__ void stencil_points(const Dev& d, int32_t c, bool skip_self,
                                               int32_t* pref, int32_t* off, Seg& s) {
  int32_t st, len;
  stencil_cells<NR>(d, c, skip_self, st, len);
  int32_t ps = 0, pl = 0;
  if (len > 0) {
    ps = d.cell_start[st];
    pl = d.cell_start[st + len] - ps;
  }
  seg_build(ps, pl, pref, off, s);
}

// ------------------------------------------------------------ predicates --
// exact FP64 predicate centred on p (box bounds lo/hi precomputed from p)
template <int D>
__device__ __forceinline__ bool pred64(const double* p, const double* lo, const double* hi,
                                       const double* q, double eps2) {
  bool in = true;
#pragma unroll
  for (int k = 0; k < D; ++k) in = in && !(lo[k] > q[k] || q[k] > hi[k]);
  double t = p[0] - q[0];
  double s = t * t;
#pragma unroll
  for (int k = 1; k < D; ++k) {
    double u = p[k] - q[k];
    double v = u * u;
    s = s + v;
  }
  return in && s <= eps2;
}

// exact predicate centred on p: box bounds fl(p -+ eps) formed here
template <int D>
__device__ __forceinline__ bool pred64c(const double* p, const double* q, double eps,
                                        double eps2) {
  double lo[D], hi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    lo[k] = p[k] - eps;
    hi[k] = p[k] + eps;
  }
  return pred64<D>(p, lo, hi, q, eps2);
}

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

struct Tally {
  long long band = 0, flip = 0, asym = 0, eval = 0, loads = 0;
  long long pops = 0, batch = 0, xchunk = 0, xrel = 0;
};

// Query data held by a lane (one query per lane); the iteration over
// queries reads one lane's copy into wave-uniform registers.  The FP64 form
// keeps the FP64 coordinates, the FP32 form only the fixed-point ones (its
// rare guard-band rechecks load the FP64 coordinates from memory).
template <int D, bool F32>
struct Query {
  double x[D];
  uint32_t g[D];
};

template <int D, bool F32>
__device__ __forceinline__ void load_query_lane(const Dev& d, int32_t pt, bool active,
                                                Query<D, F32>& q) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (F32)
      q.g[k] = active ? d.G[k][pt] : 0u;
    else
      q.x[k] = active ? d.X[k][pt] : 0.0;
  }
}

template <int D, bool F32>
__device__ __forceinline__ void read_query(const Query<D, F32>& mine, int qi, Query<D, F32>& u) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (F32)
      u.g[k] = rl_u(mine.g[k], qi);
    else
      u.x[k] = rl_d(mine.x[k], qi);
  }
}

// Candidate coordinates held by a lane.
template <int D, bool F32>
struct Cand {
  double x[D];
  uint32_t g[D];
};

template <int D, bool F32>
__device__ __forceinline__ void load_cand(const Dev& d, int32_t pt, bool active, Cand<D, F32>& c) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (F32)
      c.g[k] = active ? d.G[k][pt] : 0u;
    else
      c.x[k] = active ? d.X[k][pt] : 0.0;
  }
}

// FP32 decision for the fixed-point pair (a, b), symmetric in a and b.
// Returns true inside; sets amb for the guard band.
template <int D>
__device__ __forceinline__ bool f32_decide(const Inst& I, const uint32_t* a, const uint32_t* b,
                                           bool active, bool& amb, float& s) {
  s = d2fix<D>(a, b);
  const bool in = s < I.d2_in;
  amb = active && !in && s <= I.d2_out;
  return in;
}

// P(query -> candidate): the query (point index qpt, wave-uniform) is the
// centre, as in its own neighbour count or expansion.  The FP32 form
// decides outside the guard band and re-decides inside it with the exact
// FP64 predicate (FP64 coordinates fetched only then).  With ASYM, pairs
// whose box test differs between the two directions are counted (d^2 is
// symmetric; outside the guard band the box test cannot differ).
template <int D, bool F32, bool ASYM>
__device__ __forceinline__ bool pred_qc(const Dev& d, const Inst& I, int32_t qpt,
                                        const Query<D, F32>& uq, const Cand<D, F32>& c,
                                        int32_t pt, bool active, Tally& t) {
  if (!F32) {
    const bool fwd = active && pred64c<D>(uq.x, c.x, I.eps, I.eps2);
    if (ASYM) {
      const bool rev = active && pred64c<D>(c.x, uq.x, I.eps, I.eps2);
      t.asym += fwd != rev;
    }
    return fwd;
  }
  bool amb;
  float s;
  bool in = f32_decide<D>(I, uq.g, c.g, active, amb, s);
  if (ballot(amb)) {
    double px[D], qx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) px[k] = d.X[k][qpt];
    if (amb) {
#pragma unroll
      for (int k = 0; k < D; ++k) qx[k] = d.X[k][pt];
      bool ex = pred64c<D>(px, qx, I.eps, I.eps2);
      if (ASYM) t.asym += ex != pred64c<D>(qx, px, I.eps, I.eps2);
      t.band++;
      t.flip += (ex != (s <= I.d2_mid));
      in = ex;
    }
  }
  return active && in;
}

// ------------------------------------------------------------------ core --
template <int D, bool F32>
__device__ void core_cell(const Dev& d, const Inst& I, int32_t c, int32_t* pref, int32_t* off,
                          Tally& t, long long& ncore, long long& dense) {
  constexpr int NR = D == 3 ? 25 : 5;
  const int lane = lane_id();
  const int32_t qs = d.cell_start[c], qe = d.cell_start[c + 1];
  const int32_t nc = qe - qs;
  if (nc >= I.minpts) {
    // a clique of at least minpts points: every point is core
    for (int32_t q = qs + lane; q < qe; q += 64) d.core[q] = 1;
    if (lane == 0) d.cell_ncore[c] = nc;
    ncore += nc;
    dense++;
    return;
  }
  Seg st;
  stencil_points<NR>(d, c, true, pref, off, st);
  int32_t ncc = 0;
  for (int32_t q0 = qs; q0 < qe; q0 += 64) {
    const int nq = (qe - q0) < 64 ? (qe - q0) : 64;
    Query<D, F32> mq;
    load_query_lane<D, F32>(d, q0 + lane, lane < nq, mq);
    int cnt = nc;  // the own cell is a clique, the point itself included
    unsigned long long done = nq == 64 ? 0ull : (~0ull << nq);
    for (int32_t m0 = 0; m0 < st.M && done != ~0ull; m0 += 64) {
      const int32_t m = m0 + lane;
      const bool valid = m < st.M;
      int k;
      const int32_t pt = valid ? seg_at(st, m, k) : 0;
      Cand<D, F32> cd;
      load_cand<D, F32>(d, pt, valid, cd);
      unsigned long long todo = ~done;
      const int nvalid = __popcll(ballot(valid));
      t.eval += (long long)__popcll(todo) * nvalid;
      t.loads += nvalid;
      while (todo) {
        const int qi = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        Query<D, F32> uq;
        read_query<D, F32>(mq, qi, uq);
        bool hit = pred_qc<D, F32, false>(d, I, q0 + qi, uq, cd, pt, valid, t);
        int pc = __popcll(ballot(hit));
        cnt += (lane == qi) ? pc : 0;
      }
      done |= ballot(cnt >= I.minpts);
    }
    const bool isc = lane < nq && cnt >= I.minpts;
    if (lane < nq) d.core[q0 + lane] = isc;
    ncc += __popcll(ballot(isc));
  }
  if (lane == 0) d.cell_ncore[c] = ncc;
  ncore += ncc;
}

template <bool F32>
__global__ __launch_bounds__(256) void k_core(Dev d) {
  __shared__ int32_t s_seg[4][2][kSegMax];
  const int wv = threadIdx.x >> 6;
  const int32_t C = (int32_t)d.ctl[C_NCELL];
  const int nwb = blockDim.x >> 6;
  const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * nwb + wv;
  const int64_t W = (int64_t)gridDim.x * nwb;
  constexpr int CH = 16;  // cells per wave step
  Tally t;
  long long ncore = 0, dense = 0;
  for (int64_t c0 = gw * CH; c0 < C; c0 += W * CH) {
    const int32_t ce = (int32_t)((c0 + CH) < C ? (c0 + CH) : C);
    for (int32_t c = (int32_t)c0; c < ce; ++c) {
      const Inst& I = d.inst[d.cell_inst[c]];
      if (I.dim == 3)
        core_cell<3, F32>(d, I, c, s_seg[wv][0], s_seg[wv][1], t, ncore, dense);
      else
        core_cell<2, F32>(d, I, c, s_seg[wv][0], s_seg[wv][1], t, ncore, dense);
    }
  }
  if (lane_id() == 0) {
    if (ncore) atomicAdd((unsigned long long*)&d.ctl[C_NCORE], (unsigned long long)ncore);
    if (dense) atomicAdd((unsigned long long*)&d.ctl[C_DENSE], (unsigned long long)dense);
    if (t.eval) atomicAdd((unsigned long long*)&d.ctl[C_EVAL], (unsigned long long)t.eval);
    if (t.loads) atomicAdd((unsigned long long*)&d.ctl[C_CANDLOAD], (unsigned long long)t.loads);
  }
  wave_add(&d.ctl[C_BAND_CORE], t.band);
  wave_add(&d.ctl[C_BAND], t.band);
  wave_add(&d.ctl[C_FLIP], t.flip);
}

// ---------------------------------------------------------------- expand --
// Per-wave LDS working set of the expansion.
struct ExpLds {
  int32_t* pref;   // [kSegMax] stream table
  int32_t* off;    // [kSegMax]
  int32_t* rcell;  // [kSegMax] cells to test (core points, not own chain)
  int32_t* rlab;   // [kSegMax] their labels when looked up
  int32_t* radj;   // [kSegMax] 1 once found adjacent to the expanded cell
  int32_t* qlist;  // [64] query points (core points of the expanded cell)
};

struct SeedList {
  int32_t* list;  // seed cells
  int S;
  int sl;
  int32_t* qout;
  long long* qout_n;
  int32_t cache;  // lanes 0..7: recently recorded foreign chains
  int cache_pos;
  int2* pbuf;     // this wave's collision-pair buffer (EXTRA_COLLISION_SIZE
  int pcnt;       // per chain in dbscan.cu:202-211), no shared counter
  int pcap;
};

// Record a chain-chain collision in the wave's own buffer; when it is full
// the two chains are merged inline (same final result).  Called by lane 0.
__device__ __forceinline__ void record_pair(const Dev& d, SeedList& L, int32_t a, int32_t b) {
  if (L.pcnt < L.pcap)
    L.pbuf[L.pcnt] = make_int2(a, b);
  else
    uf_union(d.parent, a, b);
}

// squared FP32 distance (fixed point) from point g to the box [lo, hi]
// (box of one cell's points: no wrap inside a cell)
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

// Expand cell A of chain `own`: find the stencil cells with core points
// that are not yet own, decide which are adjacent (some core pair within
// eps), claim the unclaimed ones and record the foreign ones.
template <int D, bool F32>
__device__ void expand_cell(const Dev& d, const Inst& I, int32_t A, int32_t own, SeedList& L,
                            ExpLds& X, Tally& t, long long& ovf) {
  constexpr int NR = D == 3 ? 25 : 5;
  const int lane = lane_id();
  // 1. stencil cells (A excluded) and the ones worth testing
  Seg cs;
  {
    int32_t st, len;
    stencil_cells<NR>(d, A, true, st, len);
    seg_build(st, len, X.pref, X.off, cs);
  }
  int nrel = 0;
  int32_t rcell[2] = {0, 0}, rlab[2] = {0, 0};
  bool rel[2] = {false, false};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int32_t j = h * 64 + lane;
    if (j < cs.M) {
      int k;
      const int32_t cj = seg_at(cs, j, k);
      const int32_t lab = d.cell_label[cj];
      const int32_t nco = d.cell_ncore[cj];
      rel[h] = nco > 0 && lab != own;
      rcell[h] = cj;
      rlab[h] = lab;
    }
  }
  wave_sync();  // the cell table is about to be overwritten
  t.xchunk++;
  {
    const unsigned long long m0 = ballot(rel[0]), m1 = ballot(rel[1]);
    const int n0 = __popcll(m0);
    nrel = n0 + __popcll(m1);
    if (nrel == 0) return;
    t.xrel++;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (rel[h]) {
        const int k = (h ? n0 + mbcnt(m1) : mbcnt(m0));
        X.rcell[k] = rcell[h];
        X.rlab[k] = rlab[h];
        X.radj[k] = 0;
      }
    }
    wave_sync();
  }
  // point runs of the cells to test, one segment per cell (cell k <-> segment k)
  Seg ps;
  {
    int32_t s0 = 0, l0 = 0, s1 = 0, l1 = 0;
    if (lane < nrel) {
      const int32_t cc = X.rcell[lane];
      s0 = d.cell_start[cc];
      l0 = d.cell_start[cc + 1] - s0;
    }
    if (lane + 64 < nrel) {
      const int32_t cc = X.rcell[lane + 64];
      s1 = d.cell_start[cc];
      l1 = d.cell_start[cc + 1] - s1;
    }
    // two rounds of seg_build into one table
    const unsigned long long m0 = ballot(l0 > 0);
    const int32_t i0 = wave_incl_scan(l0);
    const int32_t e0 = i0 - l0;
    if (l0 > 0) {
      const int k = mbcnt(m0);
      X.pref[k] = e0;
      X.off[k] = s0 - e0;
    }
    const int K0 = __popcll(m0);
    const int32_t M0 = rl_i(i0, 63);
    const unsigned long long m1 = ballot(l1 > 0);
    const int32_t i1 = wave_incl_scan(l1);
    const int32_t e1 = M0 + i1 - l1;
    if (l1 > 0) {
      const int k = K0 + mbcnt(m1);
      X.pref[k] = e1;
      X.off[k] = s1 - e1;
    }
    ps.pref = X.pref;
    ps.off = X.off;
    ps.K = K0 + __popcll(m1);
    ps.M = M0 + rl_i(i1, 63);
    wave_sync();
  }
  // 2. queries: core points of A, up to 64 at a time
  const int32_t as = d.cell_start[A], ae = d.cell_start[A + 1];
  const bool a_all_core = d.cell_ncore[A] == ae - as;
  for (int32_t q0 = as; q0 < ae; q0 += 64) {
    const bool qa = q0 + lane < ae && (a_all_core || d.core[q0 + lane]);
    const unsigned long long qm = ballot(qa);
    if (!qm) continue;
    if (qa) X.qlist[mbcnt(qm)] = q0 + lane;
    wave_sync();
    const int nq = __popcll(qm);
    const int32_t qpt = lane < nq ? X.qlist[lane] : 0;
    Query<D, F32> mq;
    load_query_lane<D, F32>(d, qpt, lane < nq, mq);
    // bounding box of the queries (fixed point; one cell, so no wrap)
    uint32_t blo[D], bhi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const uint32_t g = lane < nq ? d.G[k][qpt] : 0u;
      blo[k] = wave_min(lane < nq ? g : 0xffffffffu);
      bhi[k] = wave_max(lane < nq ? g : 0u);
    }
    // 3. stream the candidate cells' points
    for (int32_t m0 = 0; m0 < ps.M; m0 += 64) {
      const int32_t m = m0 + lane;
      const bool valid = m < ps.M;
      int k = 0;
      const int32_t pt = valid ? seg_at(ps, m, k) : 0;
      bool act = valid && !X.radj[k];
      if (!ballot(act)) continue;
      uint32_t cg[D];
#pragma unroll
      for (int a = 0; a < D; ++a) cg[a] = act ? d.G[a][pt] : 0u;
      if (act) act = d.core[pt] && !(d2box<D>(cg, blo, bhi) > I.d2_out);
      if (!ballot(act)) continue;
      Cand<D, F32> cd;
      if (F32) {
#pragma unroll
        for (int a = 0; a < D; ++a) cd.g[a] = cg[a];
      } else {
        load_cand<D, F32>(d, pt, act, cd);
      }
      bool hit = false;
      for (int qi = 0; qi < nq; ++qi) {
        Query<D, F32> uq;
        read_query<D, F32>(mq, qi, uq);
        hit |= pred_qc<D, F32, true>(d, I, rl_i(qpt, qi), uq, cd, pt, act && !hit, t);
        if (!ballot(act && !hit)) break;
      }
      if (hit) X.radj[k] = 1;
      wave_sync();
    }
  }
  // 4. claims and collisions
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int j = h * 64 + lane;
    const bool adj = j < nrel && X.radj[j];
    const int32_t cj = adj ? X.rcell[j] : 0;
    int32_t lab = adj ? X.rlab[j] : own;
    bool claimed = false;
    if (adj && lab == kUnclaimed) {
      const int32_t old = atomicCAS(&d.cell_label[cj], kUnclaimed, own);
      claimed = old == kUnclaimed;
      lab = claimed ? own : old;
    }
    bool coll = adj && !claimed && lab != own;
    unsigned long long cm = ballot(coll);
    while (cm) {
      const int l = __ffsll((long long)cm) - 1;
      const int32_t v = rl_i(lab, l);
      const bool seen = ballot(lane < 8 && L.cache == v) != 0;
      if (!seen) {
        if (lane == 0) record_pair(d, L, own, v);
        L.pcnt++;
        if (lane == L.cache_pos) L.cache = v;
        L.cache_pos = (L.cache_pos + 1) & 7;
      }
      coll = coll && lab != v;
      cm = ballot(coll);
    }
    const unsigned long long pm = ballot(claimed);
    if (pm) {
      const int cnt = __popcll(pm);
      const int pos = L.sl + mbcnt(pm);
      const int over = L.sl + cnt - L.S;
      long long base = 0;
      if (over > 0) {
        if (lane == 0) base = (long long)atomicAdd((unsigned long long*)L.qout_n, (unsigned long long)over);
        base = __shfl(base, 0, 64);
        ovf += over;
      }
      if (claimed) {
        if (pos < L.S)
          L.list[pos] = cj;
        else
          L.qout[base + pos - L.S] = cj;
      }
      L.sl = (L.sl + cnt) < L.S ? (L.sl + cnt) : L.S;
    }
  }
  wave_sync();
}

// Persistent expansion pass.  Wave w scans its own contiguous segment of
// the cells for unclaimed cells with core points and starts a new chain at
// each (device-side seeding; the reference pops host-side seeds,
// dbscan.cu:142-160); it also drains the correction queue left by the
// previous pass.  After `budget` pops a wave hands its remaining list to
// the next pass and exits, so one long chain cannot hold a pass; a wave
// also exits when its list is empty and no seed source is left.  Every wave
// reaches the end of the kernel: there is no inter-wave waiting.
template <bool F32>
__global__ __launch_bounds__(256) void k_expand(Dev d, int pass, int S, int budget) {
  extern __shared__ int32_t s_lists[];
  __shared__ int32_t s_x[4][5][kSegMax];
  __shared__ int32_t s_q[4][64];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const int nwb = blockDim.x >> 6;
  const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * nwb + wv;
  const int64_t W = (int64_t)gridDim.x * nwb;
  const int32_t C = (int32_t)d.ctl[C_NCELL];
  ExpLds X{s_x[wv][0], s_x[wv][1], s_x[wv][2], s_x[wv][3], s_x[wv][4], s_q[wv]};
  SeedList L;
  L.list = s_lists + wv * S;
  L.S = S;
  L.sl = 0;
  L.qout = d.q[(pass + 1) & 1];
  L.qout_n = &d.ctl[C_Q0_N + ((pass + 1) & 1)];
  L.cache = -1;
  L.cache_pos = 0;
  L.pcap = d.pair_per;
  L.pbuf = d.pairs + gw * (int64_t)d.pair_per;
  L.pcnt = rfl_i(d.pair_cnt[gw]);  // pairs recorded by earlier passes
  const int pcnt0 = L.pcnt;
  const unsigned long long t_begin = __builtin_amdgcn_s_memrealtime();
  const int32_t* qin = d.q[pass & 1];
  long long* qin_head = &d.ctl[C_Q0_HEAD + (pass & 1)];
  const long long qn = pass > 0 ? d.ctl[C_Q0_N + (pass & 1)] : 0;
  bool qlive = qn > 0;
  // the seeding cursor over this wave's segment of cells persists across passes
  int64_t seg = d.seg_pos[gw];
  const int64_t seg_end = (int64_t)C * (gw + 1) / W;
  int pops_left = budget;
  int32_t cache_own = -2;
  Tally t;
  long long ovf = 0;
  for (;;) {
    if (pops_left == 0) {
      if (L.sl > 0) {
        long long base = 0;
        if (lane == 0) base = (long long)atomicAdd((unsigned long long*)L.qout_n, (unsigned long long)L.sl);
        base = __shfl(base, 0, 64);
        for (int i = lane; i < L.sl; i += 64) L.qout[base + i] = L.list[i];
        L.sl = 0;
      }
      break;
    }
    if (L.sl == 0) {
      bool got = false;
      if (qlive) {
        long long base = 0;
        if (lane == 0) base = (long long)atomicAdd((unsigned long long*)qin_head, 64ull);
        base = __shfl(base, 0, 64);
        if (base < qn) {
          const int cnt = (int)((qn - base) < 64 ? (qn - base) : 64);
          if (lane < cnt) L.list[lane] = qin[base + lane];
          L.sl = cnt;
          got = true;
        } else {
          qlive = false;
        }
      }
      while (!got && seg < seg_end) {
        const int64_t i = seg + lane;
        const bool cand = i < seg_end && d.cell_ncore[i] > 0 && d.cell_label[i] == kUnclaimed;
        unsigned long long m = ballot(cand);
        while (m) {
          const int l = __ffsll((long long)m) - 1;
          int ok = 0;
          if (lane == 0) {
            const int32_t id = (int32_t)atomicAdd((unsigned long long*)&d.ctl[C_NCHAIN], 1ull);
            ok = atomicCAS(&d.cell_label[seg + l], kUnclaimed, id) == kUnclaimed;
          }
          ok = __shfl(ok, 0, 64);
          if (ok) {
            if (lane == 0) L.list[0] = (int32_t)(seg + l);
            L.sl = 1;
            got = true;
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
    // pop one cell
    const int32_t A = L.list[L.sl - 1];
    L.sl -= 1;
    wave_sync();
    const int32_t own = d.cell_label[A];
    if (own < 0) continue;  // cannot happen: list entries are claimed cells
    pops_left--;
    t.pops++;
    t.batch += 1;
    if (own != cache_own) {
      L.cache = -1;
      L.cache_pos = 0;
      cache_own = own;
    }
    const Inst& I = d.inst[d.cell_inst[A]];
    if (I.dim == 3)
      expand_cell<3, F32>(d, I, A, own, L, X, t, ovf);
    else
      expand_cell<2, F32>(d, I, A, own, L, X, t, ovf);
  }
  if (lane_id() == 0) {
    // 100 MHz constant clock -> microseconds of this wave's life
    const unsigned long long us = (__builtin_amdgcn_s_memrealtime() - t_begin) / 100ull;
    d.pair_cnt[gw] = L.pcnt;
    d.seg_pos[gw] = seg;
    if (seg < seg_end) atomicAdd((unsigned long long*)&d.ctl[C_SEGS_LEFT], 1ull);
    if (L.pcnt > pcnt0)
      atomicAdd((unsigned long long*)&d.ctl[C_NPAIR], (unsigned long long)(L.pcnt - pcnt0));
    atomicAdd((unsigned long long*)&d.ctl[C_BUSY_SUM], us);
    atomicMax((unsigned long long*)&d.ctl[C_BUSY_MAX], us);
    if (ovf) atomicAdd((unsigned long long*)&d.ctl[C_OVF], (unsigned long long)ovf);
    atomicAdd((unsigned long long*)&d.ctl[C_POPS], (unsigned long long)t.pops);
    atomicAdd((unsigned long long*)&d.ctl[C_BATCH], (unsigned long long)t.batch);
    atomicAdd((unsigned long long*)&d.ctl[C_XCHUNK], (unsigned long long)t.xchunk);
    atomicAdd((unsigned long long*)&d.ctl[C_XREL], (unsigned long long)t.xrel);
  }
  wave_add(&d.ctl[C_BAND], t.band);
  wave_add(&d.ctl[C_FLIP], t.flip);
  wave_add(&d.ctl[C_ASYM], t.asym);
}

// Resets the counters of the queue that pass `pass` fills and of the head
// of the queue it drains.
__global__ void k_pass_prep(Dev d, int pass) {
  d.ctl[C_Q0_N + ((pass + 1) & 1)] = 0;
  d.ctl[C_Q0_HEAD + (pass & 1)] = 0;
  d.ctl[C_SEGS_LEFT] = 0;
}

// After a pass: queue entries no wave took (every wave spent its budget)
// are appended to the queue of the next pass.  One block.
__global__ __launch_bounds__(256) void k_pass_carry(Dev d, int pass) {
  if (pass == 0) return;
  const long long qn = d.ctl[C_Q0_N + (pass & 1)];
  const long long head = d.ctl[C_Q0_HEAD + (pass & 1)];
  if (head >= qn) return;
  __shared__ long long s_base;
  if (threadIdx.x == 0)
    s_base = (long long)atomicAdd((unsigned long long*)&d.ctl[C_Q0_N + ((pass + 1) & 1)],
                                  (unsigned long long)(qn - head));
  __syncthreads();
  const int32_t* qin = d.q[pass & 1];
  int32_t* qout = d.q[(pass + 1) & 1];
  for (long long i = head + threadIdx.x; i < qn; i += blockDim.x) qout[s_base + (i - head)] = qin[i];
}

// Segment start of every expansion wave (seeding cursor over the cells).
__global__ void k_seg_init(Dev d, int64_t W) {
  const int64_t C = d.ctl[C_NCELL];
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < W;
       w += (int64_t)gridDim.x * blockDim.x)
    d.seg_pos[w] = C * w / W;
}

// ----------------------------------------------------------------- merge --
// When the chains fit in LDS (the usual case: a few thousand to a few tens
// of thousands of chains), one workgroup merges every recorded pair on an
// LDS copy of the chain forest with LDS atomics and writes the flattened
// roots back.  Otherwise the pairs are merged in global memory.
constexpr int kLdsChains = 16384;  // 64 KiB of LDS

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
  __shared__ int32_t par[kLdsChains];
  const long long nc = d.ctl[C_NCHAIN];
  if (nc > kLdsChains) return;
  // parent may already hold inline merges (pairs past a full buffer)
  for (int i = threadIdx.x; i < nc; i += blockDim.x) par[i] = d.parent[i];
  __syncthreads();
  // each wavefront of the block takes one expansion wave's buffer at a time
  // and merges its pairs across its lanes
  const int64_t nw = pair_waves(d);
  for (int64_t w = threadIdx.x >> 6; w < nw; w += blockDim.x >> 6) {
    const int cnt = min(d.pair_cnt[w], d.pair_per);
    for (int k = lane_id(); k < cnt; k += 64) {
      int2 pr = d.pairs[w * d.pair_per + k];
      int32_t a = lds_find(par, pr.x), b = lds_find(par, pr.y);
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
  __syncthreads();
  for (int i = threadIdx.x; i < nc; i += blockDim.x) {
    int32_t r = i;
    while (par[r] != r) r = par[r];
    d.parent[i] = r;
  }
}

__global__ __launch_bounds__(256) void k_union_pairs(Dev d) {
  if (d.ctl[C_NCHAIN] <= kLdsChains) return;
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
  const long long nc = d.ctl[C_NCHAIN];
  if (nc <= kLdsChains) return;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc;
       c += (long long)gridDim.x * blockDim.x)
    d.parent[c] = uf_find(d.parent, (int32_t)c);
}

// chain of a core point: the chain that claimed its cell
__device__ __forceinline__ int32_t core_chain(const Dev& d, int64_t i) {
  return d.core[i] ? d.cell_label[d.cell_of[i]] : -1;
}

// minimum input index per component (wave-aggregated per distinct root)
__global__ __launch_bounds__(256) void k_compmin(Dev d) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); i0 < d.n; i0 += stride) {
    const int64_t i = i0 + lane_id();
    const int32_t li = i < d.n ? core_chain(d, i) : -1;
    bool cr = li >= 0;  // every core cell is claimed; the guard only keeps a bug in bounds
    const int32_t r = cr ? d.parent[li] : -1;
    const int32_t o = cr ? d.orig[i] : kNoKey;
    unsigned long long m = ballot(cr);
    while (m) {
      const int l = __ffsll((long long)m) - 1;
      const int32_t rr = rl_i(r, l);
      const bool mine = cr && r == rr;
      const int32_t mn = wave_min(mine ? o : kNoKey);
      if (lane_id() == 0) atomicMin(&d.compmin[rr], mn);
      cr = cr && !mine;
      m = ballot(cr);
    }
  }
}

__global__ __launch_bounds__(256) void k_keys(Dev d) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < d.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t li = core_chain(d, i);
    if (li >= 0) {
      const int32_t k = d.compmin[d.parent[li]];
      d.keyv[i] = k;
      if (d.orig[i] == k) d.flag[k] = 1;
    } else {
      d.keyv[i] = kNoKey;
    }
  }
}

// ---------------------------------------------------------------- border --
// Non-core points of a cell (queries) take the minimum component key over
// the core candidates q of the stencil (own cell included) with
// P(q -> query), evaluated from the core point's side as its expansion in
// the reference oracle would.
template <int D, bool F32>
__device__ void border_cell(const Dev& d, const Inst& I, int32_t c, int32_t* pref, int32_t* off,
                            Tally& t) {
  constexpr int NR = D == 3 ? 25 : 5;
  const int lane = lane_id();
  const int32_t qs = d.cell_start[c], qe = d.cell_start[c + 1];
  if (d.cell_ncore[c] == qe - qs) return;  // every point is core
  Seg st;
  stencil_points<NR>(d, c, false, pref, off, st);
  for (int32_t q0 = qs; q0 < qe; q0 += 64) {
    const int nq = (qe - q0) < 64 ? (qe - q0) : 64;
    const bool isq = lane < nq && !d.core[q0 + lane];
    unsigned long long qm = ballot(isq);
    if (!qm) continue;
    double qx[D];
    uint32_t qg[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      qx[k] = (!F32 && isq) ? d.X[k][q0 + lane] : 0.0;
      qg[k] = (F32 && isq) ? d.G[k][q0 + lane] : 0u;
    }
    int32_t best = kNoKey;
    for (int32_t m0 = 0; m0 < st.M; m0 += 64) {
      const int32_t m = m0 + lane;
      const bool valid = m < st.M;
      int sk;
      const int32_t pt = valid ? seg_at(st, m, sk) : 0;
      const bool cc = valid && d.core[pt];
      if (!ballot(cc)) continue;
      double cx[D], clo[D], chi[D];
      uint32_t cg[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        cx[k] = (cc && !F32) ? d.X[k][pt] : 0.0;
        clo[k] = cx[k] - I.eps;
        chi[k] = cx[k] + I.eps;
        cg[k] = (F32 && cc) ? d.G[k][pt] : 0u;
      }
      const int32_t ck = cc ? d.keyv[pt] : kNoKey;
      unsigned long long todo = qm;
      while (todo) {
        const int qi = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        bool hit;
        if (F32) {
          uint32_t ug[D];
#pragma unroll
          for (int k = 0; k < D; ++k) ug[k] = rl_u(qg[k], qi);
          float s = d2fix<D>(cg, ug);
          bool in = s < I.d2_in;
          bool amb = cc && !in && s <= I.d2_out;
          if (ballot(amb)) {
            double ux[D];
#pragma unroll
            for (int k = 0; k < D; ++k) ux[k] = d.X[k][q0 + qi];
            if (amb) {
              double px[D], plo[D], phi[D];
#pragma unroll
              for (int k = 0; k < D; ++k) {
                px[k] = d.X[k][pt];
                plo[k] = px[k] - I.eps;
                phi[k] = px[k] + I.eps;
              }
              bool ex = pred64<D>(px, plo, phi, ux, I.eps2);
              t.band++;
              t.flip += (ex != (s <= I.d2_mid));
              in = ex;
            }
          }
          hit = cc && in;
        } else {
          double ux[D];
#pragma unroll
          for (int k = 0; k < D; ++k) ux[k] = rl_d(qx[k], qi);
          hit = cc && pred64<D>(cx, clo, chi, ux, I.eps2);
        }
        if (ballot(hit)) {
          const int32_t mn = wave_min(hit ? ck : kNoKey);
          if (lane == qi && mn < best) best = mn;
        }
      }
    }
    if (isq) d.keyv[q0 + lane] = best;
  }
}

template <bool F32>
__global__ __launch_bounds__(256) void k_border(Dev d) {
  __shared__ int32_t s_seg[4][2][kSegMax];
  const int wv = threadIdx.x >> 6;
  const int32_t C = (int32_t)d.ctl[C_NCELL];
  const int nwb = blockDim.x >> 6;
  const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * nwb + wv;
  const int64_t W = (int64_t)gridDim.x * nwb;
  constexpr int CH = 16;
  Tally t;
  for (int64_t c0 = gw * CH; c0 < C; c0 += W * CH) {
    const int32_t ce = (int32_t)((c0 + CH) < C ? (c0 + CH) : C);
    for (int32_t c = (int32_t)c0; c < ce; ++c) {
      const Inst& I = d.inst[d.cell_inst[c]];
      if (I.dim == 3)
        border_cell<3, F32>(d, I, c, s_seg[wv][0], s_seg[wv][1], t);
      else
        border_cell<2, F32>(d, I, c, s_seg[wv][0], s_seg[wv][1], t);
    }
  }
  wave_add(&d.ctl[C_BAND], t.band);
  wave_add(&d.ctl[C_FLIP], t.flip);
}

// ------------------------------------------------------------ rank + out --
// exclusive scan of the component-minimum flags (indexed by input index):
// rank[k] = number of clusters whose minimum core index is below k.
constexpr int RK_T = 256;
constexpr int RK_I = 16;
constexpr int RK_TILE = RK_T * RK_I;

__global__ __launch_bounds__(RK_T) void k_rank_count(Dev d) {
  const int64_t t0 = (int64_t)blockIdx.x * RK_TILE;
  int c = 0;
  for (int r = 0; r < RK_I; ++r) {
    int64_t i = t0 + r * RK_T + threadIdx.x;
    c += __popcll(ballot(i < d.n && d.flag[i]));
  }
  __shared__ int s[4];
  if (lane_id() == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) d.scan_tmp[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ __launch_bounds__(RK_T) void k_rank_write(Dev d) {
  const int64_t t0 = (int64_t)blockIdx.x * RK_TILE;
  __shared__ int s_w[4];
  __shared__ int64_t s_run;
  const int wv = threadIdx.x >> 6;
  if (threadIdx.x == 0) s_run = d.scan_tmp[blockIdx.x];
  __syncthreads();
  for (int r = 0; r < RK_I; ++r) {
    int64_t i = t0 + r * RK_T + threadIdx.x;
    bool f = i < d.n && d.flag[i];
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
  long long noise = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); i0 < d.n; i0 += stride) {
    const int64_t i = i0 + lane_id();
    if (i < d.n) {
      const int32_t o = d.orig[i];
      const int32_t k = d.keyv[i];
      const Inst& I = d.inst[d.cell_inst[d.cell_of[i]]];
      const int32_t lab = k == kNoKey ? 0 : d.rank[k] - d.rank[I.pt_begin] + 1;
      d.labels_out[o] = lab;
      d.core_out[o] = d.core[i];
      noise += lab == 0;
    }
  }
  wave_add(&d.ctl[C_NNOISE], noise);
}

// ------------------------------------------------------------- launchers --
void launch_core(const Dev& d, bool f32, int grid, hipStream_t s) {
  if (f32)
    hipLaunchKernelGGL(k_core<true>, dim3(grid), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(k_core<false>, dim3(grid), dim3(256), 0, s, d);
}

void launch_pass_prep(const Dev& d, int pass, hipStream_t s) {
  hipLaunchKernelGGL(k_pass_prep, dim3(1), dim3(1), 0, s, d, pass);
}

void launch_pass_carry(const Dev& d, int pass, hipStream_t s) {
  hipLaunchKernelGGL(k_pass_carry, dim3(1), dim3(256), 0, s, d, pass);
}

void launch_seg_init(const Dev& d, int64_t waves, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_init, dim3(64), dim3(256), 0, s, d, waves);
}

void launch_expand(const Dev& d, bool f32, int pass, int grid, int seeds, int threads,
                   int budget, hipStream_t s) {
  size_t lds = (size_t)(threads / 64) * seeds * sizeof(int32_t);
  if (f32)
    hipLaunchKernelGGL(k_expand<true>, dim3(grid), dim3(threads), lds, s, d, pass, seeds, budget);
  else
    hipLaunchKernelGGL(k_expand<false>, dim3(grid), dim3(threads), lds, s, d, pass, seeds, budget);
}

int expand_blocks_per_cu(bool f32, int seeds, int threads) {
  int nb = 0;
  size_t lds = (size_t)(threads / 64) * seeds * sizeof(int32_t);
  hipError_t e = f32 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_expand<true>, threads, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_expand<false>, threads, lds);
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

void launch_merge(const Dev& d, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_merge_lds, dim3(1), dim3(1024), 0, s, d);
  hipLaunchKernelGGL(k_union_pairs, dim3(grid), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_flatten, dim3(grid), dim3(256), 0, s, d);
}

void launch_border(const Dev& d, bool f32, int grid, int64_t n, hipStream_t s) {
  int g2 = (int)((n + 255) / 256);
  if (g2 > 4096) g2 = 4096;
  if (g2 < 1) g2 = 1;
  hipLaunchKernelGGL(k_compmin, dim3(g2), dim3(256), 0, s, d);
  hipLaunchKernelGGL(k_keys, dim3(g2), dim3(256), 0, s, d);
  if (f32)
    hipLaunchKernelGGL(k_border<true>, dim3(grid), dim3(256), 0, s, d);
  else
    hipLaunchKernelGGL(k_border<false>, dim3(grid), dim3(256), 0, s, d);
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
