// Clustering kernels on gfx950: the MI355X replacement for the reference's
// expansion kernel and host merge loop.
//
//   k_core     neighbour counts with early exit at minpts -> core flags.
//              (dbscan.cu:331-364 counts neighbours while expanding; here the
//              count is a separate cell-batched pass so that every later
//              kernel knows which candidates are core.)
//   k_expand   persistent seed-list expansion (dbscan.cu:254-387): each
//              wavefront is a chain with its own seed list in LDS, pops a
//              batch of seeds of one cell, streams the 3^(D-1) neighbour
//              rows, claims unclaimed core neighbours by CAS
//              (MarkAsCandidate, dbscan.cu:189-222), pushes them with
//              ballot/mbcnt, and records chain-chain collisions as pairs.
//              Seed-list overflow goes to a correction queue expanded by the
//              next pass instead of being dropped (dbscan.cu:196-199).
//   k_union_pairs / k_flatten   GPU atomic union-find over chain ids
//              (MonitorSeedPoints' CPU BFS over the collision matrix and
//              thrust::replace relabels, dbscan.cu:60-125).
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

// ------------------------------------------------------- cell streaming --
// The candidates of a cell are the concatenation of its NR neighbour rows.
// pref[r] = first stream index of row r; delta[r] telescopes the row start
// offsets so that point = m + sum_{r: pref[r] <= m} delta[r].
template <int NR>
struct Stream {
  int32_t M;
  int32_t pref[NR];
  int32_t delta[NR];
};

template <int NR>
__device__ __forceinline__ void load_stream(const Dev& d, int32_t c, Stream<NR>& s) {
  const int lane = lane_id();
  int2 r = lane < NR ? d.rows[(int64_t)c * kMaxRows + lane] : make_int2(0, 0);
  int32_t len = r.y - r.x;
  int32_t incl = wave_incl_scan(len);
  int32_t excl = incl - len;
  int32_t off = r.x - excl;
  s.M = rl_i(incl, NR - 1);
  int32_t prev = 0;
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    s.pref[k] = rl_i(excl, k);
    int32_t o = rl_i(off, k);
    s.delta[k] = o - prev;
    prev = o;
  }
}

template <int NR>
__device__ __forceinline__ int32_t stream_point(const Stream<NR>& s, int32_t m) {
  int32_t pt = m;
#pragma unroll
  for (int k = 0; k < NR; ++k) pt += (m >= s.pref[k]) ? s.delta[k] : 0;
  return pt;
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

// FP32 squared distance from fixed-point cell offsets (units of w*2^-30).
// Differences are exact in wrapped int32 arithmetic because the two points
// are less than two cells apart on every axis.
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
  long long band = 0, flip = 0, asym = 0, eval = 0;
};

// Query of the current wave-uniform iteration, read from lane qi.
template <int D, bool F32>
struct Query {
  double x[D], lo[D], hi[D];
  uint32_t g[D];
};

template <int D, bool F32>
__device__ __forceinline__ void load_query_lane(const Dev& d, const Inst& I, int32_t pt,
                                                bool active, Query<D, F32>& q) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double x = active ? d.X[k][pt] : 0.0;
    q.x[k] = x;
    q.lo[k] = x - I.eps;
    q.hi[k] = x + I.eps;
    if (F32) q.g[k] = active ? d.G[k][pt] : 0u;
  }
}

template <int D, bool F32>
__device__ __forceinline__ void read_query(const Query<D, F32>& mine, int qi, Query<D, F32>& u) {
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (F32) {
      u.g[k] = rl_u(mine.g[k], qi);
    } else {
      u.x[k] = rl_d(mine.x[k], qi);
      u.lo[k] = rl_d(mine.lo[k], qi);
      u.hi[k] = rl_d(mine.hi[k], qi);
    }
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

// P(query -> candidate): the query is the centre (its expansion / count).
// The FP32 form decides outside the guard band and re-decides inside it
// with the exact FP64 predicate (FP64 coordinates fetched only then).
template <int D, bool F32>
__device__ __forceinline__ bool pred_qc(const Dev& d, const Inst& I, const Query<D, F32>& mineq,
                                        int qi, const Query<D, F32>& uq, const Cand<D, F32>& c,
                                        int32_t pt, bool active, Tally& t) {
  if (!F32) return active && pred64<D>(uq.x, uq.lo, uq.hi, c.x, I.eps2);
  float s = d2fix<D>(uq.g, c.g);
  bool in = s < I.d2_in;
  bool amb = active && !in && s <= I.d2_out;
  if (ballot(amb)) {
    double px[D], plo[D], phi[D], qx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      px[k] = rl_d(mineq.x[k], qi);
      plo[k] = rl_d(mineq.lo[k], qi);
      phi[k] = rl_d(mineq.hi[k], qi);
    }
    if (amb) {
#pragma unroll
      for (int k = 0; k < D; ++k) qx[k] = d.X[k][pt];
      bool ex = pred64<D>(px, plo, phi, qx, I.eps2);
      t.band++;
      t.flip += (ex != (s <= I.d2_mid));
      in = ex;
    }
  }
  return active && in;
}

// ------------------------------------------------------------------ core --
template <int D, bool F32>
__device__ void core_cell(const Dev& d, const Inst& I, int32_t c, Tally& t, long long& ncore) {
  constexpr int NR = D == 3 ? 9 : 3;
  const int lane = lane_id();
  Stream<NR> st;
  load_stream<NR>(d, c, st);
  const int32_t qs = d.cell_start[c], qe = d.cell_start[c + 1];
  for (int32_t q0 = qs; q0 < qe; q0 += 64) {
    const int nq = (qe - q0) < 64 ? (qe - q0) : 64;
    Query<D, F32> mq;
    load_query_lane<D, F32>(d, I, q0 + lane, lane < nq, mq);
    int cnt = 0;
    unsigned long long done = nq == 64 ? 0ull : (~0ull << nq);
    for (int32_t m0 = 0; m0 < st.M && done != ~0ull; m0 += 64) {
      const int32_t m = m0 + lane;
      const bool valid = m < st.M;
      const int32_t pt = valid ? stream_point<NR>(st, m) : 0;
      Cand<D, F32> cd;
      load_cand<D, F32>(d, pt, valid, cd);
      unsigned long long todo = ~done;
      t.eval += (long long)__popcll(todo) * __popcll(ballot(valid));
      while (todo) {
        const int qi = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        Query<D, F32> uq;
        read_query<D, F32>(mq, qi, uq);
        bool hit = pred_qc<D, F32>(d, I, mq, qi, uq, cd, pt, valid, t);
        int pc = __popcll(ballot(hit));
        cnt += (lane == qi) ? pc : 0;
      }
      done |= ballot(cnt >= I.minpts);
    }
    if (lane < nq) d.core[q0 + lane] = cnt >= I.minpts;
    ncore += __popcll(ballot(lane < nq && cnt >= I.minpts));
  }
}

template <bool F32>
__global__ __launch_bounds__(256) void k_core(Dev d) {
  const int32_t C = (int32_t)d.ctl[C_NCELL];
  const int nwb = blockDim.x >> 6;
  const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * nwb + (threadIdx.x >> 6);
  const int64_t W = (int64_t)gridDim.x * nwb;
  constexpr int CH = 16;  // cells per wave step
  Tally t;
  long long ncore = 0;
  for (int64_t c0 = gw * CH; c0 < C; c0 += W * CH) {
    const int32_t ce = (int32_t)((c0 + CH) < C ? (c0 + CH) : C);
    for (int32_t c = (int32_t)c0; c < ce; ++c) {
      const Inst& I = d.inst[d.cell_inst[c]];
      if (I.dim == 3)
        core_cell<3, F32>(d, I, c, t, ncore);
      else
        core_cell<2, F32>(d, I, c, t, ncore);
    }
  }
  if (lane_id() == 0) {
    if (ncore) atomicAdd((unsigned long long*)&d.ctl[C_NCORE], (unsigned long long)ncore);
    if (t.eval) atomicAdd((unsigned long long*)&d.ctl[C_EVAL], (unsigned long long)t.eval);
  }
  wave_add(&d.ctl[C_BAND], t.band);
  wave_add(&d.ctl[C_FLIP], t.flip);
}

// ---------------------------------------------------------------- expand --
__device__ __forceinline__ void record_pair(const Dev& d, int32_t a, int32_t b) {
  long long i = (long long)atomicAdd((unsigned long long*)&d.ctl[C_NPAIR], 1ull);
  if (i < d.pair_cap)
    d.pairs[i] = make_int2(a, b);
  else
    uf_union(d.parent, a, b);  // buffer full: merge inline (same result)
}

struct SeedList {
  int32_t* list;
  int S;
  int sl;
  int32_t* qout;
  long long* qout_n;
  int32_t cache;  // lanes 0..7: recently recorded foreign chains
  int cache_pos;
};

template <int D, bool F32>
__device__ void expand_cell(const Dev& d, const Inst& I, int32_t c, int32_t own, int bs,
                            int32_t qpt, SeedList& L, Tally& t, long long& ovf) {
  constexpr int NR = D == 3 ? 9 : 3;
  const int lane = lane_id();
  Stream<NR> st;
  load_stream<NR>(d, c, st);
  Query<D, F32> mq;
  load_query_lane<D, F32>(d, I, qpt, lane < bs, mq);
  for (int32_t m0 = 0; m0 < st.M; m0 += 64) {
    const int32_t m = m0 + lane;
    const bool valid = m < st.M;
    const int32_t pt = valid ? stream_point<NR>(st, m) : 0;
    int32_t lab = valid ? d.label[pt] : own;
    const bool rel = valid && lab != own && d.core[pt];
    if (!ballot(rel)) continue;
    Cand<D, F32> cd;
    load_cand<D, F32>(d, pt, rel, cd);
    bool hit = false;
    int hq = 0;
    for (int qi = 0; qi < bs; ++qi) {
      Query<D, F32> uq;
      read_query<D, F32>(mq, qi, uq);
      bool h = pred_qc<D, F32>(d, I, mq, qi, uq, cd, pt, rel && !hit, t);
      if (h) {
        hit = true;
        hq = qi;
      }
      if (!ballot(rel && !hit)) break;
    }
    // box-test symmetry of the core-core pair just found (FP64 coordinates)
    if (ballot(hit)) {
      double px[D], cx[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        px[k] = __shfl(mq.x[k], hq, 64);
        cx[k] = hit ? d.X[k][pt] : 0.0;
      }
      if (hit) {
        bool rev = true;
#pragma unroll
        for (int k = 0; k < D; ++k) rev = rev && !((cx[k] - I.eps) > px[k] || px[k] > (cx[k] + I.eps));
        bool fwd = true;
#pragma unroll
        for (int k = 0; k < D; ++k) fwd = fwd && !((px[k] - I.eps) > cx[k] || cx[k] > (px[k] + I.eps));
        t.asym += (rev != fwd);
      }
    }
    bool claimed = false;
    if (hit && lab == kUnclaimed) {
      int32_t old = atomicCAS(&d.label[pt], kUnclaimed, own);
      claimed = old == kUnclaimed;
      lab = claimed ? own : old;
    }
    bool coll = hit && !claimed && lab != own;
    unsigned long long cm = ballot(coll);
    while (cm) {
      const int l = __ffsll((long long)cm) - 1;
      const int32_t v = rl_i(lab, l);
      const bool seen = ballot(lane < 8 && L.cache == v) != 0;
      if (!seen) {
        if (lane == 0) record_pair(d, own, v);
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
          L.list[pos] = pt;
        else
          L.qout[base + pos - L.S] = pt;
      }
      L.sl = (L.sl + cnt) < L.S ? (L.sl + cnt) : L.S;
    }
  }
  wave_sync();
}

// Persistent expansion pass.  Pass 0: wave w scans its own contiguous
// segment of the cell-sorted points for unclaimed core points and starts a
// new chain at each (device-side seeding; the reference pops host-side
// seeds, dbscan.cu:142-160).  Pass p > 0: waves drain the correction queue
// filled by pass p-1 (points claimed into a full seed list).  A wave exits
// when its list is empty and no seed source is left, so every wave reaches
// the end of the kernel.
template <bool F32>
__global__ __launch_bounds__(256) void k_expand(Dev d, int pass, int S) {
  extern __shared__ int32_t s_lists[];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const int nwb = blockDim.x >> 6;
  const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * nwb + wv;
  const int64_t W = (int64_t)gridDim.x * nwb;
  SeedList L;
  L.list = s_lists + wv * S;
  L.S = S;
  L.sl = 0;
  L.qout = d.q[(pass + 1) & 1];
  L.qout_n = &d.ctl[C_Q0_N + ((pass + 1) & 1)];
  L.cache = -1;
  L.cache_pos = 0;
  const int32_t* qin = d.q[pass & 1];
  long long* qin_head = &d.ctl[C_Q0_HEAD + (pass & 1)];
  const long long qn = pass > 0 ? d.ctl[C_Q0_N + (pass & 1)] : 0;
  bool qlive = pass > 0;
  int64_t seg = 0, seg_end = 0;
  if (pass == 0) {
    seg = d.n * gw / W;
    seg_end = d.n * (gw + 1) / W;
  }
  int32_t cache_own = -2;
  Tally t;
  long long ovf = 0;
  for (;;) {
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
        const bool cand = i < seg_end && d.core[i] && d.label[i] == kUnclaimed;
        unsigned long long m = ballot(cand);
        while (m) {
          const int l = __ffsll((long long)m) - 1;
          int ok = 0;
          if (lane == 0) {
            int32_t id = (int32_t)atomicAdd((unsigned long long*)&d.ctl[C_NCHAIN], 1ull);
            ok = atomicCAS(&d.label[seg + l], kUnclaimed, id) == kUnclaimed;
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
    // pop the top seed and the run of seeds below it from the same cell
    // and chain: they share one neighbourhood stream
    const int top = L.sl - 1;
    const int32_t e = lane <= top ? L.list[top - lane] : -1;
    const int32_t ce = e >= 0 ? d.cell_of[e] : -1;
    const int32_t le = e >= 0 ? d.label[e] : -3;
    const int32_t c = rfl_i(ce);
    const int32_t own = rfl_i(le);
    const unsigned long long okm = ballot(e >= 0 && ce == c && le == own);
    const int bs = (~okm == 0ull) ? 64 : (__ffsll((long long)~okm) - 1);
    L.sl -= bs;
    wave_sync();
    if (own < 0 || c < 0) continue;  // cannot happen: list entries are claimed points
    if (own != cache_own) {
      L.cache = -1;
      L.cache_pos = 0;
      cache_own = own;
    }
    const Inst& I = d.inst[d.cell_inst[c]];
    if (I.dim == 3)
      expand_cell<3, F32>(d, I, c, own, bs, e, L, t, ovf);
    else
      expand_cell<2, F32>(d, I, c, own, bs, e, L, t, ovf);
  }
  if (lane_id() == 0 && ovf) atomicAdd((unsigned long long*)&d.ctl[C_OVF], (unsigned long long)ovf);
  wave_add(&d.ctl[C_BAND], t.band);
  wave_add(&d.ctl[C_FLIP], t.flip);
  wave_add(&d.ctl[C_ASYM], t.asym);
}

// Resets the counters of the queue that pass `pass` fills and of the head
// of the queue it drains.
__global__ void k_pass_prep(Dev d, int pass) {
  d.ctl[C_Q0_N + ((pass + 1) & 1)] = 0;
  d.ctl[C_Q0_HEAD + (pass & 1)] = 0;
}

// ----------------------------------------------------------------- merge --
__global__ __launch_bounds__(256) void k_union_pairs(Dev d) {
  long long np = d.ctl[C_NPAIR];
  if (np > d.pair_cap) np = d.pair_cap;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < np;
       i += (long long)gridDim.x * blockDim.x) {
    int2 p = d.pairs[i];
    uf_union(d.parent, p.x, p.y);
  }
}

__global__ __launch_bounds__(256) void k_flatten(Dev d) {
  const long long nc = d.ctl[C_NCHAIN];
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < nc;
       c += (long long)gridDim.x * blockDim.x)
    d.parent[c] = uf_find(d.parent, (int32_t)c);
}

// minimum input index per component (wave-aggregated per distinct root)
__global__ __launch_bounds__(256) void k_compmin(Dev d) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63); i0 < d.n; i0 += stride) {
    const int64_t i = i0 + lane_id();
    const int32_t li = (i < d.n && d.core[i]) ? d.label[i] : -1;
    bool cr = li >= 0;  // every core point is claimed; the guard only keeps a bug in bounds
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
    const int32_t li = d.core[i] ? d.label[i] : -1;
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
// the core candidates q with P(q -> query), evaluated from the core
// point's side as its expansion in the reference oracle would.
template <int D, bool F32>
__device__ void border_cell(const Dev& d, const Inst& I, int32_t c, Tally& t) {
  constexpr int NR = D == 3 ? 9 : 3;
  const int lane = lane_id();
  const int32_t qs = d.cell_start[c], qe = d.cell_start[c + 1];
  bool any = false;
  for (int32_t q0 = qs; q0 < qe; q0 += 64)
    any |= ballot(q0 + lane < qe && !d.core[q0 + lane]) != 0;
  if (!any) return;
  Stream<NR> st;
  load_stream<NR>(d, c, st);
  for (int32_t q0 = qs; q0 < qe; q0 += 64) {
    const int nq = (qe - q0) < 64 ? (qe - q0) : 64;
    const bool isq = lane < nq && !d.core[q0 + lane];
    unsigned long long qm = ballot(isq);
    if (!qm) continue;
    double qx[D];
    uint32_t qg[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      qx[k] = isq ? d.X[k][q0 + lane] : 0.0;
      qg[k] = (F32 && isq) ? d.G[k][q0 + lane] : 0u;
    }
    int32_t best = kNoKey;
    for (int32_t m0 = 0; m0 < st.M; m0 += 64) {
      const int32_t m = m0 + lane;
      const bool valid = m < st.M;
      const int32_t pt = valid ? stream_point<NR>(st, m) : 0;
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
            for (int k = 0; k < D; ++k) ux[k] = rl_d(qx[k], qi);
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
  const int32_t C = (int32_t)d.ctl[C_NCELL];
  const int nwb = blockDim.x >> 6;
  const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * nwb + (threadIdx.x >> 6);
  const int64_t W = (int64_t)gridDim.x * nwb;
  constexpr int CH = 16;
  Tally t;
  for (int64_t c0 = gw * CH; c0 < C; c0 += W * CH) {
    const int32_t ce = (int32_t)((c0 + CH) < C ? (c0 + CH) : C);
    for (int32_t c = (int32_t)c0; c < ce; ++c) {
      const Inst& I = d.inst[d.cell_inst[c]];
      if (I.dim == 3)
        border_cell<3, F32>(d, I, c, t);
      else
        border_cell<2, F32>(d, I, c, t);
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

void launch_expand(const Dev& d, bool f32, int pass, int grid, int seeds, int threads,
                   hipStream_t s) {
  size_t lds = (size_t)(threads / 64) * seeds * sizeof(int32_t);
  if (f32)
    hipLaunchKernelGGL(k_expand<true>, dim3(grid), dim3(threads), lds, s, d, pass, seeds);
  else
    hipLaunchKernelGGL(k_expand<false>, dim3(grid), dim3(threads), lds, s, d, pass, seeds);
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
