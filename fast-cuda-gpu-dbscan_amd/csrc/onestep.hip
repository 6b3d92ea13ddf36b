// One-step expansion baseline: CUDA-DClust (Böhm et al., 2009) as the
// reference ships it in cuda-dclust/dbscan_indexing.cu, on this library's
// grid index and FP64 predicate.  It exists to time the design that
// CUDA-DClust+ improved on, on the same GPU and inputs; it is not the
// clustering path.
//
// What is kept from cuda-dclust/dbscan_indexing.cu:
//   - one workgroup per chain, one seed popped (LIFO) per chain per launch,
//     its neighbours counted by the whole workgroup (DBSCAN kernel,
//     :529-637, seed pop :559);
//   - a host round trip after every launch that reads the seed-list lengths,
//     and new seeds only when every list is empty (host loop :442-472,
//     MonitorSeedPoints :639-672);
//   - refill seeds taken from the unprocessed points from the back of the
//     input order (:154-157, :738-753);
//   - the defaults THREAD_BLOCKS 256, THREAD_COUNT 256, MAX_SEEDS 128
//     (:28-31).
// What differs, so that the result is an exact DBSCAN clustering:
//   - a push that finds its seed list full leaves the point unclaimed for a
//     later refill instead of dropping it (MarkAsCandidate, :783-795, and
//     the clamp at :632-634);
//   - a contact between two chains is recorded only between two expanded
//     core points (the reference records one for any neighbour), in a list
//     instead of a blocks x blocks matrix, so chain ids stay unique across
//     refills and the union runs once at the end (MonitorSeedPoints' BFS
//     over the matrix and thrust::replace per chain, :681-728);
//   - a popped seed that is not core stays a border point of the chain that
//     claimed it (the reference sets it to noise, :619-627); a noise point
//     that a core point reaches later becomes a border point of its chain,
//     and a final pass gives every remaining noise point the chain of a core
//     point within eps, if any.
//
// Point state (cell-sorted order): label -1 unclassified, -2 noise, >= 0
// chain; state 0 not expanded, 1 expanded and not core, 2 expanded core.
// A core point publishes state 2 before it reads any neighbour's state and
// both are agent-scope atomics, so of two adjacent core points expanded in
// the same launch at least one sees the other and records the contact.
#include "internal.h"
#include "kernels.h"
#include "stencil.h"

namespace dclust {

constexpr int kOsDedup = 32;  // other chains remembered per expansion

static int os_grid(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

__device__ __forceinline__ int32_t ld_agent(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(int32_t* p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void k_os_init(OneStep o, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    o.label[i] = -1;
    o.state[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < 2) o.ctr[threadIdx.x] = 0;
}

// first[b] >= 0: the refill seed of chain b, else the chain stays empty
__global__ __launch_bounds__(256) void k_os_refill(OneStep o, const int32_t* first) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= o.blocks) return;
  const int32_t p = first[b];
  o.seeds[(int64_t)b * o.S] = p;
  o.len[b] = p >= 0 ? 1 : 0;
}

template <int D>
__device__ void os_step(const Dev& d, const OneStep& o, int32_t chain) {
  __shared__ int32_t s_p, s_own, s_len, s_cnt, s_nc;
  __shared__ int32_t s_other[kOsDedup];
  const int b = blockIdx.x, tid = threadIdx.x, T = blockDim.x;
  int32_t* const seeds = o.seeds + (int64_t)b * o.S;
  if (tid == 0) {
    int32_t len = o.len[b];
    int32_t p = -1, own = 0;
    while (len > 0 && p < 0) {
      const int32_t q = seeds[--len];
      if (q < 0) continue;  // a slot whose claim another chain won
      const int32_t l = ld_agent(o.label + q);
      if (l == chain) {
        p = q;  // pushed by this chain
      } else if (l == -1 && atomicCAS(o.label + q, -1, chain) == -1) {
        p = q;  // the chain's refill seed
        own = 1;
      }  // else: another chain took the refill seed first
    }
    s_p = p;
    s_own = own;
    s_len = len;
    s_cnt = 0;
    s_nc = 0;
  }
  if (tid < kOsDedup) s_other[tid] = -1;  // chain ids are >= 0
  __syncthreads();
  const int32_t p = s_p;
  if (p < 0) {
    if (tid == 0) o.len[b] = 0;
    return;
  }
  const Inst& I = *d.inst;
  const double eps = I.eps, eps2 = I.eps2;
  constexpr int NR = D == 3 ? 25 : 5;
  double px[D];
  load_x<D, false>(d, p, px);
  const RowsOf rows = rows_of(d, d.cell_of[p]);

  // neighbour count of p, self included (dbscan_indexing.cu:581-617)
  int cnt = 0;
  for (int g = 0; g < NR; g += 5) {
    int32_t aa[5], ll[5];
    line_group(d, rows, g, aa, ll);
    Stream5 st;
    st.set(aa, ll);
    for (int32_t i = tid; i < st.tot; i += T) {
      int jr;
      const int32_t q = st.at(i, jr);
      double qx[D];
      load_x<D, false>(d, q, qx);
      cnt += box_64<D>(px, qx, eps) && d2_64<D>(px, qx) <= eps2;
    }
  }
  cnt = wave_sum(cnt);
  if (lane_id() == 0) atomicAdd(&s_cnt, cnt);
  __syncthreads();
  const bool core = s_cnt >= I.minpts;
  if (tid == 0) {
    __threadfence();  // the claim of the refill seed before its state
    st_agent(o.state + p, core ? 2 : 1);
    if (!core && s_own) st_agent(o.label + p, -2);  // a refill seed that is not core
    __threadfence();  // the state before any neighbour's state is read
  }
  __syncthreads();
  if (!core) {
    if (tid == 0) o.len[b] = s_len;
    return;
  }

  // claim and push the unclassified neighbours, take the noise ones as
  // border points, record contacts with expanded core points of other
  // chains (MarkAsCandidate, dbscan_indexing.cu:783-807)
  for (int g = 0; g < NR; g += 5) {
    int32_t aa[5], ll[5];
    line_group(d, rows, g, aa, ll);
    Stream5 st;
    st.set(aa, ll);
    for (int32_t i = tid; i < st.tot; i += T) {
      int jr;
      const int32_t q = st.at(i, jr);
      if (q == p) continue;
      double qx[D];
      load_x<D, false>(d, q, qx);
      if (!(box_64<D>(px, qx, eps) && d2_64<D>(px, qx) <= eps2)) continue;
      const int32_t lq = ld_agent(o.label + q);
      if (lq == -1) {
        // a slot first, then the claim: a full list leaves q unclaimed
        const int32_t slot = atomicAdd(&s_len, 1);
        if (slot < o.S) {
          const bool won = atomicCAS(o.label + q, -1, chain) == -1;
          seeds[slot] = won ? q : -1;
        }
      } else if (lq == -2) {
        atomicCAS(o.label + q, -2, chain);
      } else if (lq != chain && ld_agent(o.state + q) == 2) {
        bool seen = false;
        const int m = min(*(volatile int32_t*)&s_nc, kOsDedup);
        for (int k = 0; k < m && !seen; ++k) seen = ((volatile int32_t*)s_other)[k] == lq;
        if (!seen) {
          const int k = atomicAdd(&s_nc, 1);
          if (k < kOsDedup) s_other[k] = lq;
          const int32_t at = atomicAdd(o.ctr, 1);
          if (at < o.cap) o.contacts[at] = make_int2(chain, lq);
        }
      }
    }
  }
  __syncthreads();
  if (tid == 0) o.len[b] = min(s_len, o.S);
}

__global__ __launch_bounds__(1024) void k_os_step(Dev d, OneStep o, int32_t chain0) {
  if (d.inst->dim == 3)
    os_step<3>(d, o, chain0 + (int32_t)blockIdx.x);
  else
    os_step<2>(d, o, chain0 + (int32_t)blockIdx.x);
}

// every noise point takes the chain of the first core point (in stencil
// order) whose expansion would reach it; one lane per point
template <int D>
__device__ void os_border(const Dev& d, const OneStep& o, int32_t p) {
  const Inst& I = *d.inst;
  constexpr int NR = D == 3 ? 25 : 5;
  double px[D];
  load_x<D, false>(d, p, px);
  const RowsOf rows = rows_of(d, d.cell_of[p]);
  for (int g = 0; g < NR; g += 5) {
    int32_t aa[5], ll[5];
    line_group(d, rows, g, aa, ll);
    Stream5 st;
    st.set(aa, ll);
    for (int32_t i = 0; i < st.tot; ++i) {
      int jr;
      const int32_t q = st.at(i, jr);
      if (o.state[q] != 2) continue;
      double qx[D];
      load_x<D, false>(d, q, qx);
      if (box_64<D>(qx, px, I.eps) && d2_64<D>(qx, px) <= I.eps2) {
        o.label[p] = o.label[q];
        return;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_os_border(Dev d, OneStep o) {
  const bool three = d.inst->dim == 3;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < d.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (o.label[i] != -2) continue;
    if (three)
      os_border<3>(d, o, (int32_t)i);
    else
      os_border<2>(d, o, (int32_t)i);
  }
}

void launch_os_init(const OneStep& o, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_os_init, dim3(os_grid(n)), dim3(256), 0, s, o, n);
}
void launch_os_refill(const OneStep& o, const int32_t* first, hipStream_t s) {
  hipLaunchKernelGGL(k_os_refill, dim3((o.blocks + 255) / 256), dim3(256), 0, s, o, first);
}
void launch_os_step(const Dev& d, const OneStep& o, int32_t chain0, int threads,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_os_step, dim3(o.blocks), dim3(threads), 0, s, d, o, chain0);
}
void launch_os_border(const Dev& d, const OneStep& o, hipStream_t s) {
  hipLaunchKernelGGL(k_os_border, dim3(os_grid(d.n)), dim3(256), 0, s, d, o);
}

}  // namespace dclust
