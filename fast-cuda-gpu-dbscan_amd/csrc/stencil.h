// Device helpers shared by the clustering kernels and the one-step
// baseline: coordinate loads, the FP64 neighbour predicate, and the
// candidate streams over a cell's stencil lines.
#pragma once
#include "internal.h"

namespace dclust {

// ------------------------------------------------------------ predicates --
// FP64 coordinates of cell-sorted point i.  The FP32 form keeps no sorted
// FP64 copy (only its guard-band rechecks need FP64): they are read from
// the caller's rows at the point's row offset.
template <int D, bool F32>
__device__ __forceinline__ void load_x(const Dev& d, int32_t i, double* x) {
  if constexpr (F32) {
    const double* p = d.xyz + d.xoff[i];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = p[k];
  } else {
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = d.X[k][i];
  }
}
template <int D>
__device__ __forceinline__ void load_g(const Dev& d, int32_t i, uint32_t* g) {
#pragma unroll
  for (int k = 0; k < D; ++k) g[k] = d.G[k][i];
}

// fl(fl(fl(d0^2) + fl(d1^2)) [+ fl(d2^2)]), left to right
template <int D>
__device__ __forceinline__ double d2_64(const double* a, const double* b) {
  double t = a[0] - b[0];
  double s = t * t;
#pragma unroll
  for (int k = 1; k < D; ++k) {
    double u = a[k] - b[k];
    double v = u * u;
    s = s + v;
  }
  return s;
}

// the oracle's box of a range query centred on c: fl(c-eps) <= q <= fl(c+eps)
template <int D>
__device__ __forceinline__ bool box_64(const double* c, const double* q, double eps) {
  bool in = true;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double lo = c[k] - eps, hi = c[k] + eps;
    in = in && lo <= q[k] && q[k] <= hi;
  }
  return in;
}

// ------------------------------------------------------ candidate streams --
// Up to five index ranges [a_j, a_j + l_j) read as one stream: position t
// lies in range j = #{k >= 1 : P_k <= t}, P_k = l_0 + ... + l_{k-1}.
struct Stream5 {
  int32_t a[5];
  int32_t P[5];
  int32_t tot;
  __device__ __forceinline__ void set(const int32_t* aa, const int32_t* ll) {
    int32_t s = 0;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      a[j] = aa[j];
      P[j] = s;
      s += ll[j];
    }
    tot = s;
  }
  __device__ __forceinline__ int32_t at(int32_t t, int& j) const {
    int32_t q = a[0] + t;
    j = 0;
#pragma unroll
    for (int k = 1; k < 5; ++k)
      if (t >= P[k]) {
        q = a[k] + (t - P[k]);
        j = k;
      }
    return q;
  }
};

// point ranges of five stencil lines given their packed rows
__device__ __forceinline__ void line_ranges(const Dev& d, const int32_t* pk, int32_t* a,
                                            int32_t* l) {
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const int32_t len = pk[j] >> kRowCntShift, st = pk[j] & kRowCellMask;
    int32_t x = 0, y = 0;
    if (len) {
      x = d.cell_start[st];
      y = d.cell_start[st + len];
    }
    a[j] = x;
    l[j] = y - x;
  }
}

// point ranges of five stencil lines
__device__ __forceinline__ void line_group(const Dev& d, const RowsOf& rows, int g, int32_t* a,
                                           int32_t* l) {
  int32_t pk[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) pk[j] = rows[g + j];
  line_ranges(d, pk, a, l);
}

}  // namespace dclust
