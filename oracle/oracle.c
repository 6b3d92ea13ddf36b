/*
 * CPU oracle for the DBSCAN hot path -- TEST INFRASTRUCTURE ONLY.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg, as the checker or the timed CPU baseline.  The product library never
 * links it.
 *
 * Build: must be compiled with -ffp-contract=off (see oracle/Makefile) so
 * that fl(fl(dx*dx) + fl(dy*dy)) is evaluated exactly as the reference's
 * g++/x86-64 build does (no FMA contraction).
 *
 * Pinning (tests/test_oracle.py):
 *  - oracle_seq_dbscan vs the compiled reference oracle (oracle/_ref, built
 *    from cpu-dbscan/dbscan_rtree_exp.cpp in place) on 2-D fixtures, and vs
 *    the committed golden labels in tests/golden/;
 *  - oracle_fast_dbscan vs oracle_seq_dbscan on 2-D and 3-D inputs;
 *  - core flags vs oracle_brute_counts (O(N^2), independent enumeration);
 *  - oracle_pred vs hand-built boundary pairs (closed box, exact eps^2 ties).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Predicate.  cpu-dbscan/dbscan_rtree_exp.cpp:351-358 (distance),          */
/* :420-438 (box search then d^2 <= eps*eps), Rtree.h:1573-1587 (Overlap).  */
/* ------------------------------------------------------------------------ */
static inline int pred_inline(const double* p, const double* q, int dim,
                              double eps, double eps2) {
  for (int k = 0; k < dim; ++k) {
    double lo = p[k] - eps;
    double hi = p[k] + eps;
    if (lo > q[k] || q[k] > hi) return 0;
  }
  double dx = p[0] - q[0];
  double d = dx * dx;
  for (int k = 1; k < dim; ++k) {
    double t = p[k] - q[k];
    double s = t * t;
    d = d + s;
  }
  return d <= eps2;
}

int oracle_pred(const double* p, const double* q, int dim, double eps) {
  return pred_inline(p, q, dim, eps, eps * eps);
}

/* ------------------------------------------------------------------------ */
/* A plain uniform grid used only to enumerate a superset of the closed     */
/* search box (the role Rtree::Search plays in the reference).  The cell    */
/* function c(x) = floor((x-min)*inv) is monotone in x, so every q inside   */
/* [fl(p-eps), fl(p+eps)] lies in cells [c(fl(p-eps)), c(fl(p+eps))].       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int dim;
  double mn[3], inv[3];
  int64_t nk[3];
  int64_t ncell;
  int64_t* key;   /* sorted unique cell keys            */
  int64_t* start; /* CSR offsets into idx, ncell+1      */
  int32_t* idx;   /* point ids sorted by (cell, id)     */
} grid_t;

static inline int64_t gcell(const grid_t* g, int k, double x) {
  double t = floor((x - g->mn[k]) * g->inv[k]);
  if (!(t >= 0.0)) return 0;
  if (t >= (double)(g->nk[k] - 1)) return g->nk[k] - 1;
  return (int64_t)t;
}

static int64_t gkey(const grid_t* g, const int64_t* c) {
  int64_t key = c[0];
  for (int k = 1; k < g->dim; ++k) key = key * g->nk[k] + c[k];
  return key;
}

typedef struct { int64_t key; int32_t id; } kv_t;
static int kv_cmp(const void* a, const void* b) {
  const kv_t* x = (const kv_t*)a;
  const kv_t* y = (const kv_t*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return (x->id > y->id) - (x->id < y->id);
}

/* Order (key, id) pairs by key, then id.  The pairs arrive in id order, so
 * a stable LSD radix sort on the key (11-bit digits, as many as the key
 * space needs) gives that order; qsort with kv_cmp when memory is short. */
static int kv_sort(kv_t* kv, int64_t n, const grid_t* g) {
  int64_t cells = 1;
  for (int k = 0; k < g->dim; ++k) cells *= g->nk[k];
  int bits = 1;
  while (bits < 63 && ((int64_t)1 << bits) < cells) ++bits;
  kv_t* tmp = (kv_t*)malloc(sizeof(kv_t) * (size_t)(n > 0 ? n : 1));
  if (!tmp) {
    qsort(kv, (size_t)n, sizeof(kv_t), kv_cmp);
    return 0;
  }
  enum { RB = 11, RN = 1 << RB };
  int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * RN);
  if (!cnt) { free(tmp); return -1; }
  kv_t* src = kv;
  kv_t* dst = tmp;
  for (int sh = 0; sh < bits; sh += RB) {
    memset(cnt, 0, sizeof(int64_t) * RN);
    for (int64_t i = 0; i < n; ++i) cnt[(src[i].key >> sh) & (RN - 1)]++;
    int64_t run = 0;
    for (int d = 0; d < RN; ++d) { int64_t c = cnt[d]; cnt[d] = run; run += c; }
    for (int64_t i = 0; i < n; ++i) dst[cnt[(src[i].key >> sh) & (RN - 1)]++] = src[i];
    kv_t* t = src; src = dst; dst = t;
  }
  if (src != kv) memcpy(kv, src, sizeof(kv_t) * (size_t)n);
  free(cnt);
  free(tmp);
  return 0;
}

static int grid_build(grid_t* g, const double* xyz, int64_t n, int dim,
                      double eps) {
  memset(g, 0, sizeof(*g));
  g->dim = dim;
  double mx[3];
  for (int k = 0; k < dim; ++k) { g->mn[k] = INFINITY; mx[k] = -INFINITY; }
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < dim; ++k) {
      double v = xyz[i * dim + k];
      if (v < g->mn[k]) g->mn[k] = v;
      if (v > mx[k]) mx[k] = v;
    }
  for (int k = 0; k < dim; ++k) {
    double ext = (n > 0) ? mx[k] - g->mn[k] : 0.0;
    double cw = eps;
    /* keep at most 2^20 cells per axis so keys stay in 63 bits */
    if (ext / cw > 1048576.0) cw = ext / 1048576.0;
    g->inv[k] = 1.0 / cw;
    double t = floor(ext * g->inv[k]);
    g->nk[k] = (int64_t)t + 1;
    if (g->nk[k] < 1) g->nk[k] = 1;
  }
  kv_t* kv = (kv_t*)malloc(sizeof(kv_t) * (size_t)(n > 0 ? n : 1));
  if (!kv) return -1;
  for (int64_t i = 0; i < n; ++i) {
    int64_t c[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) c[k] = gcell(g, k, xyz[i * dim + k]);
    kv[i].key = gkey(g, c);
    kv[i].id = (int32_t)i;
  }
  if (kv_sort(kv, n, g)) { free(kv); return -1; }
  g->idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  g->key = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  g->start = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
  if (!g->idx || !g->key || !g->start) { free(kv); return -1; }
  int64_t nc = 0;
  for (int64_t i = 0; i < n; ++i) {
    g->idx[i] = kv[i].id;
    if (i == 0 || kv[i].key != kv[i - 1].key) {
      g->key[nc] = kv[i].key;
      g->start[nc] = i;
      ++nc;
    }
  }
  g->start[nc] = n;
  g->ncell = nc;
  free(kv);
  return 0;
}

static void grid_free(grid_t* g) {
  free(g->idx);
  free(g->key);
  free(g->start);
}

static int64_t lower_bound64(const int64_t* a, int64_t n, int64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* Calls fn(q, ctx) for every point q in grid cells whose coordinates lie in
 * [lo_k, hi_k] per axis.  Returns early (1) when fn returns nonzero. */
typedef int (*visit_fn)(int32_t q, void* ctx);

static int grid_visit(const grid_t* g, const int64_t* lo, const int64_t* hi,
                      visit_fn fn, void* ctx) {
  int dim = g->dim;
  int64_t c[3];
  int last = dim - 1;
  /* iterate all axes except the last; the last axis is a key range */
  int64_t a0lo = dim >= 2 ? lo[0] : 0, a0hi = dim >= 2 ? hi[0] : 0;
  int64_t a1lo = dim == 3 ? lo[1] : 0, a1hi = dim == 3 ? hi[1] : 0;
  for (int64_t c0 = a0lo; c0 <= a0hi; ++c0) {
    for (int64_t c1 = a1lo; c1 <= a1hi; ++c1) {
      if (dim == 2) { c[0] = c0; }
      else { c[0] = c0; c[1] = c1; }
      c[last] = lo[last];
      int64_t klo = gkey(g, c);
      c[last] = hi[last];
      int64_t khi = gkey(g, c);
      int64_t s = lower_bound64(g->key, g->ncell, klo);
      for (; s < g->ncell && g->key[s] <= khi; ++s)
        for (int64_t j = g->start[s]; j < g->start[s + 1]; ++j)
          if (fn(g->idx[j], ctx)) return 1;
    }
  }
  return 0;
}

/* query cell range for the closed box [fl(p-eps), fl(p+eps)], widened by
 * `slack` cells (used when the box is centred on the other point). */
static void box_range(const grid_t* g, const double* p, double eps, int slack,
                      int64_t* lo, int64_t* hi) {
  for (int k = 0; k < g->dim; ++k) {
    lo[k] = gcell(g, k, p[k] - eps) - slack;
    hi[k] = gcell(g, k, p[k] + eps) + slack;
    if (lo[k] < 0) lo[k] = 0;
    if (hi[k] > g->nk[k] - 1) hi[k] = g->nk[k] - 1;
  }
}

/* ------------------------------------------------------------------------ */
/* Faithful sequential DBSCAN::run() (dbscan_rtree_exp.cpp:360-406).        */
/* ------------------------------------------------------------------------ */
typedef struct { int32_t* v; int64_t n, cap; } vec_t;
static int vec_push(vec_t* a, int32_t x) {
  if (a->n == a->cap) {
    int64_t nc = a->cap ? a->cap * 2 : 64;
    int32_t* nv = (int32_t*)realloc(a->v, sizeof(int32_t) * (size_t)nc);
    if (!nv) return -1;
    a->v = nv;
    a->cap = nc;
  }
  a->v[a->n++] = x;
  return 0;
}

typedef struct {
  const double* xyz; int dim; double eps, eps2; const double* p; vec_t* out;
  int err;
} nbr_ctx;

static int nbr_visit(int32_t q, void* vctx) {
  nbr_ctx* c = (nbr_ctx*)vctx;
  if (pred_inline(c->p, c->xyz + (int64_t)q * c->dim, c->dim, c->eps, c->eps2))
    if (vec_push(c->out, q)) { c->err = 1; return 1; }
  return 0;
}

static int find_neighbors(const grid_t* g, const double* xyz, int dim,
                          double eps, int64_t i, vec_t* out) {
  int64_t lo[3], hi[3];
  const double* p = xyz + i * dim;
  box_range(g, p, eps, 0, lo, hi);
  nbr_ctx c = {xyz, dim, eps, eps * eps, p, out, 0};
  grid_visit(g, lo, hi, nbr_visit, &c);
  return c.err ? -1 : 0;
}

int oracle_seq_dbscan(const double* xyz, int64_t n, int dim, double eps,
                      int minpts, int32_t* labels) {
  if (dim != 2 && dim != 3) return -1;
  grid_t g;
  if (grid_build(&g, xyz, n, dim, eps)) return -2;
  int32_t* clusters = (int32_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(int32_t));
  vec_t nb = {0, 0, 0}, more = {0, 0, 0};
  int32_t cluster = 0;
  int rc = 0;
  for (int64_t i = 0; i < n && rc == 0; ++i) {
    if (clusters[i] != 0) continue;
    nb.n = 0;
    if (find_neighbors(&g, xyz, dim, eps, i, &nb)) { rc = -3; break; }
    if (nb.n < minpts) { clusters[i] = -1; continue; }
    ++cluster;
    clusters[i] = cluster;
    for (int64_t j = 0; j < nb.n; ++j) {
      int32_t d = nb.v[j];
      if (d == i) continue;
      if (clusters[d] == -1) { clusters[d] = cluster; continue; }
      if (clusters[d] != 0) continue;
      clusters[d] = cluster;
      more.n = 0;
      if (find_neighbors(&g, xyz, dim, eps, d, &more)) { rc = -3; break; }
      if (more.n >= minpts)
        for (int64_t x = 0; x < more.n; ++x)
          if (vec_push(&nb, more.v[x])) { rc = -3; break; }
    }
  }
  for (int64_t i = 0; i < n; ++i) labels[i] = clusters[i] < 0 ? 0 : clusters[i];
  free(nb.v);
  free(more.v);
  free(clusters);
  grid_free(&g);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Fast exact formulation.  Cluster id = rank of a core component by its    */
/* minimum core index; border = min id over adjacent cores (from the core's */
/* side); noise = 0.  Equal to oracle_seq_dbscan whenever the predicate is  */
/* symmetric on core-core pairs (asym_pairs_out == 0).                      */
/* ------------------------------------------------------------------------ */
static inline int32_t uf_find(int32_t* par, int32_t x) {
  for (;;) {
    int32_t p = __atomic_load_n(&par[x], __ATOMIC_RELAXED);
    if (p == x) return x;
    int32_t gp = __atomic_load_n(&par[p], __ATOMIC_RELAXED);
    if (gp != p) __atomic_store_n(&par[x], gp, __ATOMIC_RELAXED);
    x = p;
  }
}

static void uf_union(int32_t* par, int32_t a, int32_t b) {
  for (;;) {
    a = uf_find(par, a);
    b = uf_find(par, b);
    if (a == b) return;
    if (a < b) { int32_t t = a; a = b; b = t; }
    /* link the larger root under the smaller: the root stays the minimum */
    int32_t expect = a;
    if (__atomic_compare_exchange_n(&par[a], &expect, b, 0, __ATOMIC_ACQ_REL,
                                    __ATOMIC_RELAXED))
      return;
  }
}

typedef struct {
  const double* xyz; int dim; double eps, eps2; const double* p; int64_t cnt;
  int64_t stop;
} cnt_ctx;
static int cnt_visit(int32_t q, void* vctx) {
  cnt_ctx* c = (cnt_ctx*)vctx;
  if (pred_inline(c->p, c->xyz + (int64_t)q * c->dim, c->dim, c->eps, c->eps2))
    if (++c->cnt >= c->stop) return 1;
  return 0;
}

typedef struct {
  const double* xyz; int dim; double eps, eps2; int32_t self; const double* p;
  const uint8_t* core; int32_t* par; int64_t asym;
} uni_ctx;
static int uni_visit(int32_t q, void* vctx) {
  uni_ctx* c = (uni_ctx*)vctx;
  if (q == c->self || !c->core[q]) return 0;
  const double* qp = c->xyz + (int64_t)q * c->dim;
  if (pred_inline(c->p, qp, c->dim, c->eps, c->eps2)) {
    if (!pred_inline(qp, c->p, c->dim, c->eps, c->eps2)) c->asym++;
    uf_union(c->par, c->self, q);
  }
  return 0;
}

typedef struct {
  const double* xyz; int dim; double eps, eps2; const double* b;
  const uint8_t* core; const int32_t* cid; int32_t best;
} bor_ctx;
static int bor_visit(int32_t q, void* vctx) {
  bor_ctx* c = (bor_ctx*)vctx;
  if (!c->core[q]) return 0;
  /* evaluated from the core point's side, as its expansion would */
  if (pred_inline(c->xyz + (int64_t)q * c->dim, c->b, c->dim, c->eps, c->eps2)) {
    int32_t id = c->cid[q];
    if (c->best == 0 || id < c->best) c->best = id;
  }
  return 0;
}

int oracle_fast_dbscan(const double* xyz, int64_t n, int dim, double eps,
                       int minpts, int nthreads, int32_t* labels,
                       uint8_t* core_out, int64_t* asym_pairs_out) {
  if (dim != 2 && dim != 3) return -1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  grid_t g;
  if (grid_build(&g, xyz, n, dim, eps)) return -2;
  size_t nn = (size_t)(n > 0 ? n : 1);
  uint8_t* core = (uint8_t*)calloc(nn, 1);
  int32_t* par = (int32_t*)malloc(sizeof(int32_t) * nn);
  int32_t* cid = (int32_t*)calloc(nn, sizeof(int32_t));
  double eps2 = eps * eps;
  int64_t asym = 0;

#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    int64_t lo[3], hi[3];
    const double* p = xyz + i * dim;
    box_range(&g, p, eps, 0, lo, hi);
    cnt_ctx c = {xyz, dim, eps, eps2, p, 0, minpts};
    grid_visit(&g, lo, hi, cnt_visit, &c);
    core[i] = c.cnt >= minpts;
  }
  for (int64_t i = 0; i < n; ++i) par[i] = (int32_t)i;

#pragma omp parallel for schedule(dynamic, 256) reduction(+ : asym)
  for (int64_t i = 0; i < n; ++i) {
    if (!core[i]) continue;
    int64_t lo[3], hi[3];
    const double* p = xyz + i * dim;
    box_range(&g, p, eps, 0, lo, hi);
    uni_ctx c = {xyz, dim, eps, eps2, (int32_t)i, p, core, par, 0};
    grid_visit(&g, lo, hi, uni_visit, &c);
    asym += c.asym;
  }

  int32_t k = 0;
  for (int64_t i = 0; i < n; ++i)
    if (core[i] && uf_find(par, (int32_t)i) == i) cid[i] = ++k;
  for (int64_t i = 0; i < n; ++i)
    if (core[i]) cid[i] = cid[uf_find(par, (int32_t)i)];

#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    if (core[i]) { labels[i] = cid[i]; continue; }
    int64_t lo[3], hi[3];
    const double* b = xyz + i * dim;
    box_range(&g, b, eps, 1, lo, hi);
    bor_ctx c = {xyz, dim, eps, eps2, b, core, cid, 0};
    grid_visit(&g, lo, hi, bor_visit, &c);
    labels[i] = c.best;
  }
  if (core_out) memcpy(core_out, core, (size_t)n);
  if (asym_pairs_out) *asym_pairs_out = asym;
  free(core);
  free(par);
  free(cid);
  grid_free(&g);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* The same labels for dense inputs, where the union step above visits       */
/* thousands of neighbours per core point.  Core-core unions run over clique */
/* sub-cells of side wc = 0.99 eps / sqrt(dim): any two points of one        */
/* sub-cell differ by less than wc per axis, so d^2 < 0.9801 eps^2 (rounding */
/* included) and every box test passes -- the core points of a sub-cell are */
/* one set.  Two sub-cells at most ceil(eps / wc) = 2 apart per axis are     */
/* united when some core pair across them satisfies P from either side      */
/* (the union rule above); the scan of a pair stops at the first joined     */
/* pair and is skipped when the two are already one set.  Near pairs (one   */
/* cell apart) go first, so that far pairs inside a cluster are skipped.    */
/* Core flags and border labels are computed as above.                      */
/* ------------------------------------------------------------------------ */
static void key_coords(const grid_t* g, int64_t key, int64_t* c) {
  for (int k = g->dim - 1; k >= 0; --k) {
    c[k] = key % g->nk[k];
    key /= g->nk[k];
  }
}

/* Unite sub-cell a with every sub-cell b > a in its 5^D stencil whose
 * offset is near (max |d| <= 1) or far (max |d| == 2), per `far`. */
static void cells_pass(const grid_t* s, const double* xyz, int dim, double eps,
                       const uint8_t* core, const int32_t* rep, int32_t* par,
                       int far) {
  const double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t a = 0; a < s->ncell; ++a) {
    if (rep[a] < 0) continue;
    int64_t ca[3] = {0, 0, 0}, cb[3] = {0, 0, 0};
    key_coords(s, s->key[a], ca);
    const int last = dim - 1;
    for (int d0 = -2; d0 <= 2; ++d0)
      for (int d1 = (dim == 3 ? -2 : 0); d1 <= (dim == 3 ? 2 : 0); ++d1) {
        cb[0] = ca[0] + d0;
        if (dim == 3) cb[1] = ca[1] + d1;
        if (cb[0] < 0 || cb[0] >= s->nk[0]) continue;
        if (dim == 3 && (cb[1] < 0 || cb[1] >= s->nk[1])) continue;
        int64_t zlo = ca[last] - 2 < 0 ? 0 : ca[last] - 2;
        int64_t zhi = ca[last] + 2 > s->nk[last] - 1 ? s->nk[last] - 1 : ca[last] + 2;
        cb[last] = zlo;
        int64_t klo = gkey(s, cb);
        cb[last] = zhi;
        int64_t khi = gkey(s, cb);
        for (int64_t b = lower_bound64(s->key, s->ncell, klo);
             b < s->ncell && s->key[b] <= khi; ++b) {
          if (b <= a || rep[b] < 0) continue;
          int64_t cc[3] = {0, 0, 0};
          key_coords(s, s->key[b], cc);
          int m = 0;
          for (int k = 0; k < dim; ++k) {
            int64_t d = cc[k] - ca[k];
            if (d < 0) d = -d;
            if (d > m) m = (int)d;
          }
          if ((m == 2) != (far != 0)) continue;
          if (uf_find(par, rep[a]) == uf_find(par, rep[b])) continue;
          int joined = 0;
          for (int64_t i = s->start[a]; i < s->start[a + 1] && !joined; ++i) {
            const int32_t p = s->idx[i];
            if (!core[p]) continue;
            const double* pp = xyz + (int64_t)p * dim;
            for (int64_t j = s->start[b]; j < s->start[b + 1]; ++j) {
              const int32_t q = s->idx[j];
              if (!core[q]) continue;
              const double* qp = xyz + (int64_t)q * dim;
              if (pred_inline(pp, qp, dim, eps, eps2) || pred_inline(qp, pp, dim, eps, eps2)) {
                joined = 1;
                break;
              }
            }
          }
          if (joined) uf_union(par, rep[a], rep[b]);
        }
      }
  }
}

int oracle_cells_dbscan(const double* xyz, int64_t n, int dim, double eps,
                        int minpts, int nthreads, int32_t* labels,
                        uint8_t* core_out) {
  if (dim != 2 && dim != 3) return -1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const double wc = 0.99 * eps / sqrt((double)dim);
  grid_t g, s;
  if (grid_build(&g, xyz, n, dim, eps)) return -2;
  if (grid_build(&s, xyz, n, dim, wc)) { grid_free(&g); return -2; }
  /* grid_build widens cells of huge extents; the sub-cells must stay cliques */
  for (int k = 0; k < dim; ++k)
    if (s.inv[k] != 1.0 / wc) { grid_free(&g); grid_free(&s); return -4; }
  size_t nn = (size_t)(n > 0 ? n : 1);
  uint8_t* core = (uint8_t*)calloc(nn, 1);
  int32_t* par = (int32_t*)malloc(sizeof(int32_t) * nn);
  int32_t* cid = (int32_t*)calloc(nn, sizeof(int32_t));
  int32_t* rep = (int32_t*)malloc(sizeof(int32_t) * (size_t)(s.ncell > 0 ? s.ncell : 1));
  if (!core || !par || !cid || !rep) return -2;
  double eps2 = eps * eps;

  /* a point whose sub-cell holds at least minpts points is core: the
   * sub-cell is a clique, so all of them are its neighbours */
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < s.ncell; ++c)
    if (s.start[c + 1] - s.start[c] >= minpts)
      for (int64_t j = s.start[c]; j < s.start[c + 1]; ++j) core[s.idx[j]] = 1;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    if (core[i]) continue;
    int64_t lo[3], hi[3];
    const double* p = xyz + i * dim;
    box_range(&g, p, eps, 0, lo, hi);
    cnt_ctx c = {xyz, dim, eps, eps2, p, 0, minpts};
    grid_visit(&g, lo, hi, cnt_visit, &c);
    core[i] = c.cnt >= minpts;
  }
  for (int64_t i = 0; i < n; ++i) par[i] = (int32_t)i;
  /* a sub-cell's core points: one set, rooted at its smallest id (the
   * sub-cell lists ids in increasing order) */
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < s.ncell; ++c) {
    int32_t r = -1;
    for (int64_t j = s.start[c]; j < s.start[c + 1]; ++j) {
      const int32_t q = s.idx[j];
      if (!core[q]) continue;
      if (r < 0) r = q; else par[q] = r;
    }
    rep[c] = r;
  }
  cells_pass(&s, xyz, dim, eps, core, rep, par, 0);
  cells_pass(&s, xyz, dim, eps, core, rep, par, 1);

  int32_t k = 0;
  for (int64_t i = 0; i < n; ++i)
    if (core[i] && uf_find(par, (int32_t)i) == i) cid[i] = ++k;
  for (int64_t i = 0; i < n; ++i)
    if (core[i]) cid[i] = cid[uf_find(par, (int32_t)i)];

#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    if (core[i]) { labels[i] = cid[i]; continue; }
    int64_t lo[3], hi[3];
    const double* b = xyz + i * dim;
    box_range(&g, b, eps, 1, lo, hi);
    bor_ctx c = {xyz, dim, eps, eps2, b, core, cid, 0};
    grid_visit(&g, lo, hi, bor_visit, &c);
    labels[i] = c.best;
  }
  if (core_out) memcpy(core_out, core, (size_t)n);
  free(core);
  free(par);
  free(cid);
  free(rep);
  grid_free(&g);
  grid_free(&s);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Sampled check of given labels against the definition, by enumerating    */
/* each sampled point's eps-ball with the predicate (no labels of the       */
/* oracle involved).  For sampled point s with core flag core[s]:           */
/*   core s:     every core q with P(s,q) or P(q,s) has labels[q] == labels[s] */
/*               (and labels[s] > 0);                                       */
/*   non-core s: labels[s] == min{labels[q] : q core, P(q,s)}, 0 if none.   */
/* out[0] core-label violations, out[1] border violations, out[2] predicate */
/* evaluations, out[3] core points sampled.                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
  const double* xyz; int dim; double eps, eps2; int32_t self; const double* p;
  const uint8_t* core; const int32_t* lab; int64_t bad, evals; int32_t best;
} smp_ctx;

static int smp_core_visit(int32_t q, void* vctx) {
  smp_ctx* c = (smp_ctx*)vctx;
  c->evals++;
  if (q == c->self || !c->core[q]) return 0;
  const double* qp = c->xyz + (int64_t)q * c->dim;
  if (pred_inline(c->p, qp, c->dim, c->eps, c->eps2) ||
      pred_inline(qp, c->p, c->dim, c->eps, c->eps2))
    if (c->lab[q] != c->lab[c->self]) c->bad++;
  return 0;
}

static int smp_border_visit(int32_t q, void* vctx) {
  smp_ctx* c = (smp_ctx*)vctx;
  c->evals++;
  if (!c->core[q]) return 0;
  const double* qp = c->xyz + (int64_t)q * c->dim;
  if (pred_inline(qp, c->p, c->dim, c->eps, c->eps2))
    if (c->best == 0 || c->lab[q] < c->best) c->best = c->lab[q];
  return 0;
}

int oracle_sample_check(const double* xyz, int64_t n, int dim, double eps,
                        const int32_t* labels, const uint8_t* core,
                        const int64_t* sample, int64_t m, int nthreads,
                        int64_t* out) {
  if (dim != 2 && dim != 3) return -1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  grid_t g;
  if (grid_build(&g, xyz, n, dim, eps)) return -2;
  const double eps2 = eps * eps;
  int64_t bad_core = 0, bad_border = 0, evals = 0, ncore = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : bad_core, bad_border, evals, ncore)
  for (int64_t t = 0; t < m; ++t) {
    const int64_t i = sample[t];
    if (i < 0 || i >= n) { bad_core++; continue; }
    int64_t lo[3], hi[3];
    const double* p = xyz + i * dim;
    smp_ctx c = {xyz, dim, eps, eps2, (int32_t)i, p, core, labels, 0, 0, 0};
    box_range(&g, p, eps, 1, lo, hi);
    if (core[i]) {
      ncore++;
      grid_visit(&g, lo, hi, smp_core_visit, &c);
      bad_core += c.bad + (labels[i] <= 0);
    } else {
      grid_visit(&g, lo, hi, smp_border_visit, &c);
      bad_border += labels[i] != c.best;
    }
    evals += c.evals;
  }
  out[0] = bad_core;
  out[1] = bad_border;
  out[2] = evals;
  out[3] = ncore;
  grid_free(&g);
  return 0;
}

int oracle_brute_counts(const double* xyz, int64_t n, int dim, double eps,
                        int32_t* counts) {
  double eps2 = eps * eps;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int32_t c = 0;
    for (int64_t j = 0; j < n; ++j)
      c += pred_inline(xyz + i * dim, xyz + j * dim, dim, eps, eps2);
    counts[i] = c;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* multicore-cpu/dbscan.cpp:56-209 restated: FP32 coordinates, sqrt        */
/* compare, self excluded (:143), BFS queue expansion (:165-209).           */
/* ------------------------------------------------------------------------ */
enum { MC_UNCLASSIFIED = -1, MC_NOISE = -2, MC_MARKED = 0 };

static int64_t mc_row(const float* pts, int64_t n, int dim, int64_t i,
                      float radius, char* row) {
  const float* me = pts + i * dim;
  int64_t cnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt)
  for (int64_t x = 0; x < n; ++x) {
    const float* o = pts + x * dim;
    float s = 0.0f;
    for (int k = 0; k < dim; ++k) {
      float d = me[k] - o[k];
      s = s + d * d;
    }
    float dist = sqrtf(s);
    if (dist <= radius) {
      row[x] = 1;
      if (x != i) cnt++;
    } else {
      row[x] = 0;
    }
  }
  return cnt;
}

int oracle_multicore_dbscan(const double* xyz, int64_t n, int dim, double eps,
                            int minpts, int nthreads, int32_t* labels) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  size_t nn = (size_t)(n > 0 ? n : 1);
  float* pts = (float*)malloc(sizeof(float) * nn * (size_t)dim);
  int32_t* cl = (int32_t*)malloc(sizeof(int32_t) * nn);
  char* row = (char*)malloc(nn);
  int64_t* queue = (int64_t*)malloc(sizeof(int64_t) * (nn * 2 + 1));
  if (!pts || !cl || !row || !queue) return -2;
  for (int64_t i = 0; i < n * dim; ++i) pts[i] = (float)xyz[i];
  float radius = (float)eps;
  for (int64_t i = 0; i < n; ++i) cl[i] = MC_UNCLASSIFIED;
  int32_t next_id = 1;
  for (int64_t i = 0; i < n; ++i) {
    if (cl[i] != MC_UNCLASSIFIED) continue;
    int64_t c = mc_row(pts, n, dim, i, radius, row);
    if (c < minpts) { cl[i] = MC_NOISE; continue; }
    cl[i] = next_id;
    int64_t qh = 0, qt = 0;
    for (int64_t x = 0; x < n; ++x)
      if (row[x]) { queue[qt++] = x; cl[x] = MC_MARKED; }
    while (qh < qt) {
      int64_t cur = queue[qh++];
      int64_t cc = mc_row(pts, n, dim, cur, radius, row);
      if (cc >= minpts) {
        cl[cur] = next_id;
        for (int64_t x = 0; x < n; ++x)
          if (row[x]) {
            if (cl[x] == MC_UNCLASSIFIED) {
              if (qt >= (int64_t)(nn * 2)) { /* compact the consumed head */
                memmove(queue, queue + qh, sizeof(int64_t) * (size_t)(qt - qh));
                qt -= qh; qh = 0;
              }
              queue[qt++] = x;
            }
            cl[x] = MC_MARKED;
          }
      } else {
        cl[cur] = MC_NOISE;
      }
    }
    ++next_id;
  }
  for (int64_t i = 0; i < n; ++i) labels[i] = cl[i] > 0 ? cl[i] : 0;
  free(pts);
  free(cl);
  free(row);
  free(queue);
  /* the program's one result: clusters started (dbscan.cpp:84-95) */
  return (int)(next_id - 1);
}

/* ------------------------------------------------------------------------ */
/* test-results/cluster-check.cpp:41-199 restated with a contingency table  */
/* (sort of (c1,c2) pairs) instead of per-point vector copies.              */
/* ------------------------------------------------------------------------ */
static int u64_cmp(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return (x > y) - (x < y);
}

int oracle_cluster_check(const int32_t* c1, const int32_t* c2, int64_t n,
                         double* out) {
  int32_t m1 = 0, m2 = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (c1[i] < 0 || c2[i] < 0) return -1;
    if (c1[i] > m1) m1 = c1[i];
    if (c2[i] > m2) m2 = c2[i];
  }
  int64_t* s1 = (int64_t*)calloc((size_t)m1 + 1, sizeof(int64_t));
  int64_t* s2 = (int64_t*)calloc((size_t)m2 + 1, sizeof(int64_t));
  uint64_t* pk = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n > 0 ? n : 1));
  if (!s1 || !s2 || !pk) return -2;
  for (int64_t i = 0; i < n; ++i) {
    s1[c1[i]]++;
    s2[c2[i]]++;
    pk[i] = ((uint64_t)(uint32_t)c1[i] << 32) | (uint32_t)c2[i];
  }
  qsort(pk, (size_t)n, sizeof(uint64_t), u64_cmp);
  int64_t noise_err = 0, noise_eq = 0;
  double lost = 0.0, sum = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double score;
    if ((c1[i] == 0) != (c2[i] == 0)) { noise_err++; score = 0.0; }
    else if (c1[i] == 0) { noise_eq++; score = 1.0; }
    else {
      uint64_t key = ((uint64_t)(uint32_t)c1[i] << 32) | (uint32_t)c2[i];
      /* equal_range on the sorted pair keys = |C1(i) n C2(i)| */
      int64_t lo = 0, hi = n;
      while (lo < hi) { int64_t m = lo + (hi - lo) / 2; if (pk[m] < key) lo = m + 1; else hi = m; }
      int64_t a = lo;
      lo = a; hi = n;
      while (lo < hi) { int64_t m = lo + (hi - lo) / 2; if (pk[m] <= key) lo = m + 1; else hi = m; }
      int64_t inter = lo - a;
      int64_t uni = s1[c1[i]] + s2[c2[i]] - inter;
      score = (inter * 1.0) / (uni * 1.0);
      if (score != 1.0) lost += 1.0 - score;
    }
    sum += score;
  }
  out[0] = (double)noise_err;
  out[1] = (double)noise_eq;
  out[2] = n > 0 ? lost / (n * 1.0) : 0.0;
  out[3] = n > 0 ? (1.0 * sum) / (n * 1.0) : 0.0;
  free(s1);
  free(s2);
  free(pk);
  return 0;
}
