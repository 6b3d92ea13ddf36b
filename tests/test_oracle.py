"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

- against the reference's own FP64 R-tree DBSCAN compiled in place
  (oracle/_ref) and the committed golden labels it produced;
- the fast formulation against the faithful sequential restatement (2-D and
  3-D) -- the two share only the predicate;
- core flags against an O(N^2) brute-force count, labels against the
  definition evaluated by brute force (2-D and 3-D);
- the clique-sub-cell formulation (cells_dbscan, the checker for dense
  full-size inputs) against the fast one, the golden labels and brute force;
  the sampled definition check against exact labels and planted errors;
- the predicate against hand-built boundary pairs (closed box, exact eps^2
  ties, FMA-sensitive sums) whose answers are fixed by IEEE arithmetic;
- the cluster-check restatement against the reference checker compiled in
  place, and against hand-computed Jaccard scores;
- the multicore-cpu restatement (a timing baseline) by its reported cluster
  count: the minpts threshold without the point itself, the <= tie, expansion
  through cores, and the exact oracle's count where no pair lies near eps.
"""
import hashlib
import os
import subprocess

import numpy as np
import pytest

import dclust
import oracle

EX = np.array([[1, 1], [2, 2], [3, 3], [10, 10], [11, 11], [12, 12], [20, 20], [21, 21],
               [22, 22]], dtype=np.float64)

need_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


def test_ex_txt_known_answer(golden):
    # common.h defaults: 9 points, minpts 3, eps 1.5 -> 3 clusters, 0 noise
    lab = oracle.seq_dbscan(EX, 1.5, 3)
    assert lab.tolist() == [1, 1, 1, 2, 2, 2, 3, 3, 3]
    assert golden["ex_labels"].tolist() == lab.tolist()
    f, core, asym = oracle.fast_dbscan(EX, 1.5, 3)
    assert f.tolist() == lab.tolist() and core.tolist() == [0, 1, 0, 0, 1, 0, 0, 1, 0]


def test_cfg1_golden(golden):
    x, info = dclust.synth_config(1)
    assert hashlib.sha256(x.tobytes()).digest() == golden["cfg1_sha256"].tobytes()
    ref = golden["cfg1_labels"].astype(np.int32)
    assert np.array_equal(oracle.seq_dbscan(x, info["eps"], info["minpts"]), ref)
    assert np.array_equal(oracle.fast_dbscan(x, info["eps"], info["minpts"])[0], ref)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_lattice_golden(golden, k):
    x = golden[f"lat{k}_x"]
    eps, mp = golden[f"lat{k}_par"]
    ref = golden[f"lat{k}_labels"].astype(np.int32)
    assert np.array_equal(oracle.seq_dbscan(x, float(eps), int(mp)), ref)
    lab, _, asym = oracle.fast_dbscan(x, float(eps), int(mp))
    assert asym == 0 and np.array_equal(lab, ref)


@need_ref
@pytest.mark.parametrize("seed", range(4))
def test_seq_matches_compiled_reference_2d(seed):
    rng = np.random.default_rng(seed)
    n = 4000
    c = rng.uniform(0, 10, (6, 2))
    x = c[rng.integers(0, 6, n)] + rng.normal(0, 0.4, (n, 2))
    for eps, mp in [(0.1, 3), (0.25, 6), (0.5, 12)]:
        ref = oracle.ref_rtree_dbscan(x, eps, mp)
        assert np.array_equal(oracle.seq_dbscan(x, eps, mp), ref)
        assert np.array_equal(oracle.fast_dbscan(x, eps, mp)[0], ref)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", range(3))
def test_fast_matches_seq(dim, seed):
    rng = np.random.default_rng(10 + seed)
    n = 3000
    c = rng.uniform(0, 1, (5, dim))
    x = c[rng.integers(0, 5, n)] + rng.normal(0, 0.05, (n, dim))
    x[:300] = rng.uniform(0, 1, (300, dim))
    for eps, mp in [(0.02, 3), (0.05, 8)]:
        s = oracle.seq_dbscan(x, eps, mp)
        f, core, asym = oracle.fast_dbscan(x, eps, mp)
        assert asym == 0
        assert np.array_equal(s, f)


@pytest.mark.parametrize("dim", [2, 3])
def test_core_flags_vs_brute_force(dim):
    rng = np.random.default_rng(5)
    x = rng.normal(0, 1, (2500, dim))
    eps, mp = 0.2, 5
    cnt = oracle.brute_counts(x, eps)
    _, core, _ = oracle.fast_dbscan(x, eps, mp)
    assert np.array_equal(core, (cnt >= mp).astype(np.uint8))
    # a core point has a label; a point with count 1 (only itself) is noise
    lab = oracle.seq_dbscan(x, eps, mp)
    assert np.all(lab[cnt >= mp] > 0)


def _brute_labels(x, eps, mp):
    """Labels by the definition: components of the core graph (P from either
    side joins two core points), numbered in order of their smallest core
    index; a border point takes the smallest id of a core point whose
    predicate reaches it; noise 0.  O(N^2) predicate calls."""
    n = x.shape[0]
    adj = np.zeros((n, n), dtype=bool)
    for i in range(n):
        for j in range(n):
            adj[i, j] = oracle.pred(x[i], x[j], eps)
    core = adj.sum(axis=1) >= mp
    comp = np.zeros(n, dtype=np.int32)
    k = 0
    for s in range(n):
        if core[s] and comp[s] == 0:
            k += 1
            stack = [s]
            comp[s] = k
            while stack:
                u = stack.pop()
                for v in np.nonzero((adj[u] | adj[:, u]) & core)[0]:
                    if comp[v] == 0:
                        comp[v] = k
                        stack.append(v)
    lab = comp.copy()
    for b in np.nonzero(~core)[0]:
        ids = [comp[q] for q in np.nonzero(core & adj[:, b])[0]]
        lab[b] = min(ids) if ids else 0
    return lab, core.astype(np.uint8)


@pytest.mark.parametrize("dim,eps", [(2, 0.15), (3, 0.4)])
def test_labels_equal_brute_force_definition(dim, eps):
    rng = np.random.default_rng(60 + dim)
    x = rng.normal(0, 1, (700, dim))
    x[:40] = x[40:80] + rng.normal(0, 0.3 * eps, (40, dim))  # some dense spots
    mp = 4
    ref, rcore = _brute_labels(x, eps, mp)
    assert 3 < ref.max() and (ref == 0).sum() > 20 and (rcore == 0).sum() > (ref == 0).sum()
    lab, core, _ = oracle.fast_dbscan(x, eps, mp)
    assert np.array_equal(core, rcore) and np.array_equal(lab, ref)
    lab, core = oracle.cells_dbscan(x, eps, mp)
    assert np.array_equal(core, rcore) and np.array_equal(lab, ref)
    assert np.array_equal(oracle.seq_dbscan(x, eps, mp), ref)


def _dense_blobs(dim, n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 1, (6, dim))
    x = c[rng.integers(0, 6, n)] + rng.normal(0, 0.004, (n, dim))
    k = n // 50
    x[:k] = rng.uniform(0, 1, (k, dim))
    t = rng.uniform(0, 1, (n // 20, 1))
    x[k:k + n // 20] = c[0] + t * (c[1] - c[0]) + rng.normal(0, 2e-4, (n // 20, dim))
    return x


@pytest.mark.parametrize("case", ["blobs2", "blobs3", "dense2", "dense3", "dense3_mp60",
                                  "cfg3", "cfg5", "cfg2"])
def test_cells_matches_fast(case):
    if case.startswith("blobs"):
        dim = int(case[-1])
        rng = np.random.default_rng(70 + dim)
        x = rng.uniform(0, 1, (6, dim))[rng.integers(0, 6, 30000)] + rng.normal(0, 0.03,
                                                                                (30000, dim))
        eps, mp = (0.01, 5) if dim == 2 else (0.02, 5)
    elif case.startswith("dense"):
        dim = int(case[5])
        x = _dense_blobs(dim, 60000, 80 + dim)
        eps, mp = 0.002, (60 if case.endswith("mp60") else 20)
    else:
        x, info = dclust.synth_config(int(case[-1]), n=150000)
        eps, mp = info["eps"], info["minpts"]
    ref, rcore, asym = oracle.fast_dbscan(x, eps, mp)
    assert asym == 0 and ref.max() > 1
    lab, core = oracle.cells_dbscan(x, eps, mp)
    assert np.array_equal(core, rcore)
    assert np.array_equal(lab, ref)


def test_cells_golden(golden):
    x, info = dclust.synth_config(1)
    lab, _ = oracle.cells_dbscan(x, info["eps"], info["minpts"])
    assert np.array_equal(lab, golden["cfg1_labels"].astype(np.int32))
    for k in range(3):
        eps, mp = golden[f"lat{k}_par"]
        lab, _ = oracle.cells_dbscan(golden[f"lat{k}_x"], float(eps), int(mp))
        assert np.array_equal(lab, golden[f"lat{k}_labels"].astype(np.int32))
    assert oracle.cells_dbscan(EX, 1.5, 3)[0].tolist() == [1, 1, 1, 2, 2, 2, 3, 3, 3]


def test_sample_check_accepts_exact_labels_and_finds_planted_errors():
    x = _dense_blobs(3, 40000, 91)
    eps, mp = 0.002, 20
    lab, core = oracle.cells_dbscan(x, eps, mp)
    every = np.arange(x.shape[0])
    bc, bb, ev, nc = oracle.sample_check(x, eps, lab, core, every)
    assert (bc, bb) == (0, 0) and nc == int(core.sum()) and ev > x.shape[0]
    # a core point moved to another cluster
    ci = int(np.nonzero(core)[0][0])
    bad = lab.copy()
    bad[ci] = lab.max() + 1
    assert oracle.sample_check(x, eps, bad, core, every)[0] > 0
    # a border point given a larger adjacent id / noise given a cluster
    nb = np.nonzero((core == 0) & (lab > 0))[0]
    assert nb.size > 0
    bad = lab.copy()
    bad[nb[0]] = lab[nb[0]] + 1
    assert oracle.sample_check(x, eps, bad, core, every)[1] == 1
    nz = np.nonzero(lab == 0)[0]
    bad = lab.copy()
    bad[nz[:3]] = 1
    assert oracle.sample_check(x, eps, bad, core, every)[1] == 3


def test_labels_are_components_brute_force():
    # partition of core points == connected components of the core graph
    rng = np.random.default_rng(6)
    x = rng.normal(0, 1, (800, 2))
    eps, mp = 0.15, 3
    lab, core, _ = oracle.fast_dbscan(x, eps, mp)
    n = x.shape[0]
    adj = np.zeros((n, n), dtype=bool)
    for i in range(n):
        for j in range(n):
            adj[i, j] = oracle.pred(x[i], x[j], eps)
    comp = -np.ones(n, dtype=int)
    k = 0
    for s in range(n):
        if core[s] and comp[s] < 0:
            k += 1
            stack = [s]
            comp[s] = k
            while stack:
                u = stack.pop()
                for v in np.nonzero(adj[u] & core.astype(bool))[0]:
                    if comp[v] < 0:
                        comp[v] = k
                        stack.append(v)
    # ids by smallest core index == component discovery order above
    assert np.array_equal(lab[core == 1], comp[core == 1])
    # borders: smallest adjacent component id, evaluated from the core side
    for b in np.nonzero(core == 0)[0]:
        ids = [comp[q] for q in range(n) if core[q] and adj[q, b]]
        assert lab[b] == (min(ids) if ids else 0)


def test_predicate_boundaries():
    eps = 0.5
    p = np.array([0.0, 0.0])
    # exactly eps away on an axis: d^2 == eps^2 and box edge == q -> neighbour
    assert oracle.pred(p, np.array([0.5, 0.0]), eps)
    assert not oracle.pred(p, np.array([np.nextafter(0.5, 1.0), 0.0]), eps)
    # closed box uses fl(p + eps): 0.1 + 0.2 rounds up past 0.3
    assert oracle.pred(np.array([0.1, 0.0]), np.array([0.3, 0.0]), 0.2) == (0.3 <= 0.1 + 0.2)
    # d^2 compared with fl(eps*eps), sum without FMA: (3,4) at eps 5 is a tie
    assert oracle.pred(np.array([0.0, 0.0]), np.array([3.0, 4.0]), 5.0)
    assert not oracle.pred(np.array([0.0, 0.0]), np.array([3.0, np.nextafter(4.0, 5.0)]), 5.0)
    # 3-D left-to-right sum: 1+4+4 = 9 = 3^2
    assert oracle.pred(np.zeros(3), np.array([1.0, 2.0, 2.0]), 3.0)
    # a point is its own neighbour
    assert oracle.pred(np.array([1e6, -3.0]), np.array([1e6, -3.0]), 1e-9)


def test_predicate_no_fma_contraction():
    # chosen so that fl(fl(a*a) + fl(b*b)) <= fl(e*e) but the fused form is not
    rng = np.random.default_rng(9)
    found = 0
    for _ in range(200000):
        a, b = rng.uniform(0.1, 1.0, 2)
        e = np.sqrt(a * a + b * b)
        sep = np.float64(a * a) + np.float64(b * b)
        fused = np.longdouble(a) * np.longdouble(a) + np.longdouble(b) * np.longdouble(b)
        if (sep <= e * e) != (np.float64(fused) <= e * e):
            found += 1
            assert oracle.pred(np.array([0.0, 0.0]), np.array([a, b]), e) == bool(sep <= e * e)
    assert found > 0


@need_ref
def test_cluster_check_matches_reference_tool(tmp_path):
    rng = np.random.default_rng(3)
    n = 3000
    a = rng.integers(0, 8, n).astype(np.int32)
    b = a.copy()
    flip = rng.random(n) < 0.1
    b[flip] = rng.integers(0, 8, flip.sum())
    pa, pb = tmp_path / "a.txt", tmp_path / "b.txt"
    np.savetxt(pa, a, fmt="%d")
    np.savetxt(pb, b, fmt="%d")
    out = subprocess.run([oracle.REF_CHECK, str(pa), str(pb)], capture_output=True, text=True,
                         check=True).stdout
    ne, eq, lost, score = oracle.cluster_check(a, b)
    assert f"mismatched noise points: {ne}, agreement noise points: {eq}" in out
    assert f"fraction lost due to mismatches between clusters: {lost:f}" in out
    assert f"Final Error metric: {score:f}" in out


def test_cluster_check_hand_values():
    assert oracle.cluster_check([1, 1, 2, 2], [1, 1, 2, 2]) == (0, 0, 0.0, 1.0)
    # point 0,1 in cluster {0,1} vs {0,1,2}: J = 2/3; point 2: {2,3} vs {0,1,2}: 1/4
    ne, eq, lost, score = oracle.cluster_check([1, 1, 2, 2], [1, 1, 1, 0])
    assert ne == 1 and eq == 0
    assert score == pytest.approx((2 / 3 + 2 / 3 + 1 / 4 + 0) / 4)


# multicore-cpu/dbscan.cpp restatement (the timed 16-thread baseline).  That
# program's one result is the number of clusters it reports; its minpts counts
# the other points only (dbscan.cpp:143) and it compares the FP32 sqrt distance
# with <= (dbscan.cpp:131-136).  Where no pair lies near eps, its clusters are
# the components of the core graph, i.e. the exact oracle's clusters at
# minpts + 1.

def _groups(rng, sizes, dim, eps):
    """Groups of points inside cubes of side eps/4 (every pair within eps), the
    cubes 10 eps apart, shuffled."""
    pts = []
    for g, m in enumerate(sizes):
        c = np.zeros(dim)
        c[0] = 10 * eps * g
        pts.append(c + rng.uniform(0, eps / 4, size=(m, dim)))
    x = np.concatenate(pts)
    return x[rng.permutation(x.shape[0])]


@pytest.mark.parametrize("dim", [2, 3])
def test_multicore_minpts_excludes_self(dim):
    rng = np.random.default_rng(11)
    eps, mp = 0.1, 4
    # groups of 4 points: 3 others each -> noise; groups of >= 5 -> one cluster each
    x = _groups(rng, [4, 5, 5, 1, 7, 4, 12], dim, eps)
    k, lab = oracle.multicore_dbscan(x, eps, mp, nthreads=2)
    assert k == 4
    assert int(oracle.seq_dbscan(x, eps, mp + 1).max()) == 4
    # every labelled point is marked again by its own row (dbscan.cpp:187-199)
    assert not lab.any()


def test_multicore_distance_tie_is_inside():
    # sqrt(0.25) == 0.5 exactly: the pair is within eps, each has 1 other point
    assert oracle.multicore_dbscan(np.array([[0.0, 0.0], [0.5, 0.0]]), 0.5, 1)[0] == 1
    assert oracle.multicore_dbscan(np.array([[0.0, 0.0], [0.5000001, 0.0]]), 0.5, 1)[0] == 0


def test_multicore_expands_through_cores():
    # two chains of 10 points 0.6 eps apart, interior points have 2 others:
    # each chain is one cluster, reached point by point from its first core
    eps = 1.0
    t = np.arange(10) * 0.6
    a = np.stack([t, np.zeros(10)], axis=1)
    b = np.stack([t + 100.0, np.full(10, 50.0)], axis=1)
    x = np.concatenate([a, b])[np.random.default_rng(3).permutation(20)]
    assert oracle.multicore_dbscan(x, eps, 2)[0] == 2
    assert oracle.multicore_dbscan(x, eps, 3)[0] == 0


@pytest.mark.parametrize("dim,seed", [(2, 0), (3, 6)])
def test_multicore_cluster_count_matches_oracle(dim, seed):
    rng = np.random.default_rng(seed)
    cent = rng.uniform(0, 1, size=(12, dim))
    x = np.concatenate([c + rng.normal(0, 0.02, size=(250, dim)) for c in cent]
                       + [rng.uniform(0, 1, size=(300, dim))])
    eps, mp = 0.012 if dim == 2 else 0.025, 6
    # precondition: no pair within 1e-5 relative of eps, so FP32 and FP64 agree
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    assert not (np.abs(d - eps) < 1e-5 * eps).any()
    k, _ = oracle.multicore_dbscan(x, eps, mp - 1, nthreads=2)
    ref = oracle.seq_dbscan(x, eps, mp)
    assert ref.max() > 3 and k == int(ref.max())
