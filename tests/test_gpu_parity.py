"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bar: labels byte-identical to the oracle's canonical labels (cluster id =
rank of the smallest core input index, border = smallest adjacent id,
noise = 0), core flags bit-identical, in both precision forms.
"""
import hashlib

import numpy as np
import pytest

import dclust
import oracle

pytestmark = pytest.mark.gpu

PRECS = [dclust.FP64, dclust.FP32_GUARD]
EX = np.array([[1, 1], [2, 2], [3, 3], [10, 10], [11, 11], [12, 12], [20, 20], [21, 21],
               [22, 22]], dtype=np.float64)


def _check(handle, x, eps, minpts, prec, **kw):
    lab = handle.run(x, eps, minpts, precision=prec, **kw)
    ref, core, asym = oracle.fast_dbscan(x, eps, minpts)
    assert asym == 0
    got_core = handle.core_flags(x.shape[0])
    nbad_core = int(np.sum(got_core != core))
    nbad = int(np.sum(lab != ref))
    assert nbad_core == 0, f"{nbad_core} core flags differ"
    assert nbad == 0, f"{nbad} labels differ"
    st = handle.last_stats
    # every cluster holds at least one chain; a chain starts at a core cell
    assert st["n_clusters"] <= st["n_chains"] <= st["n_core"]
    return st


@pytest.mark.parametrize("prec", PRECS)
def test_ex_txt(handle, golden, prec):
    # test-datasets/ex.txt with the common.h defaults (eps 1.5, minpts 3, r 3)
    lab = handle.run(EX, 1.5, 3, r=3, precision=prec)
    assert lab.tolist() == golden["ex_labels"].tolist() == [1, 1, 1, 2, 2, 2, 3, 3, 3]
    assert handle.core_flags(9).tolist() == [0, 1, 0, 0, 1, 0, 0, 1, 0]


@pytest.mark.parametrize("prec", PRECS)
def test_cfg1_reference_labels(handle, golden, prec):
    x, info = dclust.synth_config(1)
    assert hashlib.sha256(x.tobytes()).digest() == golden["cfg1_sha256"].tobytes()
    lab = handle.run(x, info["eps"], info["minpts"], r=info["partitions_r"], precision=prec)
    assert np.array_equal(lab, golden["cfg1_labels"].astype(np.int32))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("k", [0, 1, 2])
def test_lattice_ties_reference_labels(handle, golden, prec, k):
    x = golden[f"lat{k}_x"]
    eps, mp = golden[f"lat{k}_par"]
    lab = handle.run(x, float(eps), int(mp), precision=prec)
    assert np.array_equal(lab, golden[f"lat{k}_labels"].astype(np.int32))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_blobs(handle, prec, dim, seed):
    rng = np.random.default_rng(100 + seed)
    n = 20000
    centres = rng.uniform(0, 1, (12, dim))
    x = centres[rng.integers(0, 12, n)] + rng.normal(0, 0.02, (n, dim))
    x[: n // 10] = rng.uniform(0, 1, (n // 10, dim))
    eps = 0.01 if dim == 2 else 0.02
    _check(handle, x, eps, 5, prec)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("seeds", [1, 2, 7])
def test_tiny_seed_list_forces_correction(handle, prec, seeds):
    # seed lists of 1-2 cells overflow constantly: claimed cells go to the
    # correction queue and are expanded by later passes (few expansion waves,
    # so that each grows its chains instead of seeding one chain per cell)
    rng = np.random.default_rng(7)
    x = rng.normal(0, 1, (8000, 2))
    st = _check(handle, x, 0.08, 4, prec, seeds=seeds, blocks=8)
    if seeds <= 2:
        assert st["n_overflow"] > 0 and st["n_passes"] > 1


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("budget", ["1", "3"])
def test_tiny_pop_budget_many_passes(handle, prec, budget, monkeypatch):
    # every wave hands its list on after `budget` seed batches: many passes,
    # segments resumed across passes, leftover queue entries carried over
    monkeypatch.setenv("DCLUST_EXPAND_BUDGET", budget)
    x, info = dclust.synth_config(3, n=120000)
    st = _check(handle, x, info["eps"], info["minpts"], prec, blocks=16)
    assert st["n_passes"] > 2


@pytest.mark.parametrize("prec", PRECS)
def test_tiny_pair_buffer_merges_inline(handle, prec):
    rng = np.random.default_rng(8)
    x = rng.normal(0, 1, (30000, 3))
    st = _check(handle, x, 0.15, 6, prec, extra=1, blocks=64, seeds=16)
    assert st["n_collisions"] > 64


@pytest.mark.parametrize("prec", PRECS)
def test_cfg2_subset(handle, prec):
    x, info = dclust.synth_config(2, n=200000)
    _check(handle, x, info["eps"], info["minpts"], prec)


@pytest.mark.parametrize("prec", PRECS)
def test_cfg3_subset(handle, prec):
    x, info = dclust.synth_config(3, n=400000)
    _check(handle, x, info["eps"], info["minpts"], prec)


def test_batched_matches_oracle(handle):
    arrays, params, refs = [], [], []
    for i in range(40):
        x, info = dclust.synth_batch_instance(i)
        if x.shape[0] > 30000:
            continue
        arrays.append(x)
        params.append(dclust.make_params(x.shape[1], info["eps"], info["minpts"],
                                         precision=dclust.FP32_GUARD))
        refs.append(oracle.fast_dbscan(x, info["eps"], info["minpts"])[0])
    labs, errs = handle.run_batched(arrays, params)
    assert not np.any(errs)
    for i, (a, b) in enumerate(zip(labs, refs)):
        assert np.array_equal(a, b), f"instance {i}"


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dim", [2, 3])
def test_far_links_only(handle, prec, dim):
    # tight groups 0.97 eps apart along the first axis: consecutive groups
    # sit two cells apart, so only far cell pairs join them
    rng = np.random.default_rng(11 + dim)
    eps, mp = 0.1, 4
    groups = []
    for g in range(40):
        c = np.zeros(dim)
        c[0] = g * 0.97 * eps
        c[1:] = rng.uniform(0, 1, dim - 1) * 1e-3
        groups.append(c + rng.normal(0, 0.002 * eps, (6, dim)))
    x = np.concatenate(groups + [rng.uniform(-1, 5, (300, dim))])
    st = _check(handle, x, eps, mp, prec)
    assert st["diag"][11] > 0  # components joined by far pairs


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("minpts", [5, 60])
def test_heavy_cells(handle, prec, dim, minpts):
    # dense blobs: cells with far more than 32 (and 256) points, with and
    # without all-core cells
    rng = np.random.default_rng(21 + dim + minpts)
    eps = 0.05
    parts = [rng.normal(0, 0.3 * eps, (30000, dim)), rng.normal(0.6, 1.5 * eps, (20000, dim)),
             rng.uniform(-0.5, 1.5, (5000, dim))]
    x = np.concatenate(parts)
    st = _check(handle, x, eps, minpts, prec)
    assert st["diag"][2] > 0  # heavy cells present


def _batch(limit=60, max_n=40000):
    xs, params = [], []
    for i in range(limit):
        x, info = dclust.synth_batch_instance(i)
        if x.shape[0] > max_n:
            continue
        xs.append(x)
        params.append(dclust.make_params(x.shape[1], info["eps"], info["minpts"],
                                         precision=dclust.FP32_GUARD))
    return xs, params


def test_batched_equals_per_instance_and_autotune(handle, devmem):
    xs, params = _batch()
    assert len({x.shape[1] for x in xs}) == 2  # 2-D and 3-D instances in one batch
    ns = np.array([x.shape[0] for x in xs], dtype=np.int64)
    flat = np.concatenate([x.ravel() for x in xs])
    d_x = devmem.upload(flat)
    d_lab = devmem.alloc(int(ns.sum()) * 4)
    tuned = handle.autotune_device(d_x, ns, params)
    shape = (tuned[0].seed_list_size, tuned[0].blocks, tuned[0].block_threads)
    assert shape[0] in (256, 512, 1024, 2048) and shape[2] in (128, 256) and shape[1] > 0
    assert all((t.seed_list_size, t.blocks, t.block_threads) == shape for t in tuned)
    errs = handle.run_batched_device(d_x, ns, tuned, d_lab)
    assert not np.any(errs)
    lab = devmem.download(np.empty(int(ns.sum()), dtype=np.int32), d_lab)
    off = np.concatenate([[0], np.cumsum(ns)])
    for i, (x, p) in enumerate(zip(xs, params)):
        single = handle.run(x, p.eps, p.minpts, precision=dclust.FP32_GUARD)
        assert np.array_equal(lab[off[i]:off[i + 1]], single), f"instance {i}"


def test_autotune_remembers_shape(handle, devmem):
    x, info = dclust.synth_config(3, n=200000)
    p = dclust.make_params(3, info["eps"], info["minpts"], precision=dclust.FP32_GUARD)
    d_x = devmem.upload(np.ascontiguousarray(x))
    tuned = handle.autotune_device(d_x, [x.shape[0]], [p])[0]
    q = handle.autotune(x.shape[0], dclust.make_params(3, info["eps"], info["minpts"],
                                                       precision=dclust.FP32_GUARD))
    assert (q.seed_list_size, q.blocks, q.block_threads) == \
        (tuned.seed_list_size, tuned.blocks, tuned.block_threads)
    _check(handle, x, info["eps"], info["minpts"], dclust.FP32_GUARD,
           seeds=q.seed_list_size, blocks=q.blocks, threads=q.block_threads)


def test_neighbourhood_statistics_switch(handle, monkeypatch):
    x, info = dclust.synth_config(3, n=100000)
    handle.run(x, info["eps"], info["minpts"], precision=dclust.FP32_GUARD)
    assert handle.last_stats["cand_total"] == 0
    monkeypatch.setenv("DCLUST_NBR_STATS", "1")
    handle.run(x, info["eps"], info["minpts"], precision=dclust.FP32_GUARD)
    st = handle.last_stats
    assert st["cand_total"] >= st["n_points"] and st["nbr_cells_total"] >= st["n_cells"]


def _blob_sample(k, n, dim=3):
    r = np.random.default_rng(300 + k)
    centres = r.uniform(0, 1, (10, dim))
    x = centres[r.integers(0, 10, n)] + r.normal(0, 0.02, (n, dim))
    x[: n // 10] = r.uniform(0, 1, (n // 10, dim))
    return x


@pytest.mark.parametrize("prec", PRECS)
def test_graph_replay(handle, devmem, prec, monkeypatch):
    # repeated identical calls: the first runs eagerly, the second is captured
    # into a graph and replayed, later ones replay it; the coordinates change
    # between calls at the same device pointers, and a call of another size
    # in between drops the graph
    monkeypatch.delenv("DCLUST_GRAPH", raising=False)
    n, eps, mp = 30000, 0.02, 5
    p = dclust.make_params(3, eps, mp, precision=prec)
    d_x = devmem.alloc(n * 3 * 8)
    d_lab = devmem.alloc(n * 4)
    seq = ["other", 0, 0, 1, 2, "other", 2, 3, 0]
    replayed = []
    for step, k in enumerate(seq):
        if k == "other":
            _check(handle, _blob_sample(9, n // 2), eps, mp, prec)
            replayed.append(handle.last_stats["diag"][14])
            continue
        x = _blob_sample(k, n)
        devmem.h2d(d_x, x)
        st = handle.run_device(d_x, n, p, d_lab)
        lab = devmem.download(np.empty(n, dtype=np.int32), d_lab)
        ref, core, asym = oracle.fast_dbscan(x, eps, mp)
        assert np.array_equal(lab, ref), f"call {step}: {int(np.sum(lab != ref))} labels differ"
        assert np.array_equal(handle.core_flags(n), core), f"call {step}: core flags"
        assert st["n_clusters"] == int(ref.max())
        if st["diag"][14]:
            # phase events are nodes of the graph
            assert st["t_core_ms"] > 0 and st["t_total_ms"] >= st["t_core_ms"]
        replayed.append(st["diag"][14])
    # (the first call's outcome depends on the call before this test)
    assert replayed[1:] == [0, 1, 1, 1, 0, 0, 1, 1]


def test_graph_replay_phase_times(handle, devmem, monkeypatch):
    # a replay records only the run's and the neighbour-count kernel's events:
    # its other phase times read 0; DCLUST_PHASE_TIMES=1 (part of the launch
    # state, so the graph is captured again) records every phase; the labels
    # are the same either way
    monkeypatch.delenv("DCLUST_GRAPH", raising=False)
    monkeypatch.delenv("DCLUST_PHASE_TIMES", raising=False)
    n, eps, mp = 30000, 0.02, 5
    x = _blob_sample(5, n)
    ref, _, _ = oracle.fast_dbscan(x, eps, mp)
    p = dclust.make_params(3, eps, mp, precision=dclust.FP32_GUARD)
    d_x = devmem.alloc(n * 3 * 8)
    d_lab = devmem.alloc(n * 4)
    devmem.h2d(d_x, x)
    phases = ("t_rows_ms", "t_adj_ms", "t_expand_ms", "t_merge_ms", "t_far_ms", "t_border_ms",
              "t_relabel_ms")
    for env in (None, "1", None):
        if env is None:
            monkeypatch.delenv("DCLUST_PHASE_TIMES", raising=False)
        else:
            monkeypatch.setenv("DCLUST_PHASE_TIMES", env)
        # a call of another size first drops any graph left by earlier calls
        _check(handle, _blob_sample(9, n // 2), eps, mp, dclust.FP32_GUARD)
        sts = []
        for _ in range(3):
            st = dict(handle.run_device(d_x, n, p, d_lab))
            lab = devmem.download(np.empty(n, dtype=np.int32), d_lab)
            assert np.array_equal(lab, ref)
            sts.append(st)
        assert [s["diag"][14] for s in sts] == [0, 1, 1]  # eager, then captured and replayed
        last = sts[-1]
        assert last["t_core_ms"] > 0 and last["t_total_ms"] >= last["t_index_ms"] + last["t_core_ms"]
        if env is None:
            assert all(last[k] == 0 for k in phases) and last["diag"][4] == 0
        else:
            assert all(last[k] > 0 for k in phases)
            assert last["t_total_ms"] >= sum(last[k] for k in phases) + last["t_core_ms"]
        # the eager runs record every phase event
        assert all(sts[0][k] > 0 for k in phases)


def test_graph_replay_switch(handle, monkeypatch):
    monkeypatch.setenv("DCLUST_GRAPH", "0")
    x = _blob_sample(4, 12000, dim=2)
    for _ in range(3):
        _check(handle, x, 0.01, 5, dclust.FP32_GUARD)
        assert handle.last_stats["diag"][14] == 0


@pytest.mark.parametrize("prec", PRECS)
def test_graph_replay_reruns_when_passes_run_out(handle, prec, monkeypatch):
    # a graph holds one expansion pass more than the run before its capture
    # needed; data that needs more abandons the replay on the device and the
    # call runs eagerly; the next identical call captures a longer graph
    monkeypatch.delenv("DCLUST_GRAPH", raising=False)
    # every wave hands its list on after one seed batch: clustered data needs
    # many passes
    monkeypatch.setenv("DCLUST_EXPAND_BUDGET", "1")
    dense, info = dclust.synth_config(3, n=120000)
    eps, mp = info["eps"], info["minpts"]
    rng = np.random.default_rng(12)
    sparse = rng.uniform(0, 1000 * eps, dense.shape)  # isolated points: one pass
    _check(handle, sparse[: len(sparse) // 2], eps, mp, prec, blocks=16)
    out = []
    for x in [sparse, sparse, dense, dense, dense]:
        st = _check(handle, x, eps, mp, prec, blocks=16)
        out.append((st["diag"][14], st["n_passes"]))
    assert out[0] == (0, 1) and out[1] == (1, 1)
    assert out[2][0] == 0 and out[2][1] > 2
    longer = 1 if out[2][1] < 8 else 0
    assert out[3][0] == longer and out[4][0] == longer


def test_batched_instance_errors(handle):
    # instances with non-finite coordinates, or with a bin width below eps,
    # report their own error code; the others are clustered as if alone
    xs, params = _batch(limit=40, max_n=20000)
    assert len(xs) >= 5
    nan = xs[1].copy()
    nan[5, 0] = np.nan
    inf = xs[2].copy()
    inf[7, -1] = np.inf
    narrow = dclust.make_params(xs[4].shape[1], params[4].eps, params[4].minpts, r=10**6,
                                precision=dclust.FP32_GUARD)
    labs, errs = handle.run_batched([xs[0], nan, inf, xs[3], xs[4]],
                                    [params[0], params[1], params[2], params[3], narrow])
    assert errs.tolist() == [0, dclust.E_PARAM, dclust.E_PARAM, 0, dclust.E_BINWIDTH]
    for i in (0, 3):
        ref = oracle.fast_dbscan(xs[i], params[i].eps, params[i].minpts)[0]
        assert np.array_equal(labs[i], ref), f"instance {i}"


def test_graph_replay_non_finite_input(handle, devmem, monkeypatch):
    # a replay whose coordinates turn non-finite reports the error; the next
    # replay with finite coordinates is exact again
    monkeypatch.delenv("DCLUST_GRAPH", raising=False)
    n, eps, mp = 20000, 0.02, 5
    x = _blob_sample(5, n)
    p = dclust.make_params(3, eps, mp, precision=dclust.FP32_GUARD)
    d_x = devmem.alloc(n * 3 * 8)
    d_lab = devmem.alloc(n * 4)
    ref = oracle.fast_dbscan(x, eps, mp)[0]
    _check(handle, x[: n // 2], eps, mp, dclust.FP32_GUARD)
    bad = x.copy()
    bad[123, 2] = np.nan
    seen = []
    for step, data in enumerate([x, x, bad, x]):
        devmem.h2d(d_x, data)
        if data is bad:
            with pytest.raises(dclust.DclustError) as e:
                handle.run_device(d_x, n, p, d_lab)
            assert e.value.code == dclust.E_PARAM
            continue
        st = handle.run_device(d_x, n, p, d_lab)
        lab = devmem.download(np.empty(n, dtype=np.int32), d_lab)
        assert np.array_equal(lab, ref), f"call {step}"
        seen.append(st["diag"][14])
    assert seen == [0, 1, 1]


@pytest.mark.parametrize("prec", PRECS)
def test_no_line_table(handle, prec):
    # 2-D points spread over more than 2^27 cells along x: the dense line
    # table is not built and neighbour lines are found by binary search
    rng = np.random.default_rng(21)
    n = 6000
    centres = np.array([[0.0, 0.0], [40.0, 3.0], [5.0e7, -5.0], [5.0e7 + 30.0, 0.0]])
    x = centres[rng.integers(0, 4, n)] + rng.normal(0, 2.0, (n, 2))
    x[:300, 0] = rng.uniform(0, 5.0e7, 300)
    st = _check(handle, x, 0.6, 4, prec)
    assert st["n_cells"] > 1000


def test_back_to_back_runs_stable(devmem):
    # a race in the final union-find flatten once left inner nodes as
    # "roots" (a few points took another cluster's label, in about one of
    # four sweeps over these instances); it showed in long sequences of
    # different instances on one handle.  The 4096 batch instances are
    # clustered back to back three times: every sweep must give the same
    # labels, and a sample equals the oracle.
    from concurrent.futures import ThreadPoolExecutor

    import dclust as dc

    with ThreadPoolExecutor(max_workers=8) as ex:
        insts = list(ex.map(dc.synth_batch_instance, range(4096)))
    big = list(range(len(insts)))
    xs = [insts[i][0] for i in big]
    infos = [insts[i][1] for i in big]
    ns = np.array([x.shape[0] for x in xs], dtype=np.int64)
    flat = np.concatenate([x.ravel() for x in xs])
    off_el = np.concatenate([[0], np.cumsum(ns * np.array([x.shape[1] for x in xs]))])
    off_pt = np.concatenate([[0], np.cumsum(ns)])
    d_x = devmem.upload(flat)
    d_l = devmem.alloc(int(ns.sum()) * 4)
    params = [dc.make_params(x.shape[1], f["eps"], f["minpts"], precision=dc.FP32_GUARD)
              for x, f in zip(xs, infos)]
    h = dc.Handle(0)
    try:
        first = None
        for sweep in range(3):
            for k in range(len(xs)):
                h.run_device(d_x + int(off_el[k]) * 8, int(ns[k]), params[k],
                             d_l + int(off_pt[k]) * 4)
            lab = devmem.download(np.empty(int(ns.sum()), dtype=np.int32), d_l)
            if first is None:
                first = lab
                for k in range(0, len(xs), 256):
                    ref = oracle.fast_dbscan(xs[k], infos[k]["eps"], infos[k]["minpts"])[0]
                    assert np.array_equal(lab[off_pt[k]:off_pt[k + 1]], ref), f"instance {big[k]}"
            else:
                diff = [big[k] for k in range(len(xs))
                        if not np.array_equal(lab[off_pt[k]:off_pt[k + 1]],
                                              first[off_pt[k]:off_pt[k + 1]])]
                assert not diff, f"sweep {sweep}: instances {diff} changed"
    finally:
        h.close()


def _degenerate_cases():
    rng = np.random.default_rng(77)
    line = np.zeros((3000, 3))
    line[:, 0] = rng.uniform(0, 1, 3000)  # zero extent on two axes
    cases = {
        "single point": (np.array([[0.25, -3.0, 7.0]]), 0.1, 1),
        "single point, minpts 2": (np.array([[0.25, -3.0]]), 0.1, 2),
        "all identical": (np.full((5000, 3), 1.5), 0.01, 10),
        "identical, minpts above n": (np.full((40, 2), -2.0), 0.01, 41),
        "minpts 1 (every point core)": (rng.uniform(0, 1, (3000, 2)), 0.005, 1),
        "minpts above n": (rng.uniform(0, 1, (500, 3)), 0.2, 600),
        "collinear in 3-D": (line, 0.002, 4),
        "negative offset coordinates": (rng.normal(-1e4, 1.0, (6000, 2)), 0.05, 5),
        "eps covers everything": (rng.uniform(0, 1, (2000, 2)), 10.0, 3),
        "duplicates within clusters": (np.repeat(rng.uniform(0, 1, (800, 3)), 4, axis=0), 0.03, 6),
    }
    return cases


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("case", sorted(_degenerate_cases()))
def test_degenerate_inputs(handle, prec, case):
    x, eps, mp = _degenerate_cases()[case]
    _check(handle, x, eps, mp, prec)


def test_empty_input(handle):
    lab = handle.run(np.empty((0, 3)), 0.1, 4, precision=dclust.FP32_GUARD)
    assert lab.shape == (0,) and handle.last_stats["n_points"] == 0


@pytest.mark.parametrize("prec", PRECS)
def test_graph_replay_reruns_when_the_key_widens(handle, prec, monkeypatch):
    # a graph holds the radix passes of the key width it was captured after;
    # coordinates whose cell key needs more bits abandon the replay on the
    # device and the call runs eagerly; the next identical call recaptures
    monkeypatch.delenv("DCLUST_GRAPH", raising=False)
    rng = np.random.default_rng(44)
    n, eps, mp = 5000, 0.01, 4
    small = rng.uniform(0, 0.05, (n, 2))  # ~100 cells: one 8-bit pass
    big = rng.uniform(0, 10.0, (n, 2))    # ~4M cells: three passes
    _check(handle, small[: n // 2], eps, mp, prec)
    out = []
    for x in [small, small, big, big, small]:
        st = _check(handle, x, eps, mp, prec)
        out.append((st["diag"][14], st["key_bits"]))
    assert [r for r, _ in out] == [0, 1, 0, 1, 1]
    assert out[1][1] <= 8 and out[2][1] > 16


@pytest.mark.parametrize("prec", PRECS)
def test_graph_replay_reruns_when_the_key_passes_32_bits(handle, prec, monkeypatch):
    # keys of at most 32 bits are sorted as 32-bit keys; a replay of such a
    # graph whose cell keys need more bits is abandoned before its truncated
    # keys are used, and the call runs eagerly with 64-bit keys; a graph
    # captured for 64-bit keys then replays the narrower input too
    monkeypatch.delenv("DCLUST_GRAPH", raising=False)
    rng = np.random.default_rng(46)
    n, eps, mp = 4000, 0.001, 4
    centres = rng.uniform(0.05, 0.45, (8, 3))
    narrow = centres[rng.integers(0, 8, n)] + rng.normal(0, 0.002, (n, 3))
    narrow[0] = [0.0, 0.0, 0.0]
    narrow[1] = [0.5, 0.5, 0.5]  # extent 0.5: 1000 cells per axis, 30-bit keys
    wide = narrow.copy()
    wide[1] = [1.0, 1.0, 1.0]  # extent 1.0: 2000 cells per axis, 33-bit keys
    _check(handle, narrow[: n // 2], eps, mp, prec)
    out = []
    for x in [narrow, narrow, wide, wide, narrow]:
        st = _check(handle, x, eps, mp, prec)
        out.append((st["diag"][14], st["key_bits"]))
    assert [r for r, _ in out] == [0, 1, 0, 1, 1]
    assert out[0][1] <= 32 < out[2][1]


def test_point_count_limit(handle, devmem):
    # cell indices are packed into 28 bits: 2^28 points are refused on the
    # host, before any buffer is touched
    p = dclust.make_params(3, 0.01, 4)
    d_x, d_lab = devmem.alloc(64), devmem.alloc(64)
    for run in (handle.run_device, handle.run_onestep_device):
        with pytest.raises(dclust.DclustError) as e:
            run(d_x, 1 << 28, p, d_lab)
        assert e.value.code == dclust.E_CAPACITY


@pytest.mark.parametrize("prec", PRECS)
def test_axis_limit_3d(handle, prec):
    # a 3-D axis holds at most 2^21 cells of width eps/2: 5000 / 0.001 is more
    rng = np.random.default_rng(91)
    x = rng.uniform(0, 1, (2000, 3))
    x[0] = [5000.0, 0.5, 0.5]
    with pytest.raises(dclust.DclustError) as e:
        handle.run(x, 0.002, 4, precision=prec)
    assert e.value.code == dclust.E_CAPACITY


@pytest.mark.parametrize("prec", PRECS)
def test_wide_2d_extent(handle, prec):
    # two groups 10^6 apart with eps 10^-3: about 2*10^9 cells on x (below the
    # 2-D limit of 2^31), a cell key wider than 32 bits (five radix passes) and
    # no line table
    rng = np.random.default_rng(92)
    a = rng.normal(0.5, 0.01, (1500, 2))
    b = rng.normal(0.5, 0.01, (1500, 2)) + np.array([1.0e6, 0.0])
    x = np.concatenate([a, b])[rng.permutation(3000)]
    st = _check(handle, x, 0.001, 4, prec)
    assert st["key_bits"] > 32


@pytest.mark.parametrize("prec", PRECS)
def test_batched_ragged_with_empty_instances(handle, prec):
    # empty, one-point and two-point instances between ordinary ones: each
    # instance is clustered as if alone, the empty ones give empty labels
    rng = np.random.default_rng(93)
    xs, params = _batch(limit=12, max_n=20000)
    inst = [
        (np.empty((0, 2)), 0.1, 3),
        (np.array([[0.3, 0.4, 0.5]]), 0.1, 1),
        (xs[0], params[0].eps, params[0].minpts),
        (np.empty((0, 3)), 0.05, 4),
        (np.array([[0.0, 0.0], [5.0, 5.0]]), 0.5, 1),
        (rng.uniform(0, 1, (3000, 3)), 0.05, 5),
        (xs[1], params[1].eps, params[1].minpts),
    ]
    plist = [dclust.make_params(x.shape[1], eps, mp, precision=prec) for x, eps, mp in inst]
    labs, errs = handle.run_batched([x for x, _, _ in inst], plist)
    assert errs.tolist() == [0] * len(inst)
    for i, (x, eps, mp) in enumerate(inst):
        assert labs[i].shape == (x.shape[0],)
        if x.shape[0]:
            ref, _, asym = oracle.fast_dbscan(x, eps, mp)
            assert asym == 0 and np.array_equal(labs[i], ref), f"instance {i}"
    assert labs[1].tolist() == [1] and labs[4].tolist() == [1, 2]


def test_guard_band_decides_pairs_fp32_cannot_resolve(handle):
    # 1000 isolated pairs at distance eps*(1 + d), |d| <= 3e-8: below FP32's
    # resolution, so every pair falls in the guard band and is re-decided in
    # FP64; with minpts 2 a pair is a cluster exactly when the FP64 predicate
    # holds.  Some FP32 decisions must be flipped, and the labels equal the
    # oracle's (and the FP64 form's).
    rng = np.random.default_rng(94)
    eps = 1e-3
    g = np.stack(np.meshgrid(*[np.arange(10)] * 3, indexing="ij"), -1).reshape(-1, 3)
    a = 0.05 + g * 0.01
    u = rng.normal(size=(1000, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    d = rng.uniform(-3e-8, 3e-8, (1000, 1))
    b = a + u * eps * (1.0 + d)
    x = np.concatenate([a, b])[rng.permutation(2000)]
    ref, _, asym = oracle.fast_dbscan(x, eps, 2)
    assert asym == 0 and 200 < ref.max() < 800  # both outcomes occur
    lab32 = handle.run(x, eps, 2, precision=dclust.FP32_GUARD)
    st = handle.last_stats
    assert np.array_equal(lab32, ref)
    assert st["band_rechecks"] >= 1000 and st["band_flips"] > 0
    assert np.array_equal(handle.run(x, eps, 2, precision=dclust.FP64), ref)
