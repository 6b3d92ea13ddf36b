"""Every synthetic configuration at its stated size, through the C ABI.

  config 2   1,000,000 2-D points, eps 0.002, minpts 4
  config 3   4,000,000 3-D points, eps 0.001425, minpts 10
  config 5  32,000,000 3-D points (dense blobs + filaments), eps 0.001,
            minpts 20, seed list 64 (many list overflows: the correction
            queue and the union-find merge carry the clustering)
  config 4   4,096 instances (113.7 M points) in one batched call

Bar: labels and core flags byte-identical to the FP64 oracle in both
precision forms.  The labels are those of cpu-dbscan/dbscan_rtree_exp.cpp:360-406
(cluster = rank of the smallest core index, border = smallest adjacent id,
noise 0).  Config 5 is checked against oracle.cells_dbscan (pinned equal to
oracle.fast_dbscan in tests/test_oracle.py) and, independently of any oracle
labels, at 20,000 sampled points against the definition.
"""
import numpy as np
import pytest

import dclust
import oracle

pytestmark = pytest.mark.gpu

PRECS = [dclust.FP64, dclust.FP32_GUARD]


def _say(msg):
    print(f"[fullsize] {msg}", flush=True)


@pytest.fixture(scope="module")
def cfg2():
    x, info = dclust.synth_config(2)
    ref, core, asym = oracle.fast_dbscan(x, info["eps"], info["minpts"])
    assert asym == 0
    return x, info, ref, core


@pytest.fixture(scope="module")
def cfg3():
    x, info = dclust.synth_config(3)
    ref, core, asym = oracle.fast_dbscan(x, info["eps"], info["minpts"])
    assert asym == 0
    return x, info, ref, core


def _run_full(handle, x, info, prec):
    _say(f"n={x.shape[0]} dim={x.shape[1]} precision={prec}")
    lab = handle.run(x, info["eps"], info["minpts"], r=info["partitions_r"],
                     seeds=info["seed_list_size"], precision=prec)
    return lab, handle.core_flags(x.shape[0]), handle.last_stats


def _canonical_numbering(lab, core):
    """Cluster ids 1..K appear in order of their smallest core index."""
    first = lab[np.nonzero(core)[0]]
    _, idx = np.unique(first, return_index=True)
    order = first[np.sort(idx)]
    return np.array_equal(order, np.arange(1, order.size + 1))


@pytest.mark.parametrize("prec", PRECS)
def test_cfg2_full_size(handle, cfg2, prec):
    x, info, ref, core = cfg2
    assert x.shape == (1_000_000, 2)
    lab, got_core, st = _run_full(handle, x, info, prec)
    assert int(np.sum(got_core != core)) == 0
    assert int(np.sum(lab != ref)) == 0
    assert st["n_clusters"] == ref.max() and st["asym_pairs"] == 0


@pytest.mark.parametrize("prec", PRECS)
def test_cfg3_full_size(handle, cfg3, prec):
    x, info, ref, core = cfg3
    assert x.shape == (4_000_000, 3)
    lab, got_core, st = _run_full(handle, x, info, prec)
    assert int(np.sum(got_core != core)) == 0
    assert int(np.sum(lab != ref)) == 0
    assert st["n_clusters"] == ref.max() and st["asym_pairs"] == 0
    if prec == dclust.FP32_GUARD:
        assert st["band_rechecks"] > 0


def test_cfg5_full_size(handle):
    x, info = dclust.synth_config(5)
    assert x.shape == (32_000_000, 3) and info["seed_list_size"] == 64
    eps, mp = info["eps"], info["minpts"]
    lab64, core64, st64 = _run_full(handle, x, info, dclust.FP64)
    lab32, core32, st32 = _run_full(handle, x, info, dclust.FP32_GUARD)
    # the small seed lists overflow: the correction queue is exercised
    assert st64["n_overflow"] > 0 and st32["n_overflow"] > 0
    _say("oracle (clique sub-cells)")
    ref, rcore = oracle.cells_dbscan(x, eps, mp)
    assert int(np.sum(core64 != rcore)) == 0
    assert int(np.sum(core32 != rcore)) == 0
    assert int(np.sum(lab64 != ref)) == 0
    assert int(np.sum(lab32 != ref)) == 0
    assert _canonical_numbering(lab64, core64)
    # the definition at sampled points: 10,000 uniform, 10,000 non-core
    _say("sampled eps-ball checks")
    rng = np.random.default_rng(5)
    nonc = np.nonzero(core64 == 0)[0]
    smp = np.concatenate([rng.choice(x.shape[0], 10_000, replace=False),
                          rng.choice(nonc, min(10_000, nonc.size), replace=False)])
    bad_core, bad_border, evals, ncore = oracle.sample_check(x, eps, lab64, core64, smp)
    assert ncore > 5_000 and evals > 10 ** 8
    assert (bad_core, bad_border) == (0, 0)


def test_cfg4_full_batch(handle, devmem):
    # all 4,096 instances in one batched call (FP32 + guard band) against one
    # plain call per instance, and a spread of instances against the oracle
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=16) as ex:
        insts = list(ex.map(dclust.synth_batch_instance, range(4096)))
    xs = [x for x, _ in insts]
    infos = [i for _, i in insts]
    ns = np.array([x.shape[0] for x in xs], dtype=np.int64)
    assert len(xs) == 4096 and {x.shape[1] for x in xs} == {2, 3}
    params = [dclust.make_params(i["dim"], i["eps"], i["minpts"], precision=dclust.FP32_GUARD)
              for i in infos]
    tot = int(ns.sum())
    _say(f"{len(xs)} instances, {tot} points")
    d_x = devmem.upload(np.concatenate([x.ravel() for x in xs]))
    d_b = devmem.alloc(tot * 4)
    d_1 = devmem.alloc(tot * 4)
    errs = handle.run_batched_device(d_x, ns, params, d_b)
    assert not np.any(errs)
    off_el = np.concatenate([[0], np.cumsum(ns * np.array([x.shape[1] for x in xs]))])
    off_pt = np.concatenate([[0], np.cumsum(ns)])
    for i in range(len(xs)):
        handle.run_device(d_x + int(off_el[i]) * 8, int(ns[i]), params[i],
                          d_1 + int(off_pt[i]) * 4)
    lab_b = devmem.download(np.empty(tot, dtype=np.int32), d_b)
    lab_1 = devmem.download(np.empty(tot, dtype=np.int32), d_1)
    differ = [i for i in range(len(xs))
              if not np.array_equal(lab_b[off_pt[i]:off_pt[i + 1]], lab_1[off_pt[i]:off_pt[i + 1]])]
    assert not differ, differ[:10]
    order = np.argsort(ns)
    sample = sorted(set(order[np.linspace(0, len(xs) - 1, 64).astype(int)].tolist()))
    _say(f"oracle on {len(sample)} instances")
    for i in sample:
        ref, _ = oracle.cells_dbscan(xs[i], infos[i]["eps"], infos[i]["minpts"])
        assert np.array_equal(lab_b[off_pt[i]:off_pt[i + 1]], ref), f"instance {i}"
