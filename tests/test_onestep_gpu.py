"""The one-step (CUDA-DClust) baseline, dclust_run_onestep_device, against
the CPU oracle.

Bar: core points and noise labelled exactly as the oracle labels them
(cluster id = rank of the smallest core input index); every border point in
the cluster of a core point whose query reaches it (the baseline does not
pick the smallest such id).  Border validity is checked with the oracle's
predicate restated in numpy: closed box around the core point, then the
squared distance summed left to right against fl(eps*eps).
"""
import numpy as np
import pytest

import dclust
import oracle

pytestmark = pytest.mark.gpu

EX = np.array([[1, 1], [2, 2], [3, 3], [10, 10], [11, 11], [12, 12], [20, 20], [21, 21],
               [22, 22]], dtype=np.float64)


def _onestep(handle, devmem, x, eps, minpts, seeds=0, blocks=0, threads=0):
    n, dim = x.shape
    d_x = devmem.upload(np.ascontiguousarray(x))
    d_lab = devmem.alloc(max(n, 1) * 4)
    p = dclust.make_params(dim, eps, minpts, 0, seeds, blocks, threads, 0, dclust.FP64)
    print(f"onestep n={n} dim={dim} eps={eps} minpts={minpts} seeds={seeds} blocks={blocks} "
          f"threads={threads}", flush=True)
    st = handle.run_onestep_device(d_x, n, p, d_lab)
    return devmem.download(np.empty(n, dtype=np.int32), d_lab), st


def _reach(x, q, p, eps):
    """P(q -> p) of the oracle for core points q (rows) and one point p."""
    lo, hi = x[q] - eps, x[q] + eps
    box = np.all((lo <= x[p]) & (x[p] <= hi), axis=1)
    d = x[q] - x[p]
    s = d[:, 0] * d[:, 0]
    for k in range(1, x.shape[1]):
        s = s + d[:, k] * d[:, k]
    return box & (s <= eps * eps)


def _verify(x, eps, minpts, lab, st):
    ref, core, asym = oracle.fast_dbscan(x, eps, minpts)
    assert asym == 0
    core = core.astype(bool)
    assert np.array_equal(lab[core], ref[core]), "core labels differ"
    assert np.array_equal(lab == 0, ref == 0), "noise differs"
    cq = np.flatnonzero(core)
    for p in np.flatnonzero(~core & (ref != 0)):
        r = _reach(x, cq, p, eps)
        assert lab[p] in set(ref[cq[r]].tolist()), f"border point {p}: label {lab[p]}"
    assert st["n_core"] == int(core.sum())
    assert st["n_clusters"] == int(ref.max(initial=0))
    assert st["n_passes"] > 0


def test_onestep_ex_txt(handle, devmem):
    lab, st = _onestep(handle, devmem, EX, 1.5, 3)
    assert lab.tolist() == [1, 1, 1, 2, 2, 2, 3, 3, 3]
    assert st["n_core"] == 3 and st["n_clusters"] == 3


def test_onestep_cfg1(handle, devmem):
    x, info = dclust.synth_config(1)
    lab, st = _onestep(handle, devmem, x, info["eps"], info["minpts"])
    _verify(x, info["eps"], info["minpts"], lab, st)


@pytest.mark.parametrize("dim,seed", [(2, 0), (3, 1)])
def test_onestep_random_blobs(handle, devmem, dim, seed):
    rng = np.random.default_rng(300 + seed)
    n = 20000
    centres = rng.uniform(0, 1, (12, dim))
    x = centres[rng.integers(0, 12, n)] + rng.normal(0, 0.02, (n, dim))
    x[: n // 10] = rng.uniform(0, 1, (n // 10, dim))
    eps = 0.01 if dim == 2 else 0.02
    lab, st = _onestep(handle, devmem, x, eps, 5)
    _verify(x, eps, 5, lab, st)


@pytest.mark.parametrize("seeds,blocks,threads", [(2, 4, 64), (1, 64, 1024), (128, 512, 128)])
def test_onestep_full_lists_and_shapes(handle, devmem, seeds, blocks, threads):
    # lists of 1-2 slots overflow at once: the points left unclaimed are
    # picked up by later refills
    rng = np.random.default_rng(7)
    n = 3000
    x = rng.normal(0, 1, (n, 2)) * np.array([1.0, 0.3])
    x[:200] = rng.uniform(-4, 4, (200, 2))
    lab, st = _onestep(handle, devmem, x, 0.12, 6, seeds=seeds, blocks=blocks, threads=threads)
    _verify(x, 0.12, 6, lab, st)


def test_onestep_thread_counts(handle, devmem):
    # THREAD_COUNT is rounded up to whole wavefronts (100 -> 128); beyond
    # 1024 threads the call is refused
    lab, st = _onestep(handle, devmem, EX, 1.5, 3, threads=100)
    assert st["eff_block_threads"] == 128
    assert lab.tolist() == [1, 1, 1, 2, 2, 2, 3, 3, 3]
    with pytest.raises(dclust.DclustError) as e:
        _onestep(handle, devmem, EX, 1.5, 3, threads=1025)
    assert e.value.code == dclust.E_PARAM


def test_onestep_contacts_folded_during_the_run(handle, devmem, monkeypatch):
    # a 4096-entry contact list fills many times over: the host folds it into
    # its union-find whenever half is used, and the labels stay exact (32
    # chains keep one launch's contacts below the 2048 entries left)
    monkeypatch.setenv("DCLUST_ONESTEP_CONTACTS", "4096")
    rng = np.random.default_rng(21)
    n = 20000
    centres = rng.uniform(0, 1, (6, 3))
    x = centres[rng.integers(0, 6, n)] + rng.normal(0, 0.03, (n, 3))
    lab, st = _onestep(handle, devmem, x, 0.02, 5, seeds=4, blocks=32)
    assert st["n_collisions"] > 4096
    _verify(x, 0.02, 5, lab, st)
