"""Synthetic workload generators: determinism and the stated mixtures."""
import hashlib

import numpy as np
import pytest

import dclust


def test_config_info():
    for cfg, (n, dim, eps, mp) in {1: (10000, 2, 1.0, 4), 2: (1000000, 2, 0.002, 4),
                                   3: (4000000, 3, 0.001425, 10),
                                   5: (32000000, 3, 0.001, 20)}.items():
        x, info = dclust.synth_config(cfg, n=1000)
        assert x.shape == (1000, dim) and info["eps"] == eps and info["minpts"] == mp


def test_deterministic(golden):
    a, _ = dclust.synth_config(1)
    b, _ = dclust.synth_config(1)
    assert np.array_equal(a, b)
    assert hashlib.sha256(a.tobytes()).digest() == golden["cfg1_sha256"].tobytes()
    c, _ = dclust.synth_config(1, seed=99)
    assert not np.array_equal(a, c)


@pytest.mark.parametrize("cfg", [2, 3, 5])
def test_unit_box_and_noise_share(cfg):
    x, info = dclust.synth_config(cfg, n=200000)
    assert x.min() >= 0.0 and x.max() <= 1.0


def test_cfg5_mixture_shares():
    x, info = dclust.synth_config(5, n=100000)
    assert x.shape == (100000, 3) and info["seed_list_size"] == 64


def test_batch_plan_distribution():
    ns, ds = dclust.synth_batch_plan(4096)
    assert ns.min() >= 200 and ns.max() <= 200000
    assert set(np.unique(ds).tolist()) == {2, 3}
    assert abs(np.mean(ds == 3) - 0.5) < 0.05
    # log-uniform: median near the geometric mean sqrt(200 * 2e5) ~ 6325
    assert 4500 < np.median(ns) < 9000


def test_batch_instance_matches_plan():
    ns, ds = dclust.synth_batch_plan(8)
    for i in range(8):
        x, info = dclust.synth_batch_instance(i)
        assert x.shape == (ns[i], ds[i]) and info["minpts"] in (4, 8, 16) and info["eps"] > 0
        y, _ = dclust.synth_batch_instance(i)
        assert np.array_equal(x, y)


def test_knn_median_exact_on_lattice():
    g = np.stack(np.meshgrid(np.arange(20.0), np.arange(20.0)), -1).reshape(-1, 2)
    # interior lattice point: 4 neighbours at 1, 4 at sqrt 2, then 4 at 2
    assert dclust.knn_median(g, k=4, sample=400) == 1.0
    assert dclust.knn_median(g, k=8, sample=400) == pytest.approx(np.sqrt(2.0))
