"""dclust_run_device_noalloc: the same clustering as dclust_run_device, in the
library's fixed per-GPU scratch array, with no device allocation.

Checked through the C ABI: labels and core flags bit-identical to
dclust_run_device on two calls in a row (the second smaller, so stale
contents of the shared array must not leak into it), both precision forms;
no change of the device's free memory across the calls on a fresh handle;
DCLUST_E_CAPACITY above the array's limit."""
import ctypes

import numpy as np
import pytest

import dclust
import oracle

pytestmark = pytest.mark.gpu


def _free_bytes():
    L = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert L.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _points(n, seed):
    x, info = dclust.synth_config(3, n=n, seed=seed)
    return x, info


@pytest.mark.parametrize("prec", [dclust.FP32_GUARD, dclust.FP64])
def test_noalloc_equals_run_device_twice(devmem, prec):
    print(f"[noalloc] precision {prec}", flush=True)
    with dclust.Handle(0) as h:
        for n, seed in ((120_000, 11), (30_000, 12)):  # the second call is smaller
            x, info = _points(n, seed)
            p = dclust.make_params(3, info["eps"], info["minpts"], precision=prec)
            d_x = devmem.upload(x)
            d_a = devmem.alloc(n * 4)
            d_b = devmem.alloc(n * 4)
            h.run_device(d_x, n, p, d_a)
            ref_core = h.core_flags(n)
            st = h.run_device_noalloc(d_x, n, p, d_b)
            got_core = h.core_flags(n)
            a = devmem.download(np.empty(n, dtype=np.int32), d_a)
            b = devmem.download(np.empty(n, dtype=np.int32), d_b)
            assert np.array_equal(a, b), f"n={n}: {int(np.sum(a != b))} labels differ"
            assert np.array_equal(ref_core, got_core)
            assert st["n_clusters"] == a.max() and st["diag"][14] == 0  # eager, no replay
            assert np.array_equal(b, oracle.fast_dbscan(x, info["eps"], info["minpts"])[0])


def test_noalloc_allocates_nothing(devmem):
    x, info = _points(100_000, 13)
    n = x.shape[0]
    p = dclust.make_params(3, info["eps"], info["minpts"], precision=dclust.FP32_GUARD)
    d_x = devmem.upload(x)
    d_b = devmem.alloc(n * 4)
    with dclust.Handle(0) as h:  # a fresh handle: its arena is empty
        # a first, small call: the runtime loads the code objects and sets up
        # the stream's queue on first use
        h.run_device_noalloc(d_x, 5_000, p, d_b)
        before = _free_bytes()
        # larger calls: a run that took its buffers from an arena would have
        # to grow it here
        for m in (50_000, n, n):
            h.run_device_noalloc(d_x, m, p, d_b)
        after = _free_bytes()
    assert after == before
    lab = devmem.download(np.empty(n, dtype=np.int32), d_b)
    assert np.array_equal(lab, oracle.fast_dbscan(x, info["eps"], info["minpts"])[0])


def test_noalloc_rejects_inputs_beyond_the_array(devmem):
    # all points at the origin: one all-core cell, a run of a few kernels
    n_max = 2_000_000
    d_x = devmem.upload(np.zeros((n_max, 3)))
    d_b = devmem.alloc(4 * n_max)
    p = dclust.make_params(3, 0.01, 10, precision=dclust.FP32_GUARD)
    with dclust.Handle(0) as h:
        with pytest.raises(dclust.DclustError) as e:
            h.run_device_noalloc(d_x, n_max, p, d_b)
        assert e.value.code == dclust.E_CAPACITY
        # the largest input the array takes (the size check precedes any launch)
        lo, hi = 1_000, n_max
        while hi - lo > 1_000:
            mid = (lo + hi) // 2
            try:
                st = h.run_device_noalloc(d_x, mid, p, d_b)
                assert st["n_clusters"] == 1 and st["n_core"] == mid
                lo = mid
            except dclust.DclustError as err:
                assert err.code == dclust.E_CAPACITY
                hi = mid
        print(f"[noalloc] largest 3-D FP32 input: {lo}-{hi} points", flush=True)
        assert lo > 150_000
        lab = devmem.download(np.empty(lo, dtype=np.int32), d_b)
        assert np.all(lab == 1)
