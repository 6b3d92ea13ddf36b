"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or the timed CPU baseline.  The product path
(fast-cuda-gpu-dbscan_amd/) never imports it.

  liboracle.so            C restatement of the reference's CPU paths (oracle.c)
  _ref/libref_oracle.so   the reference's own cpu-dbscan/dbscan_rtree_exp.cpp,
                          compiled in place from the reference tree (2-D only)
  _ref/ref_check          the reference's test-results/cluster-check.cpp,
                          compiled in place
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref_oracle.so")
REF_CHECK = os.path.join(HERE, "_ref", "ref_check")

_L = None
_R = None
_P = C.c_void_p


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", HERE, "-s", "liboracle.so"], check=True)
        L = C.CDLL(LIB)
        L.oracle_pred.argtypes = [_P, _P, C.c_int, C.c_double]
        L.oracle_seq_dbscan.argtypes = [_P, C.c_int64, C.c_int, C.c_double, C.c_int, _P]
        L.oracle_fast_dbscan.argtypes = [_P, C.c_int64, C.c_int, C.c_double, C.c_int, C.c_int,
                                         _P, _P, _P]
        L.oracle_cells_dbscan.argtypes = [_P, C.c_int64, C.c_int, C.c_double, C.c_int, C.c_int,
                                          _P, _P]
        L.oracle_sample_check.argtypes = [_P, C.c_int64, C.c_int, C.c_double, _P, _P, _P,
                                          C.c_int64, C.c_int, _P]
        L.oracle_brute_counts.argtypes = [_P, C.c_int64, C.c_int, C.c_double, _P]
        L.oracle_multicore_dbscan.argtypes = [_P, C.c_int64, C.c_int, C.c_double, C.c_int,
                                              C.c_int, _P]
        L.oracle_cluster_check.argtypes = [_P, _P, C.c_int64, _P]
        _L = L
    return _L


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib():
    global _R
    if _R is None:
        R = C.CDLL(REF_LIB)
        R.ref_rtree_dbscan.argtypes = [_P, C.c_int64, C.c_double, C.c_int, _P]
        _R = R
    return _R


def _a(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def pred(p, q, eps) -> bool:
    p, q = _a(p), _a(q)
    return bool(lib().oracle_pred(p.ctypes.data, q.ctypes.data, p.shape[0], eps))


def seq_dbscan(xyz, eps, minpts):
    x = _a(xyz)
    lab = np.empty(x.shape[0], dtype=np.int32)
    rc = lib().oracle_seq_dbscan(x.ctypes.data, x.shape[0], x.shape[1], eps, minpts,
                                 lab.ctypes.data)
    assert rc == 0, rc
    return lab


def fast_dbscan(xyz, eps, minpts, nthreads=0):
    """(labels, core flags, asymmetric core-core pairs)"""
    x = _a(xyz)
    lab = np.empty(x.shape[0], dtype=np.int32)
    core = np.empty(x.shape[0], dtype=np.uint8)
    asym = C.c_int64(0)
    rc = lib().oracle_fast_dbscan(x.ctypes.data, x.shape[0], x.shape[1], eps, minpts, nthreads,
                                  lab.ctypes.data, core.ctypes.data, C.byref(asym))
    assert rc == 0, rc
    return lab, core, asym.value


def cells_dbscan(xyz, eps, minpts, nthreads=0):
    """(labels, core flags): the labels of fast_dbscan, with the core-core
    unions done between clique sub-cells (dense inputs)."""
    x = _a(xyz)
    lab = np.empty(x.shape[0], dtype=np.int32)
    core = np.empty(x.shape[0], dtype=np.uint8)
    rc = lib().oracle_cells_dbscan(x.ctypes.data, x.shape[0], x.shape[1], eps, minpts, nthreads,
                                   lab.ctypes.data, core.ctypes.data)
    assert rc == 0, rc
    return lab, core


def sample_check(xyz, eps, labels, core, sample, nthreads=0):
    """Violations of the DBSCAN definition at the sampled points:
    (core-label violations, border violations, predicate evaluations,
    sampled core points)."""
    x = _a(xyz)
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    cf = np.ascontiguousarray(core, dtype=np.uint8)
    smp = np.ascontiguousarray(sample, dtype=np.int64)
    out = np.zeros(4, dtype=np.int64)
    rc = lib().oracle_sample_check(x.ctypes.data, x.shape[0], x.shape[1], eps, lab.ctypes.data,
                                   cf.ctypes.data, smp.ctypes.data, smp.shape[0], nthreads,
                                   out.ctypes.data)
    assert rc == 0, rc
    return tuple(int(v) for v in out)


def brute_counts(xyz, eps):
    x = _a(xyz)
    c = np.empty(x.shape[0], dtype=np.int32)
    lib().oracle_brute_counts(x.ctypes.data, x.shape[0], x.shape[1], eps, c.ctypes.data)
    return c


def multicore_dbscan(xyz, eps, minpts, nthreads=0):
    """(number of clusters reported, labels) of the multicore-cpu restatement;
    minpts counts the other points only, as that program does."""
    x = _a(xyz)
    lab = np.empty(x.shape[0], dtype=np.int32)
    rc = lib().oracle_multicore_dbscan(x.ctypes.data, x.shape[0], x.shape[1], eps, minpts,
                                       nthreads, lab.ctypes.data)
    if rc < 0:
        raise MemoryError("oracle_multicore_dbscan")
    return rc, lab


def cluster_check(c1, c2):
    """(mismatched noise, agreeing noise, fraction lost, final score)"""
    a = np.ascontiguousarray(c1, dtype=np.int32)
    b = np.ascontiguousarray(c2, dtype=np.int32)
    out = np.zeros(4, dtype=np.float64)
    rc = lib().oracle_cluster_check(a.ctypes.data, b.ctypes.data, a.shape[0], out.ctypes.data)
    assert rc == 0, rc
    return int(out[0]), int(out[1]), float(out[2]), float(out[3])


def ref_rtree_dbscan(xy, eps, minpts):
    """Labels from the reference's own compiled R-tree DBSCAN (2-D only)."""
    x = _a(xy)
    assert x.shape[1] == 2
    lab = np.empty(x.shape[0], dtype=np.int32)
    rc = ref_lib().ref_rtree_dbscan(x.ctypes.data, x.shape[0], eps, minpts, lab.ctypes.data)
    assert rc == 0
    return lab


def canonical(labels):
    """Relabel clusters by first appearance (noise stays 0): two labelings
    are the same partition iff their canonical forms are equal."""
    lab = np.asarray(labels)
    out = np.zeros_like(lab)
    m = {}
    for i, v in enumerate(lab.tolist()):
        if v == 0:
            continue
        if v not in m:
            m[v] = len(m) + 1
        out[i] = m[v]
    return out
