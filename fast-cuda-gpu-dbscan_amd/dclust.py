"""ctypes binding of libdclust.so (include/dclust.h) and libdclust_synth.so
(include/dclust_synth.h).

Mirrors the reference's surface for its CUDA-DClust+ path: the common.h
parameters (common.h:6-25) as keyword arguments, a row-major FP64 dataset in,
per-point labels out (0 = noise, 1..K).  No fallback: if the HIP library is
missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os
from collections.abc import Mapping

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libdclust.so")
SYNTH_PATH = os.path.join(_PKG, "libdclust_synth.so")

OK = 0
E_PARAM = -1
E_BINWIDTH = -2
E_IO = -3
E_HIP = -4
E_OOM = -5
E_CAPACITY = -6
E_NODEVICE = -7

FP64 = 0
FP32_GUARD = 1


class DclustError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = _lib().dclust_strerror(code).decode() if _LIB else str(code)
        super().__init__(f"{what}: {msg} ({code})" if what else f"{msg} ({code})")


class Params(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("minpts", C.c_int32),
        ("eps", C.c_double),
        ("partitions_r", C.c_int32),
        ("seed_list_size", C.c_int32),
        ("blocks", C.c_int32),
        ("block_threads", C.c_int32),
        ("extra_collision", C.c_int32),
        ("precision", C.c_int32),
    ]


_STAT_F64 = ["t_total_ms", "t_index_ms", "t_core_ms", "t_adj_ms", "t_expand_ms", "t_merge_ms", "t_far_ms",
             "t_border_ms", "t_relabel_ms"]
_STAT_I64 = ["n_points", "n_cells", "n_core", "n_noise", "n_clusters", "n_chains",
             "n_collisions", "n_overflow", "n_passes", "cand_total", "nbr_cells_total",
             "band_rechecks", "band_flips", "asym_pairs", "pairs_evaluated",
             "band_rechecks_core", "rec_overflow"]
_STAT_I32 = ["hip_error", "key_bits"]
_STAT_SHAPE = ["eff_blocks", "eff_block_threads", "eff_seed_list_size", "eff_extra_collision"]
_STAT_COUNTS = ["n_core_cells", "n_adj_points", "n_lines"]


class Stats(C.Structure):
    _fields_ = ([(k, C.c_double) for k in _STAT_F64] + [(k, C.c_int64) for k in _STAT_I64] +
                [(k, C.c_int32) for k in _STAT_I32] + [("diag", C.c_int64 * 16)] +
                [(k, C.c_int32) for k in _STAT_SHAPE] + [("t_rows_ms", C.c_double)] +
                [(k, C.c_int64) for k in _STAT_COUNTS] + [("work", C.c_int64 * 4)])

    def as_dict(self):
        out = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("diag", "work")}
        out["diag"] = list(self.diag)
        out["work"] = list(self.work)
        return out


_STATS_KEYS = tuple(k for k, _ in Stats._fields_)
_STATS_KEYSET = frozenset(_STATS_KEYS)


class StatsView(Mapping):
    """One run's dclust_stats as a read-only mapping: a field is converted
    when it is read (the arrays `diag` and `work` as lists), so a call in a
    tight loop does not pay for building a dict of every field."""

    __slots__ = ("_st",)

    def __init__(self, st):
        self._st = st

    def __getitem__(self, k):
        if k not in _STATS_KEYSET:
            raise KeyError(k)
        v = getattr(self._st, k)
        return list(v) if k in ("diag", "work") else v

    def __iter__(self):
        return iter(_STATS_KEYS)

    def __len__(self):
        return len(_STATS_KEYS)


class SynthInfo(C.Structure):
    _fields_ = [("dim", C.c_int32), ("minpts", C.c_int32), ("partitions_r", C.c_int32),
                ("seed_list_size", C.c_int32), ("eps", C.c_double), ("n", C.c_int64),
                ("seed", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIB = None
_SYN = None

_P = C.c_void_p


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.dclust_params_default.argtypes = [C.POINTER(Params)]
        L.dclust_create.argtypes = [C.POINTER(_P), C.c_int]
        L.dclust_destroy.argtypes = [_P]
        L.dclust_destroy.restype = None
        L.dclust_run.argtypes = [_P, _P, C.c_int64, C.POINTER(Params), _P, C.POINTER(Stats)]
        L.dclust_run_device.argtypes = [_P, _P, C.c_int64, C.POINTER(Params), _P,
                                        C.POINTER(Stats)]
        L.dclust_run_batched.argtypes = [_P, C.c_int32, _P, _P, _P, _P, _P, C.POINTER(Stats)]
        L.dclust_run_batched_device.argtypes = [_P, C.c_int32, _P, _P, _P, _P, _P,
                                                C.POINTER(Stats)]
        L.dclust_run_onestep_device.argtypes = [_P, _P, C.c_int64, C.POINTER(Params), _P,
                                                C.POINTER(Stats)]
        L.dclust_run_device_noalloc.argtypes = [_P, _P, C.c_int64, C.POINTER(Params), _P,
                                                C.POINTER(Stats)]
        L.dclust_last_core_flags.argtypes = [_P, _P, C.c_int64]
        L.dclust_autotune.argtypes = [_P, C.c_int64, C.POINTER(Params)]
        L.dclust_autotune_device.argtypes = [_P, C.c_int32, _P, _P, _P]
        L.dclust_import_dataset.argtypes = [C.c_char_p, C.c_int32, C.c_int64, _P,
                                            C.POINTER(C.c_int64)]
        L.dclust_write_labels.argtypes = [C.c_char_p, _P, C.c_int64]
        L.dclust_write_dataset_f64.argtypes = [C.c_char_p, _P, C.c_int64, C.c_int32]
        L.dclust_strerror.argtypes = [C.c_int]
        L.dclust_strerror.restype = C.c_char_p
        L.dclust_version.restype = C.c_char_p
        _LIB = L
    return _LIB


def _syn():
    global _SYN
    if _SYN is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"{SYNTH_PATH} is not built (run __graft_entry__.build())")
        S = C.CDLL(SYNTH_PATH)
        S.dclust_synth_config.argtypes = [C.c_int32, C.c_int64, C.c_uint64, _P,
                                          C.POINTER(SynthInfo)]
        S.dclust_synth_batch_plan.argtypes = [C.c_int32, C.c_uint64, _P, _P]
        S.dclust_synth_batch_instance.argtypes = [C.c_uint64, C.c_int32, _P,
                                                  C.POINTER(SynthInfo)]
        S.dclust_synth_knn_median.argtypes = [_P, C.c_int64, C.c_int32, C.c_int32, C.c_int64,
                                              C.c_uint64]
        S.dclust_synth_knn_median.restype = C.c_double
        S.dclust_synth_calibrate_eps.argtypes = [_P, C.c_int64, C.c_int32, C.c_int32,
                                                 C.c_double, C.c_int64, C.c_uint64]
        S.dclust_synth_calibrate_eps.restype = C.c_double
        _SYN = S
    return _SYN


def make_params(dim=2, eps=1.5, minpts=3, r=0, seeds=0, blocks=0, threads=0, extra=0,
                precision=FP64) -> Params:
    p = Params()
    _lib().dclust_params_default(C.byref(p))
    p.dim, p.eps, p.minpts, p.partitions_r = int(dim), float(eps), int(minpts), int(r)
    p.seed_list_size, p.blocks, p.block_threads = int(seeds), int(blocks), int(threads)
    p.extra_collision = int(extra)
    if isinstance(precision, str):
        precision = FP32_GUARD if precision.lower().startswith("fp32") else FP64
    p.precision = int(precision)
    return p


def default_params() -> Params:
    p = Params()
    _check(_lib().dclust_params_default(C.byref(p)), "dclust_params_default")
    return p


def _check(rc, what):
    if rc != OK:
        raise DclustError(rc, what)


def strerror(code: int) -> str:
    return _lib().dclust_strerror(code).decode()


def version() -> str:
    return _lib().dclust_version().decode()


class Handle:
    """One device context (dclust_create / dclust_destroy)."""

    def __init__(self, device: int = 0):
        self._h = _P()
        _check(_lib().dclust_create(C.byref(self._h), int(device)), "dclust_create")
        self.last_stats = None

    def close(self):
        if self._h:
            _lib().dclust_destroy(self._h)
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, xyz, eps, minpts, r=0, precision=FP64, seeds=0, blocks=0, threads=0,
            extra=0):
        """Cluster host points (n, dim) float64 -> int32 labels (0 = noise)."""
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        if x.ndim != 2:
            raise ValueError("xyz must be (n, dim)")
        n, dim = x.shape
        p = make_params(dim, eps, minpts, r, seeds, blocks, threads, extra, precision)
        lab = np.empty(n, dtype=np.int32)
        st = Stats()
        _check(_lib().dclust_run(self._h, x.ctypes.data, n, C.byref(p), lab.ctypes.data,
                                 C.byref(st)), "dclust_run")
        self.last_stats = st.as_dict()
        return lab

    def run_device(self, d_xyz: int, n: int, params: Params, d_labels: int):
        st = Stats()
        _check(_lib().dclust_run_device(self._h, C.c_void_p(d_xyz), int(n), C.byref(params),
                                        C.c_void_p(d_labels), C.byref(st)),
               "dclust_run_device")
        self.last_stats = StatsView(st)
        return self.last_stats

    def run_device_noalloc(self, d_xyz: int, n: int, params: Params, d_labels: int):
        """dclust_run_device_noalloc: the same run in the library's fixed
        per-GPU scratch array (no allocation)."""
        st = Stats()
        _check(_lib().dclust_run_device_noalloc(self._h, C.c_void_p(d_xyz), int(n),
                                                C.byref(params), C.c_void_p(d_labels),
                                                C.byref(st)),
               "dclust_run_device_noalloc")
        self.last_stats = st.as_dict()
        return self.last_stats

    def run_onestep_device(self, d_xyz: int, n: int, params: Params, d_labels: int):
        """The one-step (CUDA-DClust) baseline on device buffers."""
        st = Stats()
        _check(_lib().dclust_run_onestep_device(self._h, C.c_void_p(d_xyz), int(n),
                                                C.byref(params), C.c_void_p(d_labels),
                                                C.byref(st)),
               "dclust_run_onestep_device")
        self.last_stats = st.as_dict()
        return self.last_stats

    def run_batched(self, arrays, params_list):
        """Cluster independent instances in one pass; returns (labels list,
        per-instance error codes)."""
        k = len(arrays)
        xs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
        ns = np.array([a.shape[0] for a in xs], dtype=np.int64)
        ptrs = (C.c_void_p * k)(*[a.ctypes.data for a in xs])
        labs = [np.empty(a.shape[0], dtype=np.int32) for a in xs]
        lptr = (C.c_void_p * k)(*[a.ctypes.data for a in labs])
        parr = (Params * k)(*params_list)
        errs = np.zeros(k, dtype=np.int32)
        st = Stats()
        rc = _lib().dclust_run_batched(self._h, k, ptrs, ns.ctypes.data, parr, lptr,
                                       errs.ctypes.data, C.byref(st))
        if rc not in (OK,) and not (rc < 0 and np.any(errs != 0)):
            raise DclustError(rc, "dclust_run_batched")
        self.last_stats = st.as_dict()
        return labs, errs

    def run_batched_device(self, d_xyz: int, ns, params_list, d_labels: int):
        k = len(ns)
        ns = np.ascontiguousarray(ns, dtype=np.int64)
        parr = (Params * k)(*params_list)
        errs = np.zeros(k, dtype=np.int32)
        st = Stats()
        rc = _lib().dclust_run_batched_device(self._h, k, C.c_void_p(d_xyz), ns.ctypes.data,
                                              parr, C.c_void_p(d_labels), errs.ctypes.data,
                                              C.byref(st))
        if rc != OK and not np.any(errs != 0):
            raise DclustError(rc, "dclust_run_batched_device")
        self.last_stats = st.as_dict()
        return errs

    def core_flags(self, n: int):
        out = np.empty(n, dtype=np.uint8)
        _check(_lib().dclust_last_core_flags(self._h, out.ctypes.data, int(n)),
               "dclust_last_core_flags")
        return out

    def autotune(self, n: int, params: Params) -> Params:
        _check(_lib().dclust_autotune(self._h, int(n), C.byref(params)), "dclust_autotune")
        return params

    def autotune_device(self, d_xyz: int, ns, params_list):
        """Measure the candidate launch shapes on device-resident instances
        and return the params with the fastest shape filled in."""
        k = len(ns)
        ns = np.ascontiguousarray(ns, dtype=np.int64)
        parr = (Params * k)(*params_list)
        _check(_lib().dclust_autotune_device(self._h, k, C.c_void_p(d_xyz), ns.ctypes.data, parr),
               "dclust_autotune_device")
        return list(parr)


def import_dataset(path: str, dim: int, max_points: int = 0):
    n = C.c_int64(0)
    _check(_lib().dclust_import_dataset(path.encode(), dim, max_points, None, C.byref(n)),
           "dclust_import_dataset")
    out = np.empty((n.value, dim), dtype=np.float64)
    _check(_lib().dclust_import_dataset(path.encode(), dim, max_points, out.ctypes.data,
                                        C.byref(n)), "dclust_import_dataset")
    return out


def write_dataset_f64(path: str, xyz) -> None:
    """Binary dataset file (24-byte header, row-major doubles), read back by
    import_dataset."""
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    if x.ndim != 2:
        raise ValueError("xyz must be (n, dim)")
    _check(_lib().dclust_write_dataset_f64(path.encode(), x.ctypes.data, x.shape[0], x.shape[1]),
           "dclust_write_dataset_f64")


def write_labels(path: str, labels) -> None:
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    _check(_lib().dclust_write_labels(path.encode(), lab.ctypes.data, lab.shape[0]),
           "dclust_write_labels")


# --------------------------------------------------------------- synth ----
def synth_config(config: int, n: int = 0, seed: int | None = None):
    """(xyz (n, dim) float64, info dict) for synthetic config 1, 2, 3 or 5."""
    S = _syn()
    info = SynthInfo()
    sd = (1 << 64) - 1 if seed is None else int(seed)
    if S.dclust_synth_config(config, n, sd, None, C.byref(info)) != 0:
        raise ValueError(f"unknown config {config}")
    x = np.empty((info.n, info.dim), dtype=np.float64)
    if S.dclust_synth_config(config, n, sd, x.ctypes.data, C.byref(info)) != 0:
        raise RuntimeError("synth failed")
    return x, info.as_dict()


def synth_batch_plan(count: int, seed: int = 3):
    ns = np.empty(count, dtype=np.int64)
    ds = np.empty(count, dtype=np.int32)
    _syn().dclust_synth_batch_plan(count, seed, ns.ctypes.data, ds.ctypes.data)
    return ns, ds


def synth_batch_instance(i: int, seed: int = 3):
    S = _syn()
    info = SynthInfo()
    S.dclust_synth_batch_instance(seed, i, None, C.byref(info))
    x = np.empty((info.n, info.dim), dtype=np.float64)
    S.dclust_synth_batch_instance(seed, i, x.ctypes.data, C.byref(info))
    return x, info.as_dict()


def knn_median(x, k=10, sample=1000, seed=0) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return _syn().dclust_synth_knn_median(x.ctypes.data, x.shape[0], x.shape[1], k, sample, seed)
