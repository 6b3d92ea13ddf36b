#!/usr/bin/env python3
"""Benchmark of the MI355X DBSCAN path (one JSON line on stdout).

Workload (default): synthetic config 3 -- 3-D, 4,000,000 points in [0,1]^3,
128 Gaussian blobs (sigma U[0.005,0.03], Zipf(1.1) sizes) + 3 % uniform
noise, eps 0.001425, minpts 10 -- the configuration the performance target
is stated on.  One step = one full clustering of the 4 M points (grid index,
neighbour counts, seed-list expansion, union-find merge, border, canonical
relabel), FP32 distances with FP64 guard-band rechecks, inputs already
resident in HBM, labels left in HBM.  value = points clustered per second
over the timed steps.

Also reported (not timed in `value`): the plain FP64 form's time on the same
input, label/core-flag parity against the FP64 CPU oracle, the neighbour-
count kernel's roofline fraction, the batched leg (4,096 instances in one
call against one call per instance), the one-step GPU baseline (the
CUDA-DClust design on a bounded spatial sample, with the main path timed on
the same sample), and two CPU baselines on bounded spatial samples of the
same points: the oracle's sequential FP64 DBSCAN (1 thread) and the
multicore-cpu restatement (up to 16 threads).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fast-cuda-gpu-dbscan_amd"))
import dclust  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

METRIC = ("Mpoints/s clustered, exact core labels vs FP64 oracle; % HBM peak; "
          "batched vs plain")

WORKLOADS = {
    "cfg3": (3, "cfg3: 3-D, 4,000,000 points in [0,1]^3, 128 Gaussian blobs "
                "(sigma U[0.005,0.03], Zipf 1.1) + 3% noise, eps 0.001425, minpts 10"),
    "cfg2": (2, "cfg2: 2-D, 1,000,000 points in [0,1]^2, 64 anisotropic blobs "
                "(Zipf 1.2) + 5% noise, eps 0.002, minpts 4"),
    "cfg1": (1, "cfg1: 2-D, 10,000 points, 8 blobs + 10% noise, eps 1.0, minpts 4, r 5"),
    "cfg5": (5, "cfg5: 3-D, 32,000,000 points, 16 dense blobs + 2 filaments + 1% noise, "
                "eps 0.001, minpts 20, seed list 64"),
}


class Hip:
    """Minimal HIP runtime access for device buffers (plumbing only)."""

    def __init__(self):
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                self.L = C.CDLL(name)
                break
            except OSError:
                continue
        L = self.L
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipSetDevice.argtypes = [C.c_int]

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: hipError {rc}")

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p.value

    def h2d(self, dst, arr):
        self.check(self.L.hipMemcpy(C.c_void_p(dst), arr.ctypes.data, arr.nbytes, 1), "hipMemcpy")

    def d2h(self, arr, src):
        self.check(self.L.hipMemcpy(arr.ctypes.data, C.c_void_p(src), arr.nbytes, 2), "hipMemcpy")

    def sync(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def free(self, p):
        self.L.hipFree(C.c_void_p(p))


def log(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def b_nbr(st, dim, bytes_per_coord):
    """Algorithmic bytes of one neighbour-count launch: every point read once
    as a query and once per neighbouring cell tile (sum over cells of n_c
    times the points of its 5^D stencil of half-eps cells), 8 bytes per
    stencil-cell range entry, 1 byte of core flag written per point, plus
    2*D*8 bytes of FP64 coordinates per guard-band recheck.  The sums come
    from one untimed run with DCLUST_NBR_STATS=1."""
    n = st["n_points"]
    return (bytes_per_coord * dim * (n + st["cand_total"]) + 8 * st["nbr_cells_total"] + n +
            16 * dim * st.get("band_rechecks_core", 0))


def measured_traffic(workload, precision):
    """HBM bytes per launch of k_core on this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated by a plain copy),
    or None when no such measurement is committed."""
    path = os.path.join(ROOT, "profiles", f"k_core_{workload}_{precision}_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def spatial_sample(x, m):
    """Points inside the sub-box [0, a]^D holding about m points (same local
    density as the full input)."""
    key = x.max(axis=1)
    a = np.partition(key, m - 1)[m - 1]
    return x[key <= a], float(a)


def cpu_baseline(x, info, budget_s=15.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # CPU baseline leg only

    eps, mp = info["eps"], info["minpts"]
    m = 20000
    while True:
        xs, a = spatial_sample(x, min(m, x.shape[0]))
        t0 = time.perf_counter()
        oracle.seq_dbscan(xs, eps, mp)
        dt = time.perf_counter() - t0
        if dt >= budget_s / 4 or xs.shape[0] >= x.shape[0]:
            break
        m = int(m * min(8.0, max(2.0, budget_s / 2 / max(dt, 1e-3))))
    return {"value": round(xs.shape[0] / dt / 1e6, 6), "unit": "Mpoints/s", "cores": 1,
            "kind": "port",
            "sample": (f"sequential FP64 DBSCAN restated from cpu-dbscan/dbscan_rtree_exp.cpp "
                       f"(oracle/oracle.c oracle_seq_dbscan), {xs.shape[0]} points of the same "
                       f"input inside the sub-box [0,{a:.4f}]^{x.shape[1]}, same eps/minpts, "
                       f"{dt:.2f} s")}


def host_cpu():
    """Model name and the cores this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "cpus_usable": len(os.sched_getaffinity(0)),
            "cpus_online": os.cpu_count()}


def cpu_baseline_multicore(x, info, budget_s=10.0):
    """multicore-cpu/dbscan.cpp restated (oracle_multicore_dbscan): FP32, one
    OpenMP scan over all points per expanded point, so O(N^2); timed on a
    sub-box sample with up to 16 threads (dbscan.cpp:115-117 uses 16)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # CPU baseline leg only

    threads = min(16, len(os.sched_getaffinity(0)))
    eps, mp = info["eps"], info["minpts"]
    m = 20000
    while True:
        xs, a = spatial_sample(x, min(m, x.shape[0]))
        t0 = time.perf_counter()
        k, _ = oracle.multicore_dbscan(xs, eps, mp - 1, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt >= budget_s / 4 or xs.shape[0] >= x.shape[0]:
            break
        m = int(m * min(8.0, max(1.5, (budget_s / 2 / max(dt, 1e-3)) ** 0.5)))
    n = x.shape[0]
    return {"value": round(xs.shape[0] / dt / 1e6, 6), "unit": "Mpoints/s", "cores": threads,
            "kind": "port",
            "sample": (f"multicore-cpu/dbscan.cpp restated (oracle/oracle.c "
                       f"oracle_multicore_dbscan: FP32, sqrt compare, minpts {mp - 1} others), "
                       f"{xs.shape[0]} points of the same input inside the sub-box "
                       f"[0,{a:.4f}]^{x.shape[1]}, {k} clusters reported, {dt:.2f} s"),
            "full_input_estimate_s": round(dt * (n / xs.shape[0]) ** 2, 1)}


def onestep_leg(hip, h, x, info, prec, budget_s=8.0):
    """The one-step baseline (CUDA-DClust design, dclust_run_onestep_device)
    and the main path on the same sub-box sample of the workload, grown until
    the baseline takes about budget_s / 4."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker only

    n, dim = x.shape
    m = 20000
    while True:
        xs, a = spatial_sample(x, min(m, n))
        ns = xs.shape[0]
        d_x, d_lab = hip.malloc(xs.nbytes), hip.malloc(ns * 4)
        hip.h2d(d_x, np.ascontiguousarray(xs))
        p1 = dclust.make_params(dim, info["eps"], info["minpts"], precision=dclust.FP64)
        hip.sync()
        t0 = time.perf_counter()
        st1 = h.run_onestep_device(d_x, ns, p1, d_lab)
        dt = time.perf_counter() - t0
        lab1 = np.empty(ns, dtype=np.int32)
        hip.d2h(lab1, d_lab)
        log(f"one-step baseline: {ns} points, {dt:.2f} s, {st1['n_passes']} launches")
        if dt >= budget_s / 4 or ns >= n:
            break
        hip.free(d_x)
        hip.free(d_lab)
        m = int(m * min(8.0, max(1.5, budget_s / 2 / max(dt, 1e-3))))
    p2 = dclust.make_params(dim, info["eps"], info["minpts"], r=info["partitions_r"],
                            seeds=info["seed_list_size"], precision=prec)
    for _ in range(2):
        h.run_device(d_x, ns, p2, d_lab)
    hip.sync()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        h.run_device(d_x, ns, p2, d_lab)
    hip.sync()
    main_s = (time.perf_counter() - t0) / reps
    ref, core, _ = oracle.fast_dbscan(xs, info["eps"], info["minpts"])
    core = core.astype(bool)
    hip.free(d_x)
    hip.free(d_lab)
    return {"design": "CUDA-DClust one-step expansion (cuda-dclust/dbscan_indexing.cu): one "
                      "seed per chain per launch, host read of the seed-list lengths after "
                      "every launch, FP64, 256 blocks x 256 threads, 128 seeds",
            "sample": f"{ns} points of the same input inside the sub-box [0,{a:.4f}]^{dim}",
            "onestep_ms": round(dt * 1e3, 2), "onestep_mpoints_s": round(ns / dt / 1e6, 4),
            "launches": st1["n_passes"], "chains": st1["n_chains"],
            "main_path_ms": round(main_s * 1e3, 4),
            "speedup_main_vs_onestep": round(dt / main_s, 1),
            "core_label_mismatches": int(np.sum(lab1[core] != ref[core])),
            "noise_mismatches": int(np.sum((lab1 == 0) != (ref == 0)))}


BATCH_DESC = ("cfg4: 4,096 independent 2-D/3-D instances, sizes log-uniform 200-200,000, "
              "2-32 Gaussian blobs + 5% noise each, eps 0.5-2x the median 10-NN distance, "
              "minpts in {4,8,16}, FP32 + guard band")


def batched_leg(hip, h, steps, warmup, check_oracle=True, count=4096):
    """One batched clustering of every instance (autotuned launch shape)
    against one plain single-instance call per instance (library default
    shape), on the same device-resident inputs; labels compared instance by
    instance, and a sample checked against the oracle."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=16) as ex:
        insts = list(ex.map(dclust.synth_batch_instance, range(count)))
    xs = [x for x, _ in insts]
    infos = [i for _, i in insts]
    ns = np.array([x.shape[0] for x in xs], dtype=np.int64)
    dims = np.array([x.shape[1] for x in xs], dtype=np.int64)
    plain = [dclust.make_params(i["dim"], i["eps"], i["minpts"], precision=dclust.FP32_GUARD)
             for i in infos]
    flat = np.concatenate([x.ravel() for x in xs])
    tot = int(ns.sum())
    d_x = hip.malloc(flat.nbytes)
    hip.h2d(d_x, flat)
    d_b = hip.malloc(tot * 4)
    d_1 = hip.malloc(tot * 4)
    t0 = time.perf_counter()
    tuned = h.autotune_device(d_x, ns, plain)
    t_tune = time.perf_counter() - t0
    for _ in range(warmup):
        h.run_batched_device(d_x, ns, tuned, d_b)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        errs = h.run_batched_device(d_x, ns, tuned, d_b)
    hip.sync()
    t_b = (time.perf_counter() - t0) / steps
    st_b = h.last_stats
    off_el = np.concatenate([[0], np.cumsum(ns * dims)])
    off_pt = np.concatenate([[0], np.cumsum(ns)])
    h.run_device(d_x, int(ns[0]), plain[0], d_1)  # warm-up
    hip.sync()
    t0 = time.perf_counter()
    for i in range(count):
        h.run_device(d_x + int(off_el[i]) * 8, int(ns[i]), plain[i], d_1 + int(off_pt[i]) * 4)
    hip.sync()
    t_pi = time.perf_counter() - t0
    lab_b = np.empty(tot, dtype=np.int32)
    lab_1 = np.empty(tot, dtype=np.int32)
    hip.d2h(lab_b, d_b)
    hip.d2h(lab_1, d_1)
    bad = [i for i in range(count)
           if not np.array_equal(lab_b[off_pt[i]:off_pt[i + 1]], lab_1[off_pt[i]:off_pt[i + 1]])]
    out = {"workload": BATCH_DESC, "instances": count, "points": tot,
           "batched_ms": round(t_b * 1e3, 3), "per_instance_ms": round(t_pi * 1e3, 3),
           "speedup_batched_vs_per_instance": round(t_pi / t_b, 2),
           "batched_mpoints_s": round(tot / t_b / 1e6, 2),
           "per_instance_mpoints_s": round(tot / t_pi / 1e6, 2),
           "instance_errors": int(np.count_nonzero(errs)),
           "instances_differing_batched_vs_per_instance": len(bad),
           "tuned_shape": {"seed_list_size": tuned[0].seed_list_size,
                           "blocks": tuned[0].blocks, "block_threads": tuned[0].block_threads},
           "autotune_s": round(t_tune, 2),
           "phases_ms": {k: round(st_b[k], 3) for k in st_b if k.startswith("t_")}}
    if check_oracle:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # checker only, untimed

        # instances where the two runs disagree: which one is wrong
        diff = []
        for i in bad[:8]:
            ref = oracle.fast_dbscan(xs[i], infos[i]["eps"], infos[i]["minpts"])[0]
            s = slice(off_pt[i], off_pt[i + 1])
            diff.append({"instance": i, "n": int(ns[i]), "dim": int(dims[i]),
                         "batched_vs_oracle": int(np.sum(lab_b[s] != ref)),
                         "per_instance_vs_oracle": int(np.sum(lab_1[s] != ref))})
        out["differing_detail"] = diff
        order = np.argsort(ns)
        sample = sorted(set(order[np.linspace(0, count - 1, 24).astype(int)].tolist()))
        mism = 0
        for i in sample:
            ref = oracle.fast_dbscan(xs[i], infos[i]["eps"], infos[i]["minpts"])[0]
            mism += int(np.sum(ref != lab_b[off_pt[i]:off_pt[i + 1]]))
        out["oracle_sample"] = {"instances": len(sample), "label_mismatches": mism}
    for ptr in (d_x, d_b, d_1):
        hip.free(ptr)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cfg3", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-vs-per-instance leg (cfg4)")
    ap.add_argument("--no-onestep", action="store_true",
                    help="skip the one-step (CUDA-DClust design) baseline leg")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a plain copy of known size, "
                         "used to calibrate the HBM byte counters of a profiler run)")
    args = ap.parse_args()
    ws = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or ws != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(how the path would shard is written up with the design notes).", file=sys.stderr)
        sys.exit(2)

    cfg, desc = WORKLOADS[args.workload]
    x, info = dclust.synth_config(cfg)
    n, dim = x.shape
    hip = Hip()
    hip.check(hip.L.hipSetDevice(0), "hipSetDevice")
    if args.calib_copy > 0:
        nb = args.calib_copy << 20
        src, dst = hip.malloc(nb), hip.malloc(nb)
        hip.check(hip.L.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), nb, 3), "hipMemcpy D2D")
        hip.sync()
        hip.free(src)
        hip.free(dst)
    d_x = hip.malloc(x.nbytes)
    d_lab = hip.malloc(n * 4)
    hip.h2d(d_x, x)
    prec = dclust.FP32_GUARD if args.precision == "fp32" else dclust.FP64
    p = dclust.make_params(dim, info["eps"], info["minpts"], r=info["partitions_r"],
                           seeds=info["seed_list_size"], precision=prec)
    h = dclust.Handle(0)

    def timed(params, steps, warmup):
        for _ in range(warmup):
            h.run_device(d_x, n, params, d_lab)
        hip.sync()
        sts = []
        t0 = time.perf_counter()
        for _ in range(steps):
            sts.append(h.run_device(d_x, n, params, d_lab))
        hip.sync()
        return time.perf_counter() - t0, sts

    # neighbourhood sizes for the algorithmic-byte count: one extra untimed run
    os.environ["DCLUST_NBR_STATS"] = "1"
    log(f"{args.workload}: {n} points generated")
    st_nbr = h.run_device(d_x, n, p, d_lab)
    del os.environ["DCLUST_NBR_STATS"]
    dt, sts = timed(p, args.steps, args.warmup)
    log(f"timed leg done: {dt / args.steps * 1e3:.3f} ms per clustering")
    st = dict(sts[-1])
    st["cand_total"] = st_nbr["cand_total"]
    st["nbr_cells_total"] = st_nbr["nbr_cells_total"]
    ms_step = dt / args.steps * 1e3
    value = n * args.steps / dt / 1e6
    t_core = float(np.mean([s["t_core_ms"] for s in sts]))
    phases = {k: round(float(np.mean([s[k] for s in sts])), 4)
              for k in ("t_index_ms", "t_core_ms", "t_adj_ms", "t_expand_ms", "t_merge_ms", "t_far_ms",
                        "t_border_ms", "t_relabel_ms", "t_total_ms")}
    bpc = 4 if prec == dclust.FP32_GUARD else 8
    bytes_core = b_nbr(st, dim, bpc)
    achieved = bytes_core / (t_core * 1e-3) / 1e9
    traffic = measured_traffic(args.workload, args.precision)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpoints/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == dclust.FP32_GUARD else "f64", "data": "synthetic",
        "config": {"workload": desc, "n_points": n, "dim": dim, "eps": info["eps"],
                   "minpts": info["minpts"],
                   "precision": "fp32+fp64 guard band" if bpc == 4 else "fp64"},
        "roofline": {"kernel": "k_core (neighbour count)", "bound": "hbm",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     # the HBM bytes the profiler measured per launch over this launch time:
                     # the stencil re-reads counted in `achieved` are served by the caches
                     "hbm_achieved": round(traffic / (t_core * 1e-3) / 1e9, 1) if traffic else None,
                     "hbm_frac": (round(traffic / (t_core * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                  if traffic else None),
                     "algorithmic_bytes": int(bytes_core), "avg_launch_ms": round(t_core, 5),
                     # pairs a scan of every point's 5^D half-eps stencil would test, and the
                     # distance evaluations the kernel makes, per second of its launch
                     "stencil_pairs_per_s": round(st["cand_total"] / (t_core * 1e-3), 1),
                     "distance_evals_per_s": round(st["pairs_evaluated"] / (t_core * 1e-3), 1)},
        "phases_ms": phases,
        "counters": {k: st[k] for k in ("n_cells", "n_core", "n_noise", "n_clusters", "n_chains",
                                        "n_collisions", "n_overflow", "n_passes",
                                        "band_rechecks", "band_flips", "asym_pairs",
                                        "pairs_evaluated", "cand_total", "key_bits",
                                        "rec_overflow", "diag")},
    }
    if not args.no_fp64_compare and prec == dclust.FP32_GUARD:
        p64 = dclust.make_params(dim, info["eps"], info["minpts"], r=info["partitions_r"],
                                 seeds=info["seed_list_size"], precision=dclust.FP64)
        dt64, s64 = timed(p64, max(2, args.steps // 2), 1)
        out["fp64_ms_per_step"] = round(dt64 / max(2, args.steps // 2) * 1e3, 4)
        out["speedup_fp32_guard_vs_fp64"] = round(out["fp64_ms_per_step"] / ms_step, 3)
        # leave the FP32 labels in the buffer for the parity check below
        h.run_device(d_x, n, p, d_lab)
        hip.sync()
    if not args.no_parity:
        log("parity check against the oracle")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # checker only, untimed

        if n <= 8_000_000:
            lab = np.empty(n, dtype=np.int32)
            hip.d2h(lab, d_lab)
            core = h.core_flags(n)
            xo, io = x, info
            what = "the timed input"
        else:
            # the oracle's union step over thousands of neighbours per point
            # does not finish in minutes at this size: a 1M-point draw of the
            # same generator, clustered by the same entry point, is checked
            xo, io = dclust.synth_config(cfg, n=1_000_000)
            lab = h.run(xo, io["eps"], io["minpts"], r=io["partitions_r"],
                        seeds=io["seed_list_size"], precision=prec)
            core = h.core_flags(xo.shape[0])
            what = "a 1,000,000-point draw of the same generator"
        ref, rcore, asym = oracle.fast_dbscan(xo, io["eps"], io["minpts"])
        out["parity"] = {"label_mismatches": int(np.sum(lab != ref)),
                         "core_mismatches": int(np.sum(core != rcore)),
                         "oracle_asym_pairs": int(asym), "checked_on": what}
    # the same clustering from host arrays (H2D + D2H included), one run
    t0 = time.perf_counter()
    h.run(x, info["eps"], info["minpts"], r=info["partitions_r"], seeds=info["seed_list_size"],
          precision=prec)
    out["pcie_inclusive_mpoints_s"] = round(n / (time.perf_counter() - t0) / 1e6, 2)
    if not args.no_batched:
        log("batched leg")
        out["batched"] = batched_leg(hip, h, max(2, args.steps // 2), 1,
                                     check_oracle=not args.no_parity)
    if not args.no_onestep:
        log("one-step (CUDA-DClust) baseline")
        out["onestep_baseline"] = onestep_leg(hip, h, x, info, prec)
    if not args.no_cpu_baseline:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(x, info)
        out["cpu_baseline"]["host"] = host_cpu()
        log("CPU baseline, multicore restatement")
        out["cpu_baseline_multicore"] = cpu_baseline_multicore(x, info)
    h.close()
    hip.free(d_x)
    hip.free(d_lab)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
