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


def stencil_bytes(st, dim, bytes_per_coord):
    """Bytes a scan of every point's whole 5^D stencil of half-eps cells
    would read: every point once as a query and once per neighbouring cell
    tile (sum over cells of n_c times the points of its stencil), 8 bytes per
    stencil-cell range entry, 1 byte of core flag per point, plus 2*D*8 bytes
    of FP64 coordinates per guard-band recheck.  The neighbour-count kernel
    stops at minpts and skips all-core cells, so it reads far less: this is a
    size of the search problem, not the kernel's traffic.  The sums come from
    one untimed run with DCLUST_NBR_STATS=1."""
    n = st["n_points"]
    return (bytes_per_coord * dim * (n + st["cand_total"]) + 8 * st["nbr_cells_total"] + n +
            16 * dim * st.get("band_rechecks_core", 0))


def kernel_bytes(st, dim, f32):
    """Algorithmic bytes of one launch of each neighbour kernel: the work the
    kernel reported for this run (stats.work, pairs_evaluated, diag) times the
    bytes one unit of that work reads or writes, with no credit for caches.

    k_core   per point: cell_of 4, its cell's two starts 8, its coordinates
             s*D, core flag 1; per line group 5 packed rows (20); per
             non-empty line its point range (two starts, 8); per candidate
             its coordinates s*D; per cell the core count 4 and mask clear 8;
             per listed core point of a light cell the list entry 4.
    k_adj    per lane (core point of a light cell): list entry 4, cell_of 4,
             cell_z 4, fixed-point coordinates 4*D (+ FP64 8*D in the FP64
             form), the packed rows of the near lines above the cell (5 in
             3-D, 2 in 2-D); per non-empty near line its first start 4; per
             cell on it: next start, cell_z, core count (12); per pair test:
             core flag 1 + coordinates s*D; per core cell the mask 8; per
             adjacency bit found (half the directed bits, diag[3] / 2, exact
             when no cell is heavy) the mirror: the row of each cell 8 and
             the other cell's mask 8.
    k_rows   per cell and group of 5 lines: cell_xyz 8, the cell's sort key
             staged 8, five packed rows written 20 (the line-table entries
             of a tile's span and the keys of the extra lines around it are
             not counted: a lower bound).
    k_expand per popped cell: chain label 4, adjacency mask 8, 9 near rows
             36; per adjacency bit: the neighbour's label 4 and its claim 4;
             per cell the seeding scan: core count 4 and label 4.
    s = 4 (fixed point, FP32 form) or 8 (FP64 form)."""
    n, c = st["n_points"], st["n_cells"]
    s = 4 if f32 else 8
    w = st["work"]
    lines_all = 25 if dim == 3 else 5
    upper = 5 if dim == 3 else 2
    groups = lines_all // 5
    a = st["n_adj_points"]
    dg = st["diag"]
    core = (n * (4 + 8 + s * dim + 1) + 20 * w[0] + 8 * w[1] + s * dim * st["pairs_evaluated"] +
            c * (4 + 8) + 4 * a)
    lane = 12 + 4 * dim + (0 if f32 else 8 * dim) + 4 * upper
    adj = (a * lane + 4 * w[2] + 12 * w[3] + dg[8] * (1 + s * dim) + 8 * st["n_core_cells"] +
           16 * (dg[3] // 2))
    rows = c * groups * (8 + 8 + 20)
    expand = dg[0] * (4 + 8 + 36) + dg[3] * 8 + c * 8
    return {"k_core": core, "k_adj": adj, "k_rows": rows, "k_expand": expand}


KERNEL_TIME = {"k_core": "t_core_ms", "k_adj": "t_adj_ms", "k_rows": "t_rows_ms",
               "k_expand": "t_expand_ms"}


def measured_traffic(workload, precision):
    """HBM bytes per launch of each profiled kernel on this workload (rocprofv3
    FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by what the counter
    reads for a plain device-to-device copy of known size), from the
    committed summary, or {} when none is committed."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}_{precision}.json")
    try:
        with open(path) as f:
            return {k: v["traffic_bytes"] for k, v in json.load(f)["kernels"].items()}
    except (OSError, KeyError, ValueError, TypeError):
        return {}


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


def batched_leg(hip, h, steps, warmup, check_oracle=True, count=4096, online_tune=False):
    """One batched clustering of every instance against one plain
    single-instance call per instance, on the same device-resident inputs;
    labels compared instance by instance, and a sample checked against the
    oracle.  Both use the library's built-in shape table; with online_tune
    the batch runs with the shape dclust_autotune_device measures instead."""
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
    t_tune = None
    tuned = plain
    if online_tune:
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
           "shape": {"seed_list_size": st_b["eff_seed_list_size"], "blocks": st_b["eff_blocks"],
                     "block_threads": st_b["eff_block_threads"],
                     "source": "online sweep" if online_tune else "built-in table"},
           "autotune_s": round(t_tune, 2) if t_tune is not None else None,
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


def _short(name):
    """Kernel name without namespace, template and arguments (template
    arguments are dropped first: they may hold spaces)."""
    n = name.split("(")[0].split("<")[0].split(" ")[-1]
    return n.split("::")[-1]


def summarize_profile(path, skip_runs, traffic=None):
    """Markdown summary of rocprofv3 CSV output under `path` (one profiled
    bench command per subdirectory or file prefix): per kernel, its time and
    counters per clustering run, averaged over the runs after the first
    `skip_runs` (a run starts at its k_init_head dispatch).  FETCH_SIZE and
    WRITE_SIZE are divided by the factor the same counter shows for the
    1 GiB device-to-device copy that --calib-copy 1024 makes before the runs."""
    import csv
    import glob

    out = []
    for f in sorted(glob.glob(os.path.join(path, "**", "*_kernel_trace.csv"), recursive=True)):
        rows = sorted(csv.DictReader(open(f)), key=lambda r: int(r["Start_Timestamp"]))
        runs, cur = [], None
        for r in rows:
            k = _short(r["Kernel_Name"])
            if k == "k_init_head":
                cur = {}
                runs.append(cur)
            if cur is None:
                continue
            d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            t, c = cur.get(k, (0, 0))
            cur[k] = (t + d, c + 1)
        timed = runs[skip_runs:]
        tot = {}
        for run in timed:
            for k, (t, c) in run.items():
                a = tot.setdefault(k, [0, 0, []])
                a[0] += t
                a[1] += c
                a[2].append(t)
        out.append(f"### Kernel times per clustering run: `{os.path.relpath(f, path)}`\n")
        out.append(f"{len(runs)} runs in the trace, the first {skip_runs} left out, "
                   f"{len(timed)} averaged.\n")
        out.append("| Kernel | dispatches / run | avg us / run | min us | max us | share |")
        out.append("|---|---:|---:|---:|---:|---:|")
        whole = sum(v[0] for v in tot.values())
        for k, (t, c, per) in sorted(tot.items(), key=lambda kv: -kv[1][0]):
            out.append(f"| `{k}` | {c / len(timed):.2f} | {t / len(timed) / 1e3:.1f} | "
                       f"{min(per) / 1e3:.1f} | {max(per) / 1e3:.1f} | {100 * t / whole:.1f} % |")
        out.append(f"\nSum of kernel time per run: {whole / len(timed) / 1e3:.1f} us\n")
    for f in sorted(glob.glob(os.path.join(path, "**", "*_counter_collection.csv"),
                              recursive=True)):
        rows = list(csv.DictReader(open(f)))
        disp = {}
        for r in rows:
            e = disp.setdefault(int(r["Dispatch_Id"]), {"k": _short(r["Kernel_Name"]), "c": {}})
            e["c"][r["Counter_Name"]] = e["c"].get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
        order = sorted(disp)
        calib, runs, cur = {}, [], None
        for i in order:
            e = disp[i]
            if e["k"] == "k_init_head":
                cur = {}
                runs.append(cur)
            if cur is None:
                if "copyBuffer" in e["k"]:
                    for cn, v in e["c"].items():
                        calib[cn] = calib.get(cn, 0.0) + v
                continue
            a = cur.setdefault(e["k"], {})
            for cn, v in e["c"].items():
                a[cn] = a.get(cn, 0.0) + v
        timed = runs[skip_runs:]
        names = sorted({cn for run in timed for a in run.values() for cn in a})
        factor = {cn: calib[cn] / (1 << 20) for cn in ("FETCH_SIZE", "WRITE_SIZE")
                  if cn in calib and calib[cn] > 0}
        out.append(f"### Counters per clustering run: `{os.path.relpath(f, path)}`\n")
        out.append(f"{len(runs)} runs, the first {skip_runs} left out, {len(timed)} averaged.")
        for cn, fa in factor.items():
            out.append(f"{cn}: the 1 GiB copy reads {calib[cn]:.0f} KiB, factor {fa:.4f}; "
                       f"MB below = KiB / factor * 1024 / 1e6.")
        out.append("")
        hdr = "| Kernel | " + " | ".join(names) + (" | HBM MB |" if factor else " |")
        out.append(hdr)
        out.append("|---|" + "---:|" * (len(names) + (1 if factor else 0)))
        avg = {}
        for run in timed:
            for k, a in run.items():
                b = avg.setdefault(k, {})
                for cn, v in a.items():
                    b[cn] = b.get(cn, 0.0) + v / len(timed)
        # ratio metrics (percentages, occupancy, conflicts per access) are
        # per dispatch: averaged over a run's dispatches instead of summed
        ndisp = {}
        for i in order:
            e = disp[i]
            ndisp[e["k"]] = ndisp.get(e["k"], 0) + 1
        ratio = [cn for cn in names if not (cn.endswith("_SIZE") or cn.endswith("_sum"))]
        for k, b in avg.items():
            per_run = ndisp.get(k, 1) / max(1, len(runs))
            for cn in ratio:
                if cn in b:
                    b[cn] /= max(1.0, per_run)
        key = names[0] if names else None
        for k, b in sorted(avg.items(), key=lambda kv: -kv[1].get(key, 0.0)):
            cells = [f"{b.get(cn, 0.0):.4g}" for cn in names]
            mb = sum(b.get(cn, 0.0) / fa * 1024 / 1e6 for cn, fa in factor.items())
            out.append(f"| `{k}` | " + " | ".join(cells) + (f" | {mb:.1f} |" if factor else " |"))
            if traffic is not None:
                for cn, fa in factor.items():
                    traffic.setdefault(k, {})[cn] = b.get(cn, 0.0) / fa * 1024
        out.append("")
    return "\n".join(out)


def tune_table(hip, h, reps=10):
    """Measures the launch shapes behind the library's built-in shape table:
    dclust_autotune_device's sweep (seed-list length x workgroup size x
    workgroups per CU) on representative problems of each dimension mix and
    size class, each best shape timed against the library's default shape;
    then clusterings of the same points at several PARTITION_SIZE values
    (r only validates the bin width: the grid is fixed by eps)."""
    rows = []

    def timed(d_x, ns, params, d_l):
        for _ in range(2):
            h.run_batched_device(d_x, ns, params, d_l)
        hip.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            h.run_batched_device(d_x, ns, params, d_l)
        hip.sync()
        return (time.perf_counter() - t0) / reps * 1e3

    cases = [("cfg1", 1, 0), ("cfg2 100k", 2, 100_000), ("cfg2", 2, 0), ("cfg3 100k", 3, 100_000),
             ("cfg3 400k", 3, 400_000), ("cfg3", 3, 0), ("cfg5 4M", 5, 4_000_000), ("cfg5", 5, 0),
             ("cfg4 first 256", 4, 256), ("cfg4", 4, 4096)]
    for name, cfg, n in cases:
        if cfg == 4:
            xs, infos = zip(*[dclust.synth_batch_instance(i) for i in range(n)])
        else:
            x, info = dclust.synth_config(cfg, n=n)
            xs, infos = [x], [info]
        ns = np.array([x.shape[0] for x in xs], dtype=np.int64)
        flat = np.concatenate([x.ravel() for x in xs])
        d_x, d_l = hip.malloc(flat.nbytes), hip.malloc(int(ns.sum()) * 4)
        hip.h2d(d_x, flat)
        base = [dclust.make_params(i["dim"], i["eps"], i["minpts"], precision=dclust.FP32_GUARD)
                for i in infos]
        t_def = timed(d_x, ns, base, d_l)
        tuned = h.autotune_device(d_x, ns, base)
        t_best = timed(d_x, ns, tuned, d_l)
        dims = sorted({int(i["dim"]) for i in infos})
        row = {"case": name, "dims": dims, "points": int(ns.sum()),
               "log2_points": int(np.floor(np.log2(max(1, ns.sum())))),
               "best": {"seed_list_size": tuned[0].seed_list_size,
                        "block_threads": tuned[0].block_threads, "blocks": tuned[0].blocks},
               "ms_default_shape": round(t_def, 4), "ms_best_shape": round(t_best, 4)}
        log(f"tune {row}")
        rows.append(row)
        hip.free(d_x)
        hip.free(d_l)
    # PARTITION_SIZE: the same clustering at several r (it validates the bin
    # width only; the grid is fixed by eps)
    for cfg, rs in ((1, (0, 1, 2, 5)), (3, (0, 1, 10, 100, 500))):
        x, info = dclust.synth_config(cfg)
        d_x, d_l = hip.malloc(x.nbytes), hip.malloc(x.shape[0] * 4)
        hip.h2d(d_x, x)
        ref = None
        for r in rs:
            p = dclust.make_params(x.shape[1], info["eps"], info["minpts"], r=r,
                                   precision=dclust.FP32_GUARD)
            t = timed(d_x, np.array([x.shape[0]], dtype=np.int64), [p], d_l)
            lab = np.empty(x.shape[0], dtype=np.int32)
            hip.d2h(lab, d_l)
            ref = lab if ref is None else ref
            row = {"case": f"cfg{cfg} r={r}", "ms": round(t, 4),
                   "labels_equal_r0": bool(np.array_equal(lab, ref))}
            log(f"r {row}")
            rows.append(row)
        hip.free(d_x)
        hip.free(d_l)
    return rows


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
    ap.add_argument("--profile-run", action="store_true",
                    help="the timed leg only (no neighbourhood-size run, no FP64 comparison, "
                         "parity, batched, one-step or CPU legs): the command profiled "
                         "under rocprofv3")
    ap.add_argument("--summarize-profile", metavar="DIR",
                    help="print a markdown summary of the rocprofv3 CSV output under DIR "
                         "(per clustering run, warm-up runs left out) and exit")
    ap.add_argument("--traffic-out", metavar="FILE",
                    help="with --summarize-profile: write the per-kernel HBM bytes "
                         "(FETCH_SIZE + WRITE_SIZE, calibrated) as JSON")
    ap.add_argument("--skip-runs", type=int, default=1,
                    help="runs left out of --summarize-profile (the eager warm-up)")
    ap.add_argument("--online-tune", action="store_true",
                    help="batched leg: run dclust_autotune_device's sweep instead of the "
                         "built-in shape table")
    ap.add_argument("--tune-table", action="store_true",
                    help="measure the launch shapes of the built-in shape table and the "
                         "effect of PARTITION_SIZE, print them as JSON and exit")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a plain copy of known size, "
                         "used to calibrate the HBM byte counters of a profiler run)")
    args = ap.parse_args()
    if args.summarize_profile:
        traffic = {} if args.traffic_out else None
        print(summarize_profile(args.summarize_profile, args.skip_runs, traffic))
        if args.traffic_out:
            # HBM bytes per launch of the kernels whose FETCH_SIZE and WRITE_SIZE
            # were both collected (the bench's roofline `traffic`)
            ks = {k: {"traffic_bytes": int(v["FETCH_SIZE"] + v["WRITE_SIZE"]),
                      "fetch_bytes": int(v["FETCH_SIZE"]), "write_bytes": int(v["WRITE_SIZE"])}
                  for k, v in traffic.items() if "FETCH_SIZE" in v and "WRITE_SIZE" in v}
            with open(args.traffic_out, "w") as f:
                json.dump({"source": args.summarize_profile, "kernels": ks}, f, indent=1,
                          sort_keys=True)
        return
    if args.profile_run:
        args.no_cpu_baseline = args.no_parity = args.no_fp64_compare = True
        args.no_batched = args.no_onestep = True
    ws = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or ws != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it "
              "(how the path would shard is written up with the design notes).", file=sys.stderr)
        sys.exit(2)

    if args.tune_table:
        hip = Hip()
        hip.check(hip.L.hipSetDevice(0), "hipSetDevice")
        with dclust.Handle(0) as h:
            print(json.dumps(tune_table(hip, h)), flush=True)
        return
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

    log(f"{args.workload}: {n} points generated")
    st_nbr = None
    if not args.profile_run:
        # neighbourhood sizes for the stencil-byte figure: one extra untimed
        # run (its k_rows makes extra loads, so it is kept out of profiles)
        os.environ["DCLUST_NBR_STATS"] = "1"
        st_nbr = h.run_device(d_x, n, p, d_lab)
        del os.environ["DCLUST_NBR_STATS"]
    dt, sts = timed(p, args.steps, args.warmup)
    log(f"timed leg done: {dt / args.steps * 1e3:.3f} ms per clustering")
    st = dict(sts[-1])
    ms_step = dt / args.steps * 1e3
    value = n * args.steps / dt / 1e6
    # The timed leg's replays record only the events around the whole run and
    # the neighbour-count kernel (every event inside a graph idles the device
    # for a few microseconds); an untimed leg with every phase event gives the
    # phase table and the other kernels' times.
    sts_ph = sts
    if not args.profile_run:
        os.environ["DCLUST_PHASE_TIMES"] = "1"
        _, sts_ph = timed(p, max(3, args.steps // 4), 1)
        del os.environ["DCLUST_PHASE_TIMES"]
    phases = {k: round(float(np.mean([s[k] for s in sts_ph])), 4)
              for k in ("t_index_ms", "t_rows_ms", "t_core_ms", "t_adj_ms", "t_expand_ms",
                        "t_merge_ms", "t_far_ms", "t_border_ms", "t_relabel_ms", "t_total_ms")}
    f32 = prec == dclust.FP32_GUARD
    kb = kernel_bytes(st, dim, f32)
    traffic = measured_traffic(args.workload, args.precision)
    kernels = {}
    for k, tk in KERNEL_TIME.items():
        # live HIP events on the library's stream: k_core's from the timed leg
        t_k = float(np.mean([s[tk] for s in (sts if k == "k_core" else sts_ph)]))
        if t_k <= 0:
            continue  # (a profile run has no phase leg)
        kernels[k] = {"avg_launch_ms": round(t_k, 5), "algorithmic_bytes": int(kb[k]),
                      "achieved": round(kb[k] / (t_k * 1e-3) / 1e9, 1),
                      "frac": round(kb[k] / (t_k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "traffic": traffic.get(k),
                      "hbm_frac": (round(traffic[k] / (t_k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                   if traffic.get(k) else None)}
    dom = max(kernels, key=lambda k: kernels[k]["avg_launch_ms"])
    kd = kernels[dom]
    t_core = kernels["k_core"]["avg_launch_ms"]
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpoints/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if prec == dclust.FP32_GUARD else "f64", "data": "synthetic",
        "config": {"workload": desc, "n_points": n, "dim": dim, "eps": info["eps"],
                   "minpts": info["minpts"],
                   "precision": "fp32+fp64 guard band" if f32 else "fp64"},
        # the longest single kernel of the run (live HIP events): its algorithmic
        # bytes from the work it did, and the HBM bytes rocprofv3 measured
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": kd["achieved"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kd["frac"],
                     "traffic": kd["traffic"], "hbm_frac": kd["hbm_frac"],
                     "algorithmic_bytes": kd["algorithmic_bytes"],
                     "avg_launch_ms": kd["avg_launch_ms"]},
        "kernels": kernels,
        # distance evaluations of the neighbour-count kernel per second of its launch
        "distance_evals_per_s": round(st["pairs_evaluated"] / (t_core * 1e-3), 1),
        "phases_ms": phases,
        "counters": {k: st[k] for k in ("n_cells", "n_core", "n_noise", "n_clusters", "n_chains",
                                        "n_collisions", "n_overflow", "n_passes",
                                        "band_rechecks", "band_flips", "asym_pairs",
                                        "pairs_evaluated", "key_bits", "rec_overflow",
                                        "n_core_cells", "n_adj_points", "n_lines", "work",
                                        "diag")},
    }
    if st_nbr is not None:
        sb = stencil_bytes(dict(st, cand_total=st_nbr["cand_total"],
                                nbr_cells_total=st_nbr["nbr_cells_total"]), dim, 4 if f32 else 8)
        out["stencil_search"] = {"stencil_bytes": int(sb), "stencil_pairs": st_nbr["cand_total"],
                                 "stencil_pairs_per_s": round(st_nbr["cand_total"] /
                                                              (t_core * 1e-3), 1)}
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

        lab = np.empty(n, dtype=np.int32)
        hip.d2h(lab, d_lab)
        core = h.core_flags(n)
        if n <= 8_000_000:
            ref, rcore, asym = oracle.fast_dbscan(x, info["eps"], info["minpts"])
            how = "oracle.fast_dbscan"
        else:
            # dense inputs (thousands of neighbours per point): the same labels
            # with the unions done between clique sub-cells
            ref, rcore = oracle.cells_dbscan(x, info["eps"], info["minpts"])
            asym = None
            how = "oracle.cells_dbscan"
        out["parity"] = {"label_mismatches": int(np.sum(lab != ref)),
                         "core_mismatches": int(np.sum(core != rcore)),
                         "oracle_asym_pairs": None if asym is None else int(asym),
                         "checked_on": f"the timed input ({how})"}
    # the same clustering from host arrays (H2D + D2H included): two untimed
    # calls (the first call of a new launch state runs eagerly, the second
    # captures its graph), then the mean of three
    def host_call():
        h.run(x, info["eps"], info["minpts"], r=info["partitions_r"],
              seeds=info["seed_list_size"], precision=prec)

    for _ in range(2):
        host_call()
    t0 = time.perf_counter()
    for _ in range(3):
        host_call()
    out["pcie_inclusive_mpoints_s"] = round(3 * n / (time.perf_counter() - t0) / 1e6, 2)
    if not args.no_batched:
        log("batched leg")
        out["batched"] = batched_leg(hip, h, max(2, args.steps // 2), 1,
                                     check_oracle=not args.no_parity,
                                     online_tune=args.online_tune)
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
