"""The drop-in boundary on the GPU: the CLI run end to end, the reference's
own launch values (common.h:13-25), the cell-key space limit, and the FP64
predicate's rounding on pairs where a fused multiply-add would decide
differently.  Every call goes through the C ABI (libdclust.so)."""
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

import dclust
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(os.path.dirname(dclust.__file__), "bin", "dclust")
EX_TXT = "1,1\n2,2\n3,3\n10,10\n11,11\n12,12\n20,20\n21,21\n22,22\n"


def _cli(args, timeout=120):
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _labels_file(path):
    return np.loadtxt(path, dtype=np.int64, ndmin=1).astype(np.int32)


def _ref_check(a_path, b_path):
    """The reference's test-results/cluster-check.cpp (compiled in place under
    oracle/_ref) on two label files, else its restatement."""
    if os.path.exists(oracle.REF_CHECK):
        out = subprocess.run([oracle.REF_CHECK, str(a_path), str(b_path)], capture_output=True,
                             text=True, check=True, timeout=120).stdout
        line = [s for s in out.splitlines() if s.startswith("Final Error metric:")]
        return float(line[0].split(":")[1])
    return oracle.cluster_check(_labels_file(a_path), _labels_file(b_path))[3]


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cli_ex_txt_with_common_h_defaults(tmp_path, golden, prec):
    # main.exe <dataset-file> (main.cu:35-46) with the common.h parameters:
    # DIMENSION 2, EPS 1.5, MINPTS 3, PARTITION_SIZE 3, 8 blocks x 16 threads
    src = tmp_path / "ex.txt"
    src.write_text(EX_TXT)
    out = tmp_path / "labels.txt"
    text = _cli([str(src), "--precision", prec, "--out", str(out)])
    assert "Final cluster after merging: 3" in text and "Number of noises: 0" in text
    lab = _labels_file(out)
    assert lab.tolist() == golden["ex_labels"].tolist() == [1, 1, 1, 2, 2, 2, 3, 3, 3]
    assert _ref_check(out, out) == 1.0


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_cli_cfg1_label_file_matches_reference_labels(tmp_path, golden, prec):
    # the 10,000-point config-1 set as a CSV dataset file; the label file the
    # CLI writes equals the labels of the reference's own R-tree DBSCAN
    # (tests/golden) and scores 1 in the reference's checker
    x, info = dclust.synth_config(1)
    src = tmp_path / "cfg1.txt"
    src.write_text("".join(f"{a!r},{b!r}\n" for a, b in x.tolist()))
    out = tmp_path / "labels.txt"
    text = _cli([str(src), "--eps", repr(info["eps"]), "--minpts", str(info["minpts"]),
                 "--r", str(info["partitions_r"]), "--precision", prec, "--out", str(out)])
    ref = golden["cfg1_labels"].astype(np.int32)
    lab = _labels_file(out)
    assert np.array_equal(lab, ref)
    assert f"Final cluster after merging: {ref.max()}" in text
    gold = tmp_path / "golden.txt"
    np.savetxt(gold, ref, fmt="%d")
    assert _ref_check(out, gold) == 1.0
    assert oracle.cluster_check(lab, ref)[3] == 1.0


def test_cli_rejects_bin_width_below_eps(tmp_path):
    # main.cu:204-207 exits when extent / r < eps; the CLI reports the error
    src = tmp_path / "ex.txt"
    src.write_text(EX_TXT)
    r = subprocess.run([EXE, str(src), "--r", "20"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "eps" in r.stderr


def test_reference_launch_values_run_exactly(handle):
    # dclust_params_default carries common.h:13-17 (8 blocks, THREAD_COUNT 16,
    # 128 seeds, 256 extra collisions); the run uses one wavefront per block
    x, info = dclust.synth_config(3, n=60000)
    p = dclust.default_params()
    assert (p.blocks, p.block_threads, p.seed_list_size, p.extra_collision) == (8, 16, 128, 256)
    lab = handle.run(x, info["eps"], info["minpts"], seeds=p.seed_list_size, blocks=p.blocks,
                     threads=p.block_threads, extra=p.extra_collision)
    st = handle.last_stats
    assert (st["eff_blocks"], st["eff_block_threads"], st["eff_seed_list_size"],
            st["eff_extra_collision"]) == (8, 64, 128, 256)
    ref, core, _ = oracle.fast_dbscan(x, info["eps"], info["minpts"])
    assert np.array_equal(lab, ref)
    assert np.array_equal(handle.core_flags(x.shape[0]), core)


@pytest.mark.parametrize("threads,eff", [(1, 64), (16, 64), (100, 128), (200, 256),
                                         (300, 256), (1024, 256), (0, 256)])
def test_thread_count_rounds_up_to_wavefronts(handle, threads, eff):
    rng = np.random.default_rng(threads)
    x = rng.normal(0, 1, (6000, 2))
    lab = handle.run(x, 0.08, 4, threads=threads, blocks=16)
    assert handle.last_stats["eff_block_threads"] == eff
    assert np.array_equal(lab, oracle.fast_dbscan(x, 0.08, 4)[0])


def test_thread_count_above_1024_is_rejected(handle):
    with pytest.raises(dclust.DclustError) as e:
        handle.run(np.zeros((4, 2)), 1.0, 2, threads=1025)
    assert e.value.code == dclust.E_PARAM


def test_onestep_thread_count_rounds_up(handle, devmem):
    rng = np.random.default_rng(5)
    x = rng.normal(0, 1, (3000, 2))
    d_x = devmem.upload(x)
    d_l = devmem.alloc(x.shape[0] * 4)
    st = handle.run_onestep_device(d_x, x.shape[0], dclust.make_params(2, 0.1, 4, threads=16,
                                                                       blocks=32), d_l)
    assert st["eff_block_threads"] == 64 and st["eff_blocks"] == 32
    lab = devmem.download(np.empty(x.shape[0], dtype=np.int32), d_l)
    ref, core, _ = oracle.fast_dbscan(x, 0.1, 4)
    assert np.array_equal(lab == 0, ref == 0)
    assert np.array_equal(lab[core == 1], ref[core == 1])


# ---- cell-key space: at most 2^63 cells over all instances of a call ----
# 3-D, unit cube, eps 1e-6: cell side 5e-7, 2,000,001 cells per axis
# (below the 2^21 axis limit), 8.0e18 cells per instance: one instance fits
# the 63-bit key, two do not.

def _cube(seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (50, 3))
    x[0] = 0.0
    x[1] = 1.0
    return x


def test_one_instance_at_the_key_space_limit_runs(handle):
    x = _cube(1)
    lab = handle.run(x, 1e-6, 2)
    assert handle.last_stats["key_bits"] == 63
    assert np.array_equal(lab, np.zeros(50, dtype=np.int32))


def test_batch_beyond_the_key_space_fails_whole_call(handle):
    xs = [_cube(1), _cube(2), np.array([[0.0, 0.0, 0.0], [1e-3, 0.0, 0.0]])]
    ps = [dclust.make_params(3, 1e-6, 2), dclust.make_params(3, 1e-6, 2),
          dclust.make_params(3, 1.0, 2)]
    labs, errs = handle.run_batched(xs, ps)
    assert errs.tolist() == [dclust.E_CAPACITY] * 3
    # the handle stays usable
    x = _cube(3)
    assert np.array_equal(handle.run(x, 0.3, 3), oracle.fast_dbscan(x, 0.3, 3)[0])


def test_key_space_overflow_inside_a_replayed_graph(handle, devmem):
    # identical calls are replayed from a captured graph; new coordinates
    # that overflow the key space abandon the replay and fail the call
    small = [np.array([[0.0, 0.0, 0.0], [1e-5, 0.0, 0.0], [2e-5, 0.0, 0.0]])] * 2
    big = [_cube(4)[:3], _cube(5)[:3]]
    ns = np.array([3, 3], dtype=np.int64)
    ps = [dclust.make_params(3, 1e-6, 2, precision=dclust.FP32_GUARD)] * 2
    d_x = devmem.alloc(2 * 3 * 3 * 8)
    d_l = devmem.alloc(2 * 3 * 4)
    devmem.h2d(d_x, np.concatenate([a.ravel() for a in small]))
    for _ in range(3):
        assert not np.any(handle.run_batched_device(d_x, ns, ps, d_l))
    assert handle.last_stats["diag"][14] == 1  # replayed
    devmem.h2d(d_x, np.concatenate([a.ravel() for a in big]))
    errs = handle.run_batched_device(d_x, ns, ps, d_l)
    assert errs.tolist() == [dclust.E_CAPACITY] * 2
    devmem.h2d(d_x, np.concatenate([a.ravel() for a in small]))
    assert not np.any(handle.run_batched_device(d_x, ns, ps, d_l))
    lab = devmem.download(np.empty(6, dtype=np.int32), d_l)
    assert lab.tolist() == [0] * 6  # 1e-5 apart at eps 1e-6: all noise


# ---- FP64 predicate: no fused multiply-add ----
def _sums(dx, e2):
    """(unfused sum <= e2, set of fused-evaluation outcomes) for one pair."""
    sq = dx * dx
    sep = sq[0]
    for j in range(1, len(dx)):
        sep = sep + sq[j]
    if abs(sep - e2) > 8 * np.spacing(e2):  # one rounding less cannot cross e2
        return bool(sep <= e2), {bool(sep <= e2)}
    fx = [Fraction(float(v)) for v in dx]
    part = sq[0]
    for j in range(1, len(dx) - 1):
        part = part + sq[j]
    # fma on the last product, or on the first (then the rest added rounded)
    fused_last = float(fx[-1] * fx[-1] + Fraction(float(part)))
    fused_first = float(fx[0] * fx[0] + Fraction(float(sq[1])))
    if len(dx) == 3:
        fused_first = float(Fraction(fused_first) + Fraction(float(sq[2])))
    return bool(sep <= e2), {fused_last <= e2, fused_first <= e2}


def _fma_sensitive_pairs(dim, eps, count, seed):
    """Isolated pairs (p, q) whose exact differences dx = q - p make the
    oracle's rounded sum fl(fl(dx0^2) + fl(dx1^2) [+ fl(dx2^2)]) <= fl(eps^2)
    disagree with a single-rounding (fused) evaluation of the same sum.  Each
    pair sits in its own 4 eps box; the last coordinate of q is stepped a few
    ulps around |q - p| = eps until the two evaluations disagree."""
    rng = np.random.default_rng(seed)
    e2 = eps * eps
    P, Q, inside = [], [], []
    side = int(np.ceil(count ** (1.0 / dim))) + 1
    tries = 0
    while len(P) < count and tries < 200 * count:
        tries += 1
        m = len(P)
        cell = np.array([(m // side ** j) % side for j in range(dim)], dtype=np.float64)
        p = 1.0 + 4.0 * eps * cell + rng.uniform(0, eps, dim)
        u = np.abs(rng.normal(size=dim)) + 0.3
        u /= np.linalg.norm(u)
        q0 = p + u * eps
        for k in rng.permutation(np.arange(-12, 13)):
            q = q0.copy()
            q[-1] = q0[-1] + k * np.spacing(q0[-1])
            dx = q - p  # exact (Sterbenz: p and q within a factor 2)
            sep_in, fused = _sums(dx, e2)
            if any(f != sep_in for f in fused) and np.all(np.abs(dx) < eps):
                P.append(p)
                Q.append(q)
                inside.append(sep_in)
                break
    return np.array(P), np.array(Q), np.array(inside)


@pytest.mark.parametrize("dim", [2, 3])
def test_fma_sensitive_pairs_follow_the_unfused_predicate(handle, dim):
    eps = 0.37
    P, Q, inside = _fma_sensitive_pairs(dim, eps, 300, 40 + dim)
    assert len(P) == 300 and 30 < inside.sum() < 270  # both outcomes occur
    for p, q, v in zip(P, Q, inside):
        assert oracle.pred(p, q, eps) == v
    x = np.concatenate([P, Q])
    ref, core, _ = oracle.fast_dbscan(x, eps, 2)
    n = len(P)
    # with minpts 2 a pair is a cluster exactly when the unfused predicate holds
    assert np.array_equal(ref[:n] > 0, inside)
    for prec in (dclust.FP64, dclust.FP32_GUARD):
        lab = handle.run(x, eps, 2, precision=prec)
        assert np.array_equal(lab, ref), f"precision {prec}"
        assert np.array_equal(handle.core_flags(2 * n), core)


# ---- launch shape: built-in table, and PARTITION_SIZE ----
@pytest.mark.parametrize("case", ["2d_small", "3d_mid", "mixed_batch"])
def test_auto_launch_shape_comes_from_the_table(handle, case):
    # the built-in table (shapes measured per dimension mix and size class):
    # <= 2^16 points of one dimension -> 256 seeds x 128 threads; 2^18-2^19
    # of one dimension -> 2048 x 128; a 2-D/3-D batch below 2^23 -> 1024 x 256
    rng = np.random.default_rng(51)
    if case == "2d_small":
        x = rng.normal(0, 1, (20000, 2))
        lab = handle.run(x, 0.05, 4)
        want = (256, 128)
        assert np.array_equal(lab, oracle.fast_dbscan(x, 0.05, 4)[0])
    elif case == "3d_mid":
        x, info = dclust.synth_config(3, n=300000)
        lab = handle.run(x, info["eps"], info["minpts"])
        want = (2048, 128)
        assert np.array_equal(lab, oracle.fast_dbscan(x, info["eps"], info["minpts"])[0])
    else:
        xs = [rng.normal(0, 1, (5000, 2)), rng.normal(0, 1, (5000, 3))]
        ps = [dclust.make_params(2, 0.05, 4), dclust.make_params(3, 0.15, 4)]
        labs, errs = handle.run_batched(xs, ps)
        assert not np.any(errs)
        want = (1024, 256)
        for x, p, lab in zip(xs, ps, labs):
            assert np.array_equal(lab, oracle.fast_dbscan(x, p.eps, p.minpts)[0])
    st = handle.last_stats
    assert (st["eff_seed_list_size"], st["eff_block_threads"]) == want
    assert st["eff_blocks"] > 0


def test_partition_size_changes_nothing(handle, golden):
    # r only validates the bin width (main.cu:204); the grid is fixed by eps
    x, info = dclust.synth_config(1)
    ref = golden["cfg1_labels"].astype(np.int32)
    for r in (0, 1, 2, 5):
        assert np.array_equal(handle.run(x, info["eps"], info["minpts"], r=r), ref)
    with pytest.raises(dclust.DclustError) as e:
        handle.run(x, info["eps"], info["minpts"], r=200)  # extent / 200 < eps
    assert e.value.code == dclust.E_BINWIDTH
