"""The C ABI library loads and exports every function include/*.h declares;
host-only entry points behave (no GPU needed: nothing here calls HIP)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import dclust

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {
    "dclust.h": dclust.LIB_PATH,
    "dclust_synth.h": dclust.SYNTH_PATH,
}


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dclust_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_exports_every_declared_symbol(header):
    names = declared(header)
    assert len(names) >= 5
    lib = ctypes.CDLL(HEADERS[header])
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing
    nm = subprocess.run(["nm", "-D", "--defined-only", HEADERS[header]], capture_output=True,
                        text=True, check=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}$", nm, flags=re.M), n


def test_params_default_mirror_common_h():
    p = dclust.default_params()
    # common.h:13-25: THREAD_BLOCKS 8, THREAD_COUNT 16, MAX_SEEDS 128,
    # EXTRA_COLLISION_SIZE 256, DIMENSION 2, MINPTS 3, EPS 1.5, PARTITION_SIZE 3
    assert (p.blocks, p.block_threads, p.seed_list_size, p.extra_collision) == (8, 16, 128, 256)
    assert (p.dim, p.minpts, p.eps, p.partitions_r) == (2, 3, 1.5, 3)
    assert p.precision == dclust.FP64


def test_strerror_and_version():
    for code in (0, -1, -2, -3, -4, -5, -6, -7):
        assert dclust.strerror(code)
    assert "gfx950" in dclust.version()


def test_struct_layout_matches_header(tmp_path):
    # the C compiler's own layout of include/dclust.h against the ctypes mirror
    fields = {"dclust_params": dclust.Params, "dclust_stats": dclust.Stats}
    src = ["#include <stddef.h>", "#include <stdio.h>", '#include "dclust.h"', "int main(void) {"]
    for st, cls in fields.items():
        src.append(f'  printf("{st} sizeof %zu\\n", sizeof({st}));')
        for name, _ in cls._fields_:
            src.append(f'  printf("{st} {name} %zu\\n", offsetof({st}, {name}));')
    src += ["  return 0;", "}"]
    (tmp_path / "layout.c").write_text("\n".join(src) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(tmp_path / "layout.c"), "-o",
                    str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True,
                               check=True).stdout.splitlines():
        st, name, val = line.split()
        got[(st, name)] = int(val)
    for st, cls in fields.items():
        assert got[(st, "sizeof")] == ctypes.sizeof(cls), st
        for name, _ in cls._fields_:
            assert got[(st, name)] == getattr(cls, name).offset, (st, name)


def test_device_entry_points_reject_bad_arguments_before_any_gpu_call():
    lib = ctypes.CDLL(dclust.LIB_PATH)
    P = ctypes.c_void_p
    p = dclust.make_params(2, 1.5, 3)
    for fn in (lib.dclust_run_device, lib.dclust_run_onestep_device):
        fn.argtypes = [P, P, ctypes.c_int64, ctypes.POINTER(dclust.Params), P, P]
        # no handle, no parameters, a negative count
        assert fn(None, None, 0, ctypes.byref(p), None, None) == dclust.E_PARAM
        assert fn(P(1), None, 0, None, None, None) == dclust.E_PARAM
        assert fn(P(1), None, -1, ctypes.byref(p), None, None) == dclust.E_PARAM


def test_stats_view_is_a_mapping_over_the_struct():
    # run_device returns a lazy mapping over dclust_stats: every field, the
    # arrays as lists, a KeyError for an unknown name, and dict() of it equal
    # to the struct's full conversion
    st = dclust.Stats()
    st.t_total_ms = 1.5
    st.n_clusters = 7
    st.diag[14] = 1
    st.work[2] = 9
    v = dclust.StatsView(st)
    assert v["t_total_ms"] == 1.5 and v["n_clusters"] == 7
    assert v["diag"][14] == 1 and isinstance(v["diag"], list) and len(v["diag"]) == 16
    assert v["work"] == [0, 0, 9, 0]
    assert dict(v) == st.as_dict()
    assert set(v) == {k for k, _ in dclust.Stats._fields_}
    assert "t_core_ms" in v and "nope" not in v
    with pytest.raises(KeyError):
        v["nope"]
