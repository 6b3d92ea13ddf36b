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
    # common.h:19-25 defaults (DIMENSION 2, MINPTS 3, EPS 1.5, PARTITION_SIZE 3)
    assert (p.dim, p.minpts, p.eps, p.partitions_r) == (2, 3, 1.5, 3)
    assert p.precision == dclust.FP64


def test_strerror_and_version():
    for code in (0, -1, -2, -3, -4, -5, -6, -7):
        assert dclust.strerror(code)
    assert "gfx950" in dclust.version()


def test_struct_layout_matches_header():
    assert ctypes.sizeof(dclust.Params) == 40
    assert dclust.Params.eps.offset == 8
    assert ctypes.sizeof(dclust.Stats) == 9 * 8 + 17 * 8 + 2 * 4 + 16 * 8


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
