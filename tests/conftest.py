import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "fast-cuda-gpu-dbscan_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def handle():
    import dclust

    h = dclust.Handle(0)
    yield h
    h.close()


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    with np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


class DeviceMem:
    """Device buffers through the HIP runtime (plumbing for the *_device
    entry points)."""

    def __init__(self):
        import ctypes as C

        self.C = C
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def upload(self, arr):
        C = self.C
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(arr.nbytes, 1)) == 0
        assert self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        self.live.append(p.value)
        return p.value

    def alloc(self, nbytes):
        C = self.C
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 1)) == 0
        self.live.append(p.value)
        return p.value

    def h2d(self, ptr, arr):
        arr = np.ascontiguousarray(arr)
        assert self.L.hipMemcpy(self.C.c_void_p(ptr), arr.ctypes.data, arr.nbytes, 1) == 0

    def download(self, arr, ptr):
        assert self.L.hipMemcpy(arr.ctypes.data, self.C.c_void_p(ptr), arr.nbytes, 2) == 0
        return arr

    def free_all(self):
        for p in self.live:
            self.L.hipFree(self.C.c_void_p(p))
        self.live = []


@pytest.fixture
def devmem():
    m = DeviceMem()
    yield m
    m.free_all()
