import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def fp():
    import fastpcc_amd
    fastpcc_amd._lib.lib()
    return fastpcc_amd


class HipMem:
    """Device buffers and streams from the HIP runtime the library links
    (tests that drive the device-level C ABI without another runtime)."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.ptrs, self.streams = [], []

    def alloc(self, n):
        p = self.ct.c_void_p()
        assert self.h.hipMalloc(self.ct.byref(p), max(n, 1)) == 0
        self.ptrs.append(p.value)
        return p.value

    def h2d(self, d, a):
        assert self.h.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, d):
        assert self.h.hipMemcpy(a.ctypes.data, d, a.nbytes, 2) == 0

    def stream(self):
        s = self.ct.c_void_p()
        assert self.h.hipStreamCreate(self.ct.byref(s)) == 0
        self.streams.append(s.value)
        return s.value

    def sync(self, s):
        assert self.h.hipStreamSynchronize(s) == 0

    def free(self):
        for s in self.streams:
            self.h.hipStreamDestroy(s)
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs, self.streams = [], []


@pytest.fixture
def hip(fp):
    h = HipMem()
    yield h
    h.free()
