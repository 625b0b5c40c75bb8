"""fpcc_allpairs_device_noalloc (workspace in the library's fixed device scratch)
against fpcc_allpairs_device (caller's workspace) through the C ABI: bit for bit
the same output on two calls in a row, the second with fewer stations (so the
scratch still holds the first call's data), for several kinds and stacks; a job
whose workspace exceeds the scratch is refused before anything runs; a job
whose tables were never built is refused until an ordinary call builds them."""
import ctypes

import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu


def _lib(fp):
    L = fp.api._lib.lib()
    jp = ctypes.POINTER(fp.api.FpccJob)
    L.fpcc_allpairs_device_noalloc.argtypes = [jp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.fpcc_allpairs_device_noalloc.restype = ctypes.c_int
    return L


def _ordinary(fp, hip, L, job, dx, ny, st):
    ws_bytes = L.fpcc_workspace_size(ctypes.byref(job))
    dws, dy = hip.alloc(ws_bytes), hip.alloc(4 * ny)
    assert L.fpcc_allpairs_device(ctypes.byref(job), dx, dy, dws, ws_bytes, st) == 0
    hip.sync(st)
    y = np.empty(ny, np.float32)
    hip.d2h(y, dy)
    return y


def _noalloc(hip, L, job, dx, ny, st):
    dy = hip.alloc(4 * ny)
    r = L.fpcc_allpairs_device_noalloc(ctypes.byref(job), dx, dy, st)
    assert r == 0, r
    hip.sync(st)
    y = np.empty(ny, np.float32)
    hip.d2h(y, dy)
    return y


@pytest.mark.parametrize("kind,stack,W,N,lag", [("pcc2", "gram16", 24, 3600, 1000), ("pcc2", "plain", 4, 3600, 1000),
                                                ("cc1b", "gram", 16, 1200, 300), ("gncc", "plain", 3, 2000, 500)])
def test_noalloc_matches_allpairs_device(fp, hip, kind, stack, W, N, lag):
    L = _lib(fp)
    st = hip.stream()
    for S in (30, 13):  # the second call smaller: the scratch keeps the first call's contents
        print(f"noalloc {kind}/{stack} S={S} W={W} N={N} lags +-{lag}", flush=True)
        job = fp.make_job(kind, S, W, N, -lag, lag, stack=stack)
        ws_bytes = L.fpcc_workspace_size(ctypes.byref(job))
        assert 0 < ws_bytes <= (128 << 20)
        x, _ = synth.make_stations(S, W * N, lag // 2, seed=S + N)
        ny = L.fpcc_output_len(ctypes.byref(job))
        dx = hip.alloc(x.nbytes)
        hip.h2d(dx, x)
        y_ref = _ordinary(fp, hip, L, job, dx, ny, st)
        y = _noalloc(hip, L, job, dx, ny, st)
        print(f"  max|noalloc - allpairs_device| = {np.abs(y - y_ref).max():.3e}", flush=True)
        assert np.array_equal(y, y_ref)


def test_noalloc_refuses_oversized_job(fp, hip):
    """64 stations x 24 windows of 3600 on the split Gram needs ~270 MB of
    workspace: refused (-6) before any launch."""
    L = _lib(fp)
    job = fp.make_job("pcc2", 64, 24, 3600, -1000, 1000, stack="gram16")
    assert L.fpcc_workspace_size(ctypes.byref(job)) > (128 << 20)
    dx, dy = hip.alloc(64), hip.alloc(64)
    assert L.fpcc_allpairs_device_noalloc(ctypes.byref(job), dx, dy, None) == -6
    bad = fp.make_job("pcc2", 8, 1, 100, -100, 10, stack="plain")  # |lag| >= N
    assert L.fpcc_allpairs_device_noalloc(ctypes.byref(bad), dx, dy, None) == -3


def test_noalloc_needs_built_tables(fp, hip):
    """A shape whose Gram tile list was never built (37 stations) is refused (-5)
    until one ordinary call builds it; then both agree bit for bit.  (If an earlier
    test in this process built the tables already, the first call just runs.)"""
    L = _lib(fp)
    S, W, N, lag = 37, 2, 1764, 200
    job = fp.make_job("pcc2", S, W, N, -lag, lag, stack="gram")
    x, _ = synth.make_stations(S, W * N, 100, seed=37)
    ny = L.fpcc_output_len(ctypes.byref(job))
    st = hip.stream()
    dx, dy = hip.alloc(x.nbytes), hip.alloc(4 * ny)
    hip.h2d(dx, x)
    r = L.fpcc_allpairs_device_noalloc(ctypes.byref(job), dx, dy, st)
    print(f"first noalloc call on a new shape returned {r}", flush=True)
    assert r in (-5, 0)
    hip.sync(st)
    y_ref = _ordinary(fp, hip, L, job, dx, ny, st)
    y = _noalloc(hip, L, job, dx, ny, st)
    assert np.array_equal(y, y_ref)
