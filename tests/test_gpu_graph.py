"""HIP-graph replay of whole jobs: each replay, with new input uploaded between
replays (eager copies and synchronisations in between), equals the eager result
exactly; the pipeline records no memset nodes."""
import ctypes

import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu

CASES = [("pcc2", 9, 2, 2400, 300, "gram16", {}), ("pcc2", 6, 1, 3600, 500, "none", {}),
         ("pcc2", 5, 2, 2003, 500, "gram", {}), ("gncc", 4, 1, 12000, 1000, "none", {}),
         ("cc1b", 7, 2, 3000, 600, "gram", {}), ("pcc1", 4, 1, 2500, 200, "plain", {}),
         ("wpcc", 3, 1, 3000, 400, "none", dict(pmin=2.0, pmax=200.0, V=8, wtype=-1))]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_graph_replays_match_eager(fp, hip, case):
    kind, S, W, N, lag, stack, kw = CASES[case]
    L = fp.api._lib.lib()
    job = fp.make_job(kind, S, W, N, -lag, lag, stack=stack, **kw)
    ws_bytes = L.fpcc_workspace_size(ctypes.byref(job))
    ny = L.fpcc_output_len(ctypes.byref(job))
    st = hip.stream()
    dx, dy, dws = hip.alloc(4 * S * W * N), hip.alloc(4 * ny), hip.alloc(ws_bytes)
    g = ctypes.c_void_p()
    x0, _ = synth.make_stations(S, W * N, 200, seed=1)
    hip.h2d(dx, x0)
    assert L.fpcc_graph_create(ctypes.byref(job), dx, dy, dws, ws_bytes, st, ctypes.byref(g)) == 0
    try:
        for rep in range(3):
            x, _ = synth.make_stations(S, W * N, 200, seed=10 + rep)
            hip.h2d(dx, x)
            assert L.fpcc_graph_launch(g, st) == 0
            hip.sync(st)
            y = np.empty(ny, np.float32)
            hip.d2h(y, dy)
            ref = fp.allpairs(x, W, N, -lag, lag, kind=kind, stack=stack, **kw).ravel()
            print(f"{kind}/{stack} replay {rep}: max|graph - eager| = {np.abs(y - ref).max():.3e}")
            assert np.array_equal(y, ref)
    finally:
        L.fpcc_graph_destroy(g)
