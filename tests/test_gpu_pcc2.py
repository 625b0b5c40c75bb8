"""PCC2 on the GPU against the oracle (reference-precision restatement of
src/FFTapps.c:785-881) and the all-FP64 oracle."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-5  # absolute, FP32 path vs the reference-precision oracle


def _rand(Tr, N, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((Tr, N)).astype(np.float32), rng.standard_normal((Tr, N)).astype(np.float32))


@pytest.mark.parametrize("N,Tr,lag1,lag2", [
    (3600, 3, -500, 500),
    (100, 3, -40, 60),
    (64, 2, -10, 10),
    (1000, 2, 5, 300),
    (3600, 2, -3599, 3599),
    (2401, 1, -7, 0),
])
def test_pcc2_set_matches_oracle(fp, orc, N, Tr, lag1, lag2):
    print(f"pcc2_set N={N} Tr={Tr} lags=({lag1},{lag2})", flush=True)
    x1, x2 = _rand(Tr, N, N + Tr)
    y = fp.pcc2_set(x1, x2, lag1, lag2)
    yr = orc.run("pcc2", x1, x2, lag1, lag2)
    y64 = orc.run("pcc2", x1, x2, lag1, lag2, f64=True)
    err = np.abs(y - yr).max()
    print(f"  max|gpu-oracle|={err:.3e} max|gpu-fp64|={np.abs(y - y64).max():.3e}")
    assert err <= TOL


def test_pcc2_set_swapped_lags(fp, orc):
    x1, x2 = _rand(2, 500, 7)
    y = fp.pcc2_set(x1, x2, 80, -30)  # Lag1 > Lag2: window starts at min(Lag1, Lag2)
    yr = orc.run("pcc2", x1, x2, 80, -30)
    assert y.shape == yr.shape == (2, 111)
    assert np.abs(y - yr).max() <= TOL


def test_pcc2_set_rejects_large_lags(fp):
    x1, x2 = _rand(1, 100, 3)
    with pytest.raises(fp.FastPCCError) as e:
        fp.pcc2_set(x1, x2, -100, 10)
    assert e.value.code == -3


def test_allpairs_config1(fp, orc):
    """8 stations x 3600 samples, 28 pairs, lags +-500, one window."""
    x, shifts = synth.make_stations(8, 3600, 200, seed=11)
    y = fp.allpairs(x, 1, 3600, -500, 500, stack="plain")
    yr = orc.allpairs("pcc2", x, 1, 3600, -500, 500, stacked=True)
    y64 = orc.allpairs("pcc2", x, 1, 3600, -500, 500, stacked=True, f64=True)
    err = np.abs(y - yr).max()
    print(f"C1 max|gpu-oracle|={err:.3e} max|gpu-fp64|={np.abs(y - y64).max():.3e}")
    assert err <= TOL
    # each pair peaks at the true relative delay
    for p, (i, j) in enumerate(synth.pair_list(8)):
        assert np.argmax(y[p]) - 500 == shifts[j] - shifts[i]


@pytest.mark.parametrize("stack", ["none", "plain", "zsum", "gram"])
def test_allpairs_windows(fp, orc, stack):
    S, W, N = 5, 4, 3600
    x, _ = synth.make_stations(S, W * N, 300, seed=5)
    y = fp.allpairs(x, W, N, -1000, 1000, stack=stack)
    yr = orc.allpairs("pcc2", x, W, N, -1000, 1000, stacked=(stack != "none"))
    err = np.abs(y - yr).max()
    print(f"stack={stack} max|gpu-oracle|={err:.3e}")
    assert y.shape == yr.shape
    assert err <= TOL


@pytest.mark.parametrize("S,W,N,lag", [(8, 1, 3600, 500), (37, 3, 1000, 200), (50, 2, 3600, 1000), (17, 5, 640, 100)])
def test_gram_stack_matches_oracle(fp, orc, S, W, N, lag):
    """MFMA Gram stack with station counts that do not fill the 16x32 tiles,
    odd and even window counts."""
    print(f"gram S={S} W={W} N={N} lag={lag}", flush=True)
    x, _ = synth.make_stations(S, W * N, lag // 2, seed=S * 100 + W)
    y = fp.allpairs(x, W, N, -lag, lag, stack="gram")
    yp = fp.allpairs(x, W, N, -lag, lag, stack="plain")
    npr = min(y.shape[0], 60)  # oracle on a prefix of the pairs keeps the CPU time low
    yr = orc.allpairs("pcc2", x, W, N, -lag, lag, stacked=True, pair_end=npr)
    err = np.abs(y[:npr] - yr).max()
    print(f"  max|gram-oracle|={err:.3e} max|gram-plain| (all pairs)={np.abs(y - yp).max():.3e}")
    assert err <= TOL
    assert np.abs(y - yp).max() <= TOL


def test_pcc2_set_many_rows(fp, orc):
    """256 windows of 21600 samples (44 MB of input rows): the host entry gathers
    the rows on several threads in 8 MB pieces and copies the output rows back
    on several threads; every row against the oracle."""
    rng = np.random.default_rng(21600)
    x1 = rng.standard_normal((256, 21600)).astype(np.float32)
    x2 = rng.standard_normal((256, 21600)).astype(np.float32)
    print("pcc2_set Tr=256 N=21600 lags (-3000,3000)", flush=True)
    y = fp.pcc2_set(x1, x2, -3000, 3000)
    yr = orc.run("pcc2", x1, x2, -3000, 3000)
    err = np.abs(y - yr).max()
    print(f"  max|gpu-oracle|={err:.3e}")
    assert y.shape == yr.shape and err <= TOL


def test_allpairs_host_buffers_large(fp, hip):
    """fpcc_allpairs from host buffers big enough for the threaded pinned staging
    in both directions (72 stations x 86400 samples = 24.9 MB in, 2556 pairs x 2001
    lags = 20.5 MB out; the threaded copies start at two 8 MiB pieces) gives bit for
    bit the device-buffer run of the same job."""
    import ctypes
    S, W, N = 72, 24, 3600
    piece = 8 << 20
    assert 4 * S * W * N >= 2 * piece and 4 * (S * (S - 1) // 2) * 2001 >= 2 * piece
    x, _ = synth.make_stations(S, W * N, 300, seed=64)
    y_host = fp.allpairs(x, W, N, -1000, 1000, kind="pcc2", stack="gram16").ravel()
    L = fp.api._lib.lib()
    job = fp.make_job("pcc2", S, W, N, -1000, 1000, stack="gram16")
    ws_bytes = L.fpcc_workspace_size(ctypes.byref(job))
    ny = L.fpcc_output_len(ctypes.byref(job))
    st = hip.stream()
    dx, dy, dws = hip.alloc(x.nbytes), hip.alloc(4 * ny), hip.alloc(ws_bytes)
    hip.h2d(dx, x)
    assert L.fpcc_allpairs_device(ctypes.byref(job), dx, dy, dws, ws_bytes, st) == 0
    hip.sync(st)
    y_dev = np.empty(ny, np.float32)
    hip.d2h(y_dev, dy)
    print(f"host vs device buffers: {np.abs(y_host - y_dev).max():.3e}", flush=True)
    assert np.array_equal(y_host, y_dev)
