"""Phase-weighted stack (FPCC_STACK_PWS) on the GPU against the FP64 oracle's
(oracle.pws: the stack the reference hands to the external ts-PWS program,
src/PCC_fullpair_1b.c:568-569, restated with the phase of AnalyticSignal of
each window's whole circular correlation).  Bar: 1e-5 absolute."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _p(S, i, j):
    return i * S - i * (i + 1) // 2 + (j - i - 1)


@pytest.mark.parametrize("kind", ["pcc2", "cc1b", "gncc"])
@pytest.mark.parametrize("S,W,N,lag1,lag2,power", [(5, 4, 600, -100, 100, 2.0), (4, 3, 1000, -300, 120, 1.0),
                                                  (6, 6, 3600, -1000, 1000, 3.0), (3, 2, 200, 5, 60, 2.0)])
def test_pws_matches_fp64_oracle(fp, orc, kind, S, W, N, lag1, lag2, power):
    print(f"pws {kind} S={S} W={W} N={N} lags=({lag1},{lag2}) power={power}", flush=True)
    x, _ = synth.make_stations(S, W * N, 50, seed=S * W + N)
    y = fp.allpairs(x, W, N, lag1, lag2, kind=kind, stack="pws", pws_power=power)
    yr = orc.pws(kind, x, W, N, lag1, lag2, power)
    err = float(np.abs(y - yr).max())
    print(f"  max|gpu-fp64 oracle| = {err:.3e} (peak {np.abs(yr).max():.3e})", flush=True)
    assert y.shape == yr.shape and err <= TOL


@pytest.mark.parametrize("kind", ["pcc2", "cc1b", "gncc"])
def test_pws_power_zero_is_the_linear_stack(fp, kind):
    S, W, N = 6, 5, 1200
    x, _ = synth.make_stations(S, W * N, 100, seed=3)
    y0 = fp.allpairs(x, W, N, -300, 300, kind=kind, stack="pws", pws_power=0.0)
    yl = fp.allpairs(x, W, N, -300, 300, kind=kind, stack="plain")
    d = float(np.abs(y0 - yl).max())
    print(f"{kind}: max|pws(0) - plain| = {d:.3e}", flush=True)
    if kind == "pcc2":  # the same sums from another transform
        assert d <= 1e-6
    else:  # the plain stack times a coherence of exactly 1
        assert d == 0.0


def test_pws_c2_shape(fp, orc):
    """64 stations x 24 windows of 3600, nfft 8192, lags +-1000, all 2016 pairs:
    pair ranges at block boundaries against the FP64 oracle, each kind."""
    S, W, N, lag = 64, 24, 3600, 1000
    x, _ = synth.make_stations(S, W * N, 1000, seed=2016)
    P = S * (S - 1) // 2
    for kind in ("pcc2", "cc1b", "gncc"):
        y = fp.allpairs(x, W, N, -lag, lag, kind=kind, stack="pws", pws_power=2.0)
        assert y.shape == (P, 2 * lag + 1)
        worst = 0.0
        for b, e in [(0, 4), (_p(S, 0, 63), _p(S, 1, 2) + 1), (_p(S, 40, 63) - 1, _p(S, 40, 63) + 1), (P - 3, P)]:
            yr = orc.pws(kind, x, W, N, -lag, lag, 2.0, pair_begin=b, pair_end=e)
            worst = max(worst, float(np.abs(y[b:e] - yr).max()))
        print(f"C2 {kind} pws: max|gpu-fp64 oracle| = {worst:.3e}", flush=True)
        assert worst <= TOL


def test_pws_served_equals_alone(fp):
    S, W, N = 7, 4, 2000
    x, _ = synth.make_stations(S, W * N, 100, seed=44)
    alone = fp.allpairs(x, W, N, -400, 400, kind="pcc2", stack="pws", pws_power=2.0).ravel()
    with fp.Server(device_bytes=2 << 30, pinned_bytes=64 << 20) as srv:
        job = fp.make_job("pcc2", S, W, N, -400, 400, stack="pws", pws_power=2.0)
        y = np.empty(alone.size, np.float32)
        assert srv.wait(srv.submit(job, x, y))["state"] == 2
    assert np.array_equal(alone, y)


def test_pws_rejects_long_windows(fp):
    x, _ = synth.make_stations(3, 2 * 9000, 10, seed=1)
    with pytest.raises(fp.FastPCCError) as e:
        fp.allpairs(x, 2, 9000, -100, 100, kind="pcc2", stack="pws")
    assert e.value.code == -1
