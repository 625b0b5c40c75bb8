"""Split-FP16 MFMA Gram stack (FP16 hi+lo operands, FP32 accumulation) against
the oracle and the all-FP64 oracle.  Bar: 1e-4 absolute against FP64 for this
reduced-precision path; the measured errors are printed."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL16 = 1e-4


@pytest.mark.parametrize("kind,S,W,N,lag", [("pcc2", 8, 1, 3600, 500), ("pcc2", 40, 4, 3600, 1000),
                                            ("pcc2", 19, 3, 1000, 200), ("cc1b", 21, 3, 2000, 400),
                                            ("pcc2", 5, 2, 12000, 2000), ("cc1b", 5, 2, 12000, 2000)])
def test_gram16_matches_fp64(fp, orc, kind, S, W, N, lag):
    x, _ = synth.make_stations(S, W * N, lag // 2, seed=S + W + N)
    y16 = fp.allpairs(x, W, N, -lag, lag, kind=kind, stack="gram16")
    y32 = fp.allpairs(x, W, N, -lag, lag, kind=kind, stack="gram")
    npr = min(y16.shape[0], 40)
    y64 = orc.allpairs(kind, x, W, N, -lag, lag, stacked=True, f64=True, pair_end=npr)
    e64 = np.abs(y16[:npr] - y64).max()
    rms = np.sqrt(np.mean((y16[:npr] - y64) ** 2))
    e32 = np.abs(y16 - y32).max()
    print(f"{kind} S={S} W={W} N={N}: gram16 max|.-fp64|={e64:.3e} rms={rms:.3e} max|gram16-gram32|={e32:.3e}")
    assert e64 <= TOL16


@pytest.mark.parametrize("kind", ["pcc2", "cc1b"])
@pytest.mark.parametrize("lag1,lag2", [(-3599, -3000), (500, 500), (-600, 2000), (3000, 3599)])
def test_gram_stacks_uneven_lag_windows(fp, orc, kind, lag1, lag2):
    """Nz = 8192 (N = 3600, max |lag| >= 497): lag windows on one side of zero, a
    single lag and an asymmetric window through both Gram stacks and the
    specialised pair IFFT, against the FP64 oracle (1-bit: counting variant with
    the exact short-overlap lags near +-(N-1))."""
    S, W, N = 7, 3, 3600
    x, _ = synth.make_stations(S, W * N, 300, seed=lag2 - lag1 + 7)
    y64 = orc.allpairs(kind, x, W, N, lag1, lag2, stacked=True, f64=True)
    for stack, tol in (("gram", 1e-5), ("gram16", TOL16)):
        y = fp.allpairs(x, W, N, lag1, lag2, kind=kind, stack=stack)
        err = np.abs(y.reshape(y64.shape) - y64).max()
        print(f"{kind}/{stack} lags ({lag1},{lag2}): max|gpu-fp64|={err:.3e}", flush=True)
        assert err <= tol
