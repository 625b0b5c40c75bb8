"""The DMA-staged split-FP16 Gram (rows of 2W split elements, 2W a multiple of
16: PCC2 with 8, 16, 24 or 32 windows, 1-bit with 16 or 32), including station
counts that leave partial 64-station blocks, and the full 256-station headline
shape: every pair against the FP32 Gram stack, pairs on and across block
boundaries against the all-FP64 oracle.  Bar: 1e-4 absolute against FP64."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL16 = 1e-4


def _pairs(S, i, j):
    return i * S - i * (i + 1) // 2 + (j - i - 1)


def _check_ranges(orc, kind, x, W, N, lag, y16, ranges):
    worst = 0.0
    for b, e in ranges:
        y64 = orc.allpairs(kind, x, W, N, -lag, lag, stacked=True, f64=True, pair_begin=b, pair_end=e)
        worst = max(worst, float(np.abs(y16[b:e] - y64).max()))
    return worst


@pytest.mark.parametrize("kind,S,W,N,lag", [("pcc2", 70, 8, 1000, 200), ("pcc2", 130, 16, 1500, 300),
                                            ("pcc2", 67, 24, 3600, 1000), ("pcc2", 65, 32, 900, 100),
                                            ("cc1b", 70, 16, 1200, 300), ("cc1b", 40, 32, 800, 100)])
def test_dma_gram16_matches_fp64(fp, orc, kind, S, W, N, lag):
    print(f"{kind} S={S} W={W} N={N} lags +-{lag}", flush=True)
    x, _ = synth.make_stations(S, W * N, lag // 2, seed=S * W + N)
    if kind == "cc1b":
        x[:, ::97] = 0.0
    y16 = fp.allpairs(x, W, N, -lag, lag, kind=kind, stack="gram16")
    y32 = fp.allpairs(x, W, N, -lag, lag, kind=kind, stack="gram")
    P = S * (S - 1) // 2
    # first pairs, pairs of the last station block, the last pairs
    ranges = [(0, 12), (_pairs(S, 1, 63), _pairs(S, 1, 63) + 8), (P - 12, P)]
    e64 = _check_ranges(orc, kind, x, W, N, lag, y16, ranges)
    e32 = float(np.abs(y16 - y32).max())
    print(f"  max|gram16-fp64|={e64:.3e} max|gram16-gram32| over all {P} pairs={e32:.3e}")
    assert e64 <= TOL16 and e32 <= TOL16


def test_dma_gram16_headline_shape(fp, orc):
    """256 stations x 24 windows of 3600, nfft 8192, lags +-1000 (the benchmark's
    workload): all 32640 pairs against the FP32 Gram, 48 pairs against FP64."""
    S, W, N, lag = 256, 24, 3600, 1000
    x, _ = synth.make_stations(S, W * N, 1000, seed=1234)
    y16 = fp.allpairs(x, W, N, -lag, lag, kind="pcc2", stack="gram16")
    y32 = fp.allpairs(x, W, N, -lag, lag, kind="pcc2", stack="gram")
    e32 = float(np.abs(y16 - y32).max())
    P = S * (S - 1) // 2
    ranges = [(0, 8), (_pairs(S, 63, 64), _pairs(S, 63, 64) + 8), (_pairs(S, 64, 65), _pairs(S, 64, 65) + 8),
              (_pairs(S, 127, 200), _pairs(S, 127, 200) + 8), (_pairs(S, 190, 255) - 4, _pairs(S, 190, 255) + 4),
              (P - 8, P)]
    e64 = _check_ranges(orc, "pcc2", x, W, N, lag, y16, ranges)
    print(f"headline shape: max|gram16-gram32| over {P} pairs={e32:.3e}, max|gram16-fp64| (48 pairs)={e64:.3e}")
    assert e32 <= TOL16 and e64 <= TOL16


@pytest.mark.parametrize("stack", ["gram", "gram16"])
@pytest.mark.parametrize("S,W", [(2, 8), (5, 24), (17, 8), (20, 16), (31, 32), (32, 24), (33, 24)])
def test_small_network_gram_matches_fp64(fp, orc, stack, S, W):
    """Networks of at most 32 stations take the Gram kernel that gives each of its
    eight waves a bin of its own (33 stations, the 64 x 64 block kernel): every
    pair against the FP64 oracle, FP32 and split-FP16 operands."""
    N, lag = 1000, 200
    print(f"{stack} S={S} W={W} N={N} lags +-{lag}", flush=True)
    x, _ = synth.make_stations(S, W * N, lag // 2, seed=S * 7 + W)
    y = fp.allpairs(x, W, N, -lag, lag, kind="pcc2", stack=stack)
    y64 = orc.allpairs("pcc2", x, W, N, -lag, lag, stacked=True, f64=True)
    err = float(np.abs(y - y64).max())
    print(f"  max|gpu-fp64| over all {S * (S - 1) // 2} pairs = {err:.3e}")
    assert err <= (TOL16 if stack == "gram16" else 1e-5)
