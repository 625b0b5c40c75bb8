"""The station stages on windows that reach their special cases, against the FP64
oracle: the specialised kernel (hourly windows of 3600 samples, nfft 8192), the
runtime-plan kernel (2000 samples, nfft 4096) and the four-step path (5000 samples,
nfft 16384), each with a window of zeros
(AmpNormf's reciprocal is not normal, the phasors are zero), a constant window (only
bin 0, kept at the analytic gain da1 instead of doubled) and an alternating window
(only bin N/2, also kept at da1), beside noise windows; per-window output and the
split-FP16 mean stack."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


def special_network(seed, N=3600):
    S, W = 6, 4
    x, _ = synth.make_stations(S, W * N, 300, seed=seed)
    x = x.astype(np.float32)
    x[0, 1 * N:2 * N] = 0.0                                   # a silent hour
    x[3, 0:N] = 2.5                                           # constant: bin 0 only
    x[4, 2 * N:3 * N] = np.where(np.arange(N) % 2 == 0, 1.5, -1.5)  # bin N/2 only
    x[5, 3 * N:4 * N] = -0.75                                 # constant of the other sign
    x[2, 3 * N:4 * N] = 2.5                                   # the same constant as station 3
    return x, S, W, N


@pytest.mark.parametrize("N", [3600, 2000, 5000])
@pytest.mark.parametrize("stack", ["none", "gram16"])
def test_station_special_windows(fp, orc, stack, N):
    x, S, W, N = special_network(29, N)
    print(f"pcc2/{stack} S={S} W={W} N={N}: zero, constant and alternating windows", flush=True)
    y = fp.allpairs(x, W, N, -1000, 1000, kind="pcc2", stack=stack)
    y64 = orc.allpairs("pcc2", x, W, N, -1000, 1000, stacked=(stack != "none"), f64=True)
    err = float(np.abs(y.reshape(y64.shape) - y64).max())
    print(f"  max|gpu-fp64|={err:.3e}")
    assert np.isfinite(y).all()
    assert err <= (1e-4 if stack == "gram16" else TOL)
    if stack == "none":
        yw = y.reshape(S * (S - 1) // 2, W, -1)
        P = lambda i, j: i * S - i * (i + 1) // 2 + (j - i - 1)  # noqa: E731
        # pairs with the silent hour correlate to zero in that window
        assert np.abs(yw[P(0, 1), 1]).max() == 0.0
        # station 2's and station 5's last hours are constants of opposite signs: the
        # phasors are +-(1 - 1e-6) (AmpNormf's eps), and PCC2, normalised by the
        # window length, is -(N - |lag|) / N
        lags = np.arange(-1000, 1001)
        expect = -(N - np.abs(lags)) / N
        assert np.abs(yw[P(2, 5), 3] - expect).max() < 1e-5
