"""The pair inverses at nfft 8192 over lag windows on both sides of +-1024: windows
inside [-1024, 1023] take the last pass that forms only the outputs those lags
read (the transform's outputs 0..511 and 3584..4095), the others the full pass.
The Gram stacks' inverse (PCC2 and 1-bit, FP32 and split-FP16), the per-window
PCC2 path (every window, their mean, the summed spectra), the phase-weighted
stack and GNCC's FP64 inverse (every window, their mean), every pair against the FP64
oracle."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


WINDOWS = [(-1024, 1023), (-1025, 1023), (-1024, 1024), (0, 1023), (-1024, -1), (-7, 7), (-3599, 3599)]


@pytest.mark.parametrize("kind", ["pcc2", "cc1b"])
@pytest.mark.parametrize("stack", ["gram", "gram16"])
@pytest.mark.parametrize("lag1,lag2", WINDOWS)
def test_c2r_lag_windows(fp, orc, kind, stack, lag1, lag2):
    S, W, N = 6, 2, 3600
    print(f"{kind}/{stack} S={S} W={W} N={N} lags ({lag1},{lag2})", flush=True)
    x, _ = synth.make_stations(S, W * N, 300, seed=lag2 - lag1)
    if kind == "cc1b":
        x[:, ::97] = 0.0
    y = fp.allpairs(x, W, N, lag1, lag2, kind=kind, stack=stack)
    y64 = orc.allpairs(kind, x, W, N, lag1, lag2, stacked=True, f64=True)
    err = float(np.abs(y.reshape(y64.shape) - y64).max())
    print(f"  max|gpu-fp64|={err:.3e}")
    assert y.shape[-1] == lag2 - lag1 + 1
    assert err <= (1e-4 if stack == "gram16" else TOL)


@pytest.mark.parametrize("stack", ["none", "plain", "zsum"])
@pytest.mark.parametrize("lag1,lag2", WINDOWS)
def test_plain_pair_lag_windows(fp, orc, stack, lag1, lag2):
    S, W, N = 5, 3, 3600
    print(f"pcc2/{stack} S={S} W={W} N={N} lags ({lag1},{lag2})", flush=True)
    x, _ = synth.make_stations(S, W * N, 300, seed=7 + lag2 - lag1)
    y = fp.allpairs(x, W, N, lag1, lag2, kind="pcc2", stack=stack)
    y64 = orc.allpairs("pcc2", x, W, N, lag1, lag2, stacked=(stack != "none"), f64=True)
    err = float(np.abs(y.reshape(y64.shape) - y64).max())
    print(f"  max|gpu-fp64|={err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("kind", ["pcc2", "cc1b"])
@pytest.mark.parametrize("lag1,lag2", WINDOWS)
def test_pws_lag_windows(fp, orc, kind, lag1, lag2):
    S, W, N = 4, 3, 3600
    print(f"{kind}/pws S={S} W={W} N={N} lags ({lag1},{lag2})", flush=True)
    x, _ = synth.make_stations(S, W * N, 300, seed=11 + lag2 - lag1)
    y = fp.allpairs(x, W, N, lag1, lag2, kind=kind, stack="pws", pws_power=2.0)
    yr = orc.pws(kind, x, W, N, lag1, lag2, 2.0)
    err = float(np.abs(y - yr).max())
    print(f"  max|gpu-fp64|={err:.3e}")
    assert y.shape == yr.shape and err <= TOL


@pytest.mark.parametrize("stack", ["none", "plain"])
@pytest.mark.parametrize("lag1,lag2", WINDOWS)
def test_gncc_lag_windows(fp, orc, stack, lag1, lag2):
    """GNCC's FP64 pair inverse: windows inside [-1024, 1023] take the path whose
    butterflies start and end in registers, the others the LDS path."""
    S, W, N = 4, 3, 3600
    print(f"gncc/{stack} S={S} W={W} N={N} lags ({lag1},{lag2})", flush=True)
    x, _ = synth.make_stations(S, W * N, 300, seed=23 + lag2 - lag1)
    y = fp.allpairs(x, W, N, lag1, lag2, kind="gncc", stack=stack)
    y64 = orc.allpairs("gncc", x, W, N, lag1, lag2, stacked=(stack != "none"), f64=True)
    err = float(np.abs(y.reshape(y64.shape) - y64).max())
    print(f"  max|gpu-fp64|={err:.3e}")
    assert err <= TOL
