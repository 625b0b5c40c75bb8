"""The pair inverse at nfft 8192 over lag windows on both sides of +-1024: windows
inside [-1024, 1023] take the last pass that forms only the outputs those lags
read (the transform's outputs 0..511 and 3584..4095), the others the full pass.
PCC2 and 1-bit (count-normalised) stacks, FP32 and split-FP16 Gram, every pair
against the FP64 oracle."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.mark.parametrize("kind", ["pcc2", "cc1b"])
@pytest.mark.parametrize("stack", ["gram", "gram16"])
@pytest.mark.parametrize("lag1,lag2", [(-1024, 1023), (-1025, 1023), (-1024, 1024), (0, 1023), (-1024, -1),
                                       (-7, 7), (-3599, 3599)])
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
