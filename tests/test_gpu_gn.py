"""GNCC and 1-bit GNCC on the GPU against the oracle (src/FFTapps.c:883-1073)
and the committed config-1 goldens."""
import os

import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rand(Tr, N, seed, zeros=0.0):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((Tr, N)).astype(np.float32)
    b = rng.standard_normal((Tr, N)).astype(np.float32)
    if zeros:
        a[rng.random((Tr, N)) < zeros] = 0.0
        b[rng.random((Tr, N)) < zeros] = 0.0
    return a, b


@pytest.mark.parametrize("N,Tr,lag1,lag2", [(3600, 3, -1000, 1000), (100, 2, -99, 99), (1000, 2, 5, 300),
                                            (2401, 1, -7, 0), (640, 4, 60, -20)])
def test_ccgn_set_matches_oracle(fp, orc, N, Tr, lag1, lag2):
    print(f"ccgn_set N={N} Tr={Tr} lags=({lag1},{lag2})", flush=True)
    x1, x2 = _rand(Tr, N, N + 1)
    y = fp.ccgn_set(x1, x2, lag1, lag2)
    yr = orc.run("gncc", x1, x2, lag1, lag2)
    err = np.abs(y - yr).max()
    print(f"  max|gpu-oracle|={err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("N,Tr,lag1,lag2,zeros", [(3600, 3, -1000, 1000, 0.0), (100, 2, -99, 99, 0.0),
                                                  (1000, 2, -300, 300, 0.3), (512, 3, -50, 70, 0.6)])
def test_cc1b_set_matches_oracle(fp, orc, N, Tr, lag1, lag2, zeros):
    """With exact zeros the asymmetric sign rule (x1 >= 0, x2 > 0) changes the
    result by about 2/N per zero, far above the tolerance."""
    print(f"cc1b_set N={N} Tr={Tr} lags=({lag1},{lag2}) zeros={zeros}", flush=True)
    x1, x2 = _rand(Tr, N, N + 2, zeros)
    y = fp.cc1b_set(x1, x2, lag1, lag2)
    yr = orc.run("cc1b", x1, x2, lag1, lag2)
    err = np.abs(y - yr).max()
    print(f"  max|gpu-oracle|={err:.3e}")
    assert err <= TOL
    if zeros:
        # the same data with the symmetric rule would differ: check that it does
        alt = orc.run("cc1b", x2, x1, -lag2, -lag1)[:, ::-1]
        assert np.abs(alt - yr).max() > 10 * TOL


def test_cc1b_zero_records_give_minus_one(fp):
    z = np.zeros((2, 256), np.float32)
    y = fp.cc1b_set(z, z, -5, 5)
    assert np.abs(y + 1).max() < 1e-6


@pytest.mark.parametrize("kind,stack", [("gncc", "none"), ("gncc", "plain"), ("cc1b", "none"),
                                        ("cc1b", "plain"), ("cc1b", "gram")])
def test_allpairs_gn(fp, orc, kind, stack):
    S, W, N = 9, 3, 2000
    x, _ = synth.make_stations(S, W * N, 300, seed=17)
    if kind == "cc1b":
        x[:, ::97] = 0.0  # exact zeros exercise the sign rule
    y = fp.allpairs(x, W, N, -600, 600, kind=kind, stack=stack)
    yr = orc.allpairs(kind, x, W, N, -600, 600, stacked=(stack != "none"))
    err = np.abs(y - yr).max()
    print(f"{kind}/{stack} max|gpu-oracle|={err:.3e}")
    assert y.shape == yr.shape and err <= TOL


@pytest.mark.parametrize("kind,stack", [("pcc2", "plain"), ("pcc2", "gram"), ("gncc", "plain"), ("cc1b", "plain"),
                                        ("cc1b", "gram")])
def test_golden_config1(fp, kind, stack):
    x = np.load(os.path.join(G, "c1_x.npy"), allow_pickle=False)
    gold = np.load(os.path.join(G, f"c1_{kind}_fp64.npy"), allow_pickle=False)
    y = fp.allpairs(x, 1, 3600, -500, 500, kind=kind, stack=stack)
    err = np.abs(y - gold).max()
    print(f"{kind}/{stack} vs golden (FP64): {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("kind", ["pcc2", "gncc", "cc1b"])
@pytest.mark.parametrize("N,lag1,lag2", [(2000, -1500, -1200), (2000, 1200, 1500), (21600, -15000, -14000),
                                         (21600, 14000, 15000)])
def test_one_sided_lag_windows(fp, orc, kind, N, lag1, lag2):
    """Lag windows wholly on one side of zero and beyond -L (the reference's copy
    loop writes past y there; the window's values are the same), in-LDS and
    two-pass transform lengths."""
    x1, x2 = _rand(2, N, N + lag1)
    fn = {"pcc2": fp.pcc2_set, "gncc": fp.ccgn_set, "cc1b": fp.cc1b_set}[kind]
    y = fn(x1, x2, lag1, lag2)
    yr = orc.run(kind, x1, x2, lag1, lag2)
    err = np.abs(y - yr).max()
    print(f"{kind} N={N} lags ({lag1},{lag2}) max|gpu-oracle|={err:.3e}")
    assert y.shape == yr.shape and err <= TOL


@pytest.mark.parametrize("N", [100, 3000, 4000, 4096])
def test_cc1b_extreme_lags(fp, orc, N):
    """Lags up to N-1 on the in-LDS path (Nz <= 8192): the 1-bit count at an
    overlap of m samples is a small integer divided by m, and the FP32 transform's
    error there would exceed 1e-5; those lags are counted exactly from the window
    edges.  Against the FP64 oracle (the reference's own FP32 transform is 1e-5
    off at these lags)."""
    x1, x2 = _rand(3, N, N + 11)
    for l1, l2 in ((-(N - 1), -(N - N // 10)), (N - N // 10, N - 1), (-(N - 1), N - 1)):
        y = fp.cc1b_set(x1, x2, l1, l2)
        yr = orc.run("cc1b", x1, x2, l1, l2, f64=True)
        err = np.abs(y - yr).max()
        print(f"cc1b N={N} lags ({l1},{l2}) max|gpu-fp64|={err:.3e}", flush=True)
        assert y.shape == yr.shape and err <= TOL


@pytest.mark.parametrize("stack", ["none", "plain", "gram", "gram16"])
def test_cc1b_extreme_lags_allpairs(fp, orc, stack):
    """All pairs of 5 stations x 3 windows of 3000 samples at the far lags, one
    record zero in places: per window (no stack) and summed over windows (plain,
    FP32 Gram and split-FP16 Gram stacks) against the FP64 oracle."""
    S, W, N = 5, 3, 3000
    x = np.random.default_rng(9).standard_normal((S, W * N)).astype(np.float32)
    x[2, 100:700] = 0.0
    for l1, l2 in ((-(N - 1), -(N - 300)), (N - 300, N - 1)):
        y = fp.allpairs(x, W, N, l1, l2, kind="cc1b", stack=stack)
        yr = orc.allpairs("cc1b", x, W, N, l1, l2, stacked=(stack != "none"), f64=True)
        err = np.abs(y.reshape(yr.shape) - yr).max()
        print(f"cc1b/{stack} lags ({l1},{l2}) max|gpu-fp64|={err:.3e}", flush=True)
        assert err <= (1e-4 if stack == "gram16" else TOL)
