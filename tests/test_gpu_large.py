"""Windows beyond one workgroup's LDS (N > 8192 or Nz > 8192): the two-pass FFT
path against the oracle.  N = 21600 is the reference's own example size (day
files at 4 s, examples/G/*), giving Nz = 32768 for lags of +-3000."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _rand(Tr, N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((Tr, N)).astype(np.float32), rng.standard_normal((Tr, N)).astype(np.float32)


@pytest.mark.parametrize("kind", ["pcc2", "gncc", "cc1b"])
def test_reference_example_size(fp, orc, kind):
    N, Tr, l1, l2 = 21600, 2, -3000, 3000
    print(f"{kind} N={N} Tr={Tr} lags=({l1},{l2})", flush=True)
    x1, x2 = _rand(Tr, N, 21600)
    fn = {"pcc2": fp.pcc2_set, "gncc": fp.ccgn_set, "cc1b": fp.cc1b_set}[kind]
    y = fn(x1, x2, l1, l2)
    yr = orc.run(kind, x1, x2, l1, l2)
    y64 = orc.run(kind, x1, x2, l1, l2, f64=True)
    err = np.abs(y - yr).max()
    print(f"  max|gpu-oracle|={err:.3e} max|gpu-fp64|={np.abs(y - y64).max():.3e}")
    assert err <= TOL


def test_pcc1_large_window(fp, orc):
    x1, x2 = _rand(1, 21600, 5)
    y = fp.pcc1_set(x1, x2, -400, 400)
    yr = orc.run("pcc1", x1, x2, -400, 400)
    err = np.abs(y - yr).max()
    print(f"pcc1 N=21600 max|gpu-oracle|={err:.3e}")
    assert err <= TOL


def test_pcc2_one_day_at_1hz(fp, orc):
    """86400 samples, lags +-2000: Nz = 131072 (the longest serving window)."""
    x, sh = synth.make_stations(2, 86400, 1000, seed=9)
    y = fp.pcc2_set(x[:1], x[1:], -2000, 2000)
    yr = orc.run("pcc2", x[:1], x[1:], -2000, 2000)
    err = np.abs(y - yr).max()
    print(f"pcc2 N=86400 max|gpu-oracle|={err:.3e}")
    assert err <= TOL
    assert np.argmax(y[0]) - 2000 == sh[1] - sh[0]


def test_gncc_prime_length(fp, orc):
    """GNCC and 1-bit need no DFT of length N: any N works."""
    x1, x2 = _rand(2, 10007, 11)
    for kind, fn in (("gncc", fp.ccgn_set), ("cc1b", fp.cc1b_set)):
        y = fn(x1, x2, -700, 900)
        assert np.abs(y - orc.run(kind, x1, x2, -700, 900)).max() <= TOL


@pytest.mark.parametrize("kind,stack", [("pcc2", "none"), ("pcc2", "plain"), ("pcc2", "gram"), ("gncc", "none"),
                                        ("gncc", "plain"), ("cc1b", "plain"), ("cc1b", "gram"), ("pcc1", "plain")])
def test_allpairs_large(fp, orc, kind, stack):
    S, W, N = 5, 2, 12000
    x, _ = synth.make_stations(S, W * N, 500, seed=41)
    if kind == "cc1b":
        x[:, ::101] = 0.0
    lag = 300 if kind == "pcc1" else 2000
    y = fp.allpairs(x, W, N, -lag, lag, kind=kind, stack=stack)
    yr = orc.allpairs(kind, x, W, N, -lag, lag, stacked=(stack != "none"))
    err = np.abs(y - yr).max()
    print(f"{kind}/{stack} N={N} max|gpu-oracle|={err:.3e}")
    assert y.shape == yr.shape and err <= TOL


@pytest.mark.parametrize("kind,N,lag", [("pcc2", 140000, 3000), ("gncc", 140000, 3000), ("cc1b", 140000, 3000),
                                        ("gncc", 86400, 2000), ("pcc2", 300000, 1000)])
def test_compile_time_four_step_splits(fp, orc, kind, N, lag):
    """Transform lengths whose splits take the in-place column (256 points) and
    row (512 / 1024 points) passes: PCC2 at Nz 262144 (256 x 1024) and 524288
    (256 x 2048, generic rows), GNCC / 1-bit half lengths 131072 (256 x 512) and
    65536 (128 x 512)."""
    x1, x2 = _rand(1, N, N + lag)
    fn = {"pcc2": fp.pcc2_set, "gncc": fp.ccgn_set, "cc1b": fp.cc1b_set}[kind]
    print(f"{kind} N={N} lags +-{lag}", flush=True)
    y = fn(x1, x2, -lag, lag)
    yr = orc.run(kind, x1, x2, -lag, lag)
    err = np.abs(y - yr).max()
    print(f"  max|gpu-oracle|={err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("N", [8000, 10007, 21600])
@pytest.mark.parametrize("kind", ["gncc", "cc1b"])
def test_extreme_lags_two_pass(fp, orc, kind, N):
    """Lags up to N-1 on the two-pass path (Nz = 2^ceil(log2(2N-1)) > 8192): the
    normalisation divides by the energy (or count) of a short overlap, where the
    FP32 transform's error would exceed 1e-5; those lags are summed exactly.  GNCC
    against the oracle (FP64 like the reference); 1-bit against the FP64 oracle
    (the reference's FP32 transform itself is 1e-5 off at the edges)."""
    x1, x2 = _rand(2, N, N + 7)
    fn = {"gncc": fp.ccgn_set, "cc1b": fp.cc1b_set}[kind]
    for l1, l2 in ((-(N - 1), -(N - 400)), (N - 400, N - 1), (-(N - 1), N - 1)):
        y = fn(x1, x2, l1, l2)
        yr = orc.run(kind, x1, x2, l1, l2, f64=(kind == "cc1b"))
        err = np.abs(y - yr).max()
        print(f"{kind} N={N} lags ({l1},{l2}) max|gpu-{'fp64' if kind == 'cc1b' else 'oracle'}|={err:.3e}")
        assert err <= TOL


@pytest.mark.parametrize("kind,stack", [("gncc", "none"), ("gncc", "plain"), ("cc1b", "none"), ("cc1b", "plain"),
                                        ("cc1b", "gram"), ("cc1b", "gram16")])
def test_extreme_lags_allpairs(fp, orc, kind, stack):
    """All pairs of 4 stations x 3 windows of 9000 samples, lags -(N-1) .. -(N-300)
    and the far positive end: the exact short-overlap sums per window, per window
    then averaged (GNCC), and summed over windows (1-bit plain and Gram)."""
    S, W, N = 4, 3, 9000
    x, _ = synth.make_stations(S, W * N, 300, seed=77)
    for l1, l2 in ((-(N - 1), -(N - 300)), (N - 300, N - 1)):
        y = fp.allpairs(x, W, N, l1, l2, kind=kind, stack=stack)
        yr = orc.allpairs(kind, x, W, N, l1, l2, stacked=(stack != "none"), f64=True)
        err = np.abs(y.reshape(yr.shape) - yr).max()
        print(f"{kind}/{stack} lags ({l1},{l2}) max|gpu-fp64|={err:.3e}")
        tol = 1e-4 if stack == "gram16" else TOL
        assert err <= tol
