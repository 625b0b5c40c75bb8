"""WPCC on the GPU against the FP64 oracle (src/FFTapps.c:1076-1255 restated with
the filter bank of src/FWTa/wavelet_def_v7.c)."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _rand(Tr, N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((Tr, N)).astype(np.float32), rng.standard_normal((Tr, N)).astype(np.float32)


@pytest.mark.parametrize("N,Tr,l1,l2,pmin,pmax,V,wt", [
    (120, 2, -15, 15, 4.0, 16.0, 2, -1),
    (120, 2, -15, 15, 4.0, 16.0, 2, -3),
    (300, 1, -40, 25, 4.0, 30.0, 3, -2),
    (2000, 2, -300, 300, 5.0, 80.0, 4, -1),
    (1000, 1, 50, -20, 6.0, 40.0, 2, -3),
])
def test_tspcc2_set_matches_oracle(fp, orc, N, Tr, l1, l2, pmin, pmax, V, wt):
    print(f"tspcc2_set N={N} Tr={Tr} lags=({l1},{l2}) p=({pmin},{pmax}) V={V} type={wt}", flush=True)
    x1, x2 = _rand(Tr, N, N + V)
    y = fp.tspcc2_set(x1, x2, l1, l2, pmin, pmax, V, wt)
    yr = orc.run("wpcc", x1, x2, l1, l2, f64=True, pmin=pmin, pmax=pmax, V=V, wtype=wt)
    err = np.abs(y - yr).max()
    print(f"  max|gpu-fp64 oracle|={err:.3e} peak={np.abs(yr).max():.3e}")
    assert y.shape == yr.shape and err <= TOL


def test_wpcc_config4_pair(fp, orc):
    """One pair at the 32-station WPCC configuration's shape: 86400 samples, lags
    +-2000, Morlet, periods 2-200 samples, 8 voices (S = 56, Nz = 131072)."""
    x, sh = synth.make_stations(2, 86400, 1000, seed=5)
    y = fp.tspcc2_set(x[:1], x[1:], -2000, 2000, 2.0, 200.0, 8, -1)
    yr = orc.run("wpcc", x[:1], x[1:], -2000, 2000, f64=True, pmin=2.0, pmax=200.0, V=8, wtype=-1)
    err = np.abs(y - yr).max()
    print(f"C4 pair max|gpu-fp64 oracle|={err:.3e} peak={np.abs(yr).max():.3e}")
    assert err <= TOL
    assert np.argmax(y[0]) - 2000 == sh[1] - sh[0]


def test_cuda_entry_with_caller_filter_bank(fp, orc):
    N, l1, l2 = 800, -100, 100
    x1, x2 = _rand(2, N, 77)
    f = orc.wpcc_family(N, l1, l2, 5.0, 40.0, 3, -1, 0.0)
    C = np.float32(1.0 / (f["K0"] * N))
    y = fp.cuda_tspcc2_set_freq(x1, x2, f["Nz"], l1, l2, f["scale"], f["wc"], f["Ls"], f["center"], C)
    yr = orc.run("wpcc", x1, x2, l1, l2, f64=True, pmin=5.0, pmax=40.0, V=3, wtype=-1)
    # C is passed as float (as the reference): relative 6e-8 on top of the FP32 path
    assert np.abs(y - yr).max() <= TOL
    with pytest.raises(fp.FastPCCError) as e:
        fp.cuda_tspcc2_set_freq(x1, x2, f["Nz"], l1, l2, f["scale"], f["wc"], f["Ls"], f["center"], C,
                                down_smp=np.full(f["S"], 2))
    assert e.value.code == -2


def test_tspcc2_errors(fp):
    x1, x2 = _rand(1, 200, 3)
    with pytest.raises(fp.FastPCCError) as e:
        fp.tspcc2_set(x1, x2, -10, 10, 4.0, 16.0, 2, 5)
    assert e.value.code == -3
    with pytest.raises(fp.FastPCCError) as e:
        fp.tspcc2_set(x1, x2, -200, 10, 4.0, 16.0, 2, -1)
    assert e.value.code == -3


@pytest.mark.parametrize("stack", ["none", "plain"])
def test_allpairs_wpcc(fp, orc, stack):
    S, W, N = 5, 2, 1000
    x, _ = synth.make_stations(S, W * N, 100, seed=31)
    kw = dict(pmin=4.0, pmax=40.0, V=3, wtype=-1)
    y = fp.allpairs(x, W, N, -150, 150, kind="wpcc", stack=stack, **kw)
    yr = orc.allpairs("wpcc", x, W, N, -150, 150, stacked=(stack != "none"), f64=True, **kw)
    err = np.abs(y - yr).max()
    print(f"wpcc/{stack} max|gpu-fp64 oracle|={err:.3e}")
    assert y.shape == yr.shape and err <= TOL
