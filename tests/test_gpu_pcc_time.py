"""Time-domain PCC1 / PCC-v on the GPU against the oracle (src/FFTapps.c:251-336,
475-783)."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _rand(Tr, N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((Tr, N)).astype(np.float32), rng.standard_normal((Tr, N)).astype(np.float32)


def _close(a, b):
    """Equal NaN pattern (lag = +-N has an empty overlap: 0/0 in both) and close values."""
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb)
    return np.abs(a[~na] - b[~nb]).max() if (~na).any() else 0.0


@pytest.mark.parametrize("N,Tr,lag1,lag2", [(3600, 2, -1000, 1000), (100, 3, -40, 60), (1000, 2, 5, 300),
                                            (100, 2, -120, 130), (64, 1, -64, 0)])
def test_pcc1_set_matches_oracle(fp, orc, N, Tr, lag1, lag2):
    print(f"pcc1_set N={N} Tr={Tr} lags=({lag1},{lag2})", flush=True)
    x1, x2 = _rand(Tr, N, N + 3)
    y = fp.pcc1_set(x1, x2, lag1, lag2)
    yr = orc.run("pcc1", x1, x2, lag1, lag2)
    err = _close(y, yr)
    print(f"  max|gpu-oracle|={err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("v", [1.0, 1.5, 2.0, 3.0])
def test_pcc_set_matches_oracle(fp, orc, v):
    x1, x2 = _rand(2, 1500, int(v * 7))
    y = fp.pcc_set(x1, x2, v, -400, 350)
    yr = orc.run("pccv", x1, x2, -400, 350, v=v)
    err = _close(y, yr)
    print(f"pcc_set v={v} max|gpu-oracle|={err:.3e}")
    assert err <= TOL


def test_pcc_out_of_range_returns_untouched(fp):
    x1, x2 = _rand(1, 50, 1)
    y = fp.pcc1_set(x1, x2, 60, 80)  # Lag1 > N: returns 0 without output
    assert not np.any(y)
    with pytest.raises(fp.FastPCCError) as e:
        fp.pcc1_set(x1, x2, 10, 5)  # L < 0
    assert e.value.code == -1


def test_highlevel_entries_take_analytic_signals(fp, orc):
    x1, x2 = _rand(3, 900, 4)
    z1 = np.array([orc.analytic(r) for r in x1])
    z2 = np.array([orc.analytic(r) for r in x2])
    y1 = fp.pcc1_highlevel(z1, z2, -200, 200)
    assert _close(y1, orc.run("pcc1", x1, x2, -200, 200)) <= TOL
    yv = fp.pcc_highlevel(z1, z2, 1.5, -200, 200)
    assert _close(yv, orc.run("pccv", x1, x2, -200, 200, v=1.5)) <= TOL


@pytest.mark.parametrize("kind,stack", [("pcc1", "none"), ("pcc1", "plain"), ("pccv", "plain")])
def test_allpairs_time(fp, orc, kind, stack):
    S, W, N = 6, 2, 1200
    x, _ = synth.make_stations(S, W * N, 200, seed=23)
    y = fp.allpairs(x, W, N, -300, 300, kind=kind, stack=stack, v=1.5)
    yr = orc.allpairs(kind, x, W, N, -300, 300, stacked=(stack != "none"), v=1.5)
    err = np.abs(y - yr).max()
    print(f"{kind}/{stack} max|gpu-oracle|={err:.3e}")
    assert y.shape == yr.shape and err <= TOL
