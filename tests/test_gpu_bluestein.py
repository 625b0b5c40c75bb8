"""Analytic signals of lengths that are not 7-smooth (serving windows take any
length): Bluestein chirp convolutions on the GPU against the oracle, whose own DFT
is a mixed-radix transform for prime factors up to 61 (N = 858 = 2*3*11*13, an
independent algorithm) and Bluestein beyond (the primes)."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


def _rand(Tr, N, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((Tr, N)).astype(np.float32), rng.standard_normal((Tr, N)).astype(np.float32)


@pytest.mark.parametrize("N,lag", [(11, 5), (858, 300), (1801, 600), (7919, 2000), (10007, 1500), (86399, 2000)])
def test_pcc2_any_length(fp, orc, N, lag):
    assert not _smooth7(N)
    print(f"pcc2 N={N} lags=+-{lag}", flush=True)
    x1, x2 = _rand(2, N, N)
    y = fp.pcc2_set(x1, x2, -lag, lag)
    yr = orc.run("pcc2", x1, x2, -lag, lag)
    y64 = orc.run("pcc2", x1, x2, -lag, lag, f64=True)
    err = np.abs(y - yr).max()
    print(f"  max|gpu-oracle|={err:.3e} max|gpu-fp64|={np.abs(y - y64).max():.3e}")
    assert err <= TOL


def test_pcc2_any_length_delay(fp):
    """A delayed common source peaks at its delay (N = 4099, prime)."""
    x, sh = synth.make_stations(2, 4099, 150, seed=77)
    y = fp.pcc2_set(x[:1], x[1:], -400, 400)
    assert np.argmax(y[0]) - 400 == sh[1] - sh[0]


@pytest.mark.parametrize("kind", ["pcc1", "pccv"])
def test_time_domain_any_length(fp, orc, kind):
    N = 4099
    x1, x2 = _rand(1, N, 3)
    if kind == "pcc1":
        y = fp.pcc1_set(x1, x2, -300, 300)
        yr = orc.run("pcc1", x1, x2, -300, 300)
    else:
        y = fp.pcc_set(x1, x2, 1.5, -300, 300)
        yr = orc.run("pccv", x1, x2, -300, 300, v=1.5)
    err = np.abs(y - yr).max()
    print(f"{kind} N={N} max|gpu-oracle|={err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("stack", ["none", "gram", "gram16"])
def test_allpairs_any_length(fp, orc, stack):
    S, W, N = 5, 2, 2003
    x, _ = synth.make_stations(S, W * N, 300, seed=5)
    y = fp.allpairs(x, W, N, -500, 500, kind="pcc2", stack=stack)
    yr = orc.allpairs("pcc2", x, W, N, -500, 500, stacked=(stack != "none"))
    err = np.abs(y - yr).max()
    print(f"pcc2/{stack} N={N} max|gpu-oracle|={err:.3e}")
    assert y.shape == yr.shape and err <= TOL
