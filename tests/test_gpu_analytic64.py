"""FP64 AnalyticSignal (src/FFTapps.c:45-80) on the GPU against scipy's FP64
analytic signal (scipy.signal.hilbert: the same DC/positive/Nyquist weights) and
the oracle's FP64 restatement; FFT lengths (7-smooth, N <= 4096) and direct DFT
lengths (prime or longer)."""
import numpy as np
import pytest
import scipy.signal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [1, 2, 3, 1000, 3600, 4096, 4099, 10007, 21600])
def test_analytic_signal_fp64(fp, orc, N):
    x = np.random.default_rng(N).standard_normal(N)
    y = fp.analytic_signal(x)
    ref = scipy.signal.hilbert(x)
    err = np.abs(y - ref).max()
    print(f"N={N} max|gpu - scipy| = {err:.2e}")
    assert err <= 1e-10 * max(1.0, np.abs(ref).max())
    assert np.allclose(y.real, x, atol=1e-10)  # the real part is the signal itself
    yo = orc.analytic(x, f64=True)
    assert np.abs(y - yo).max() <= 1e-10 * max(1.0, np.abs(ref).max())


def test_analytic_signal_tone(fp):
    """cos(w n) -> exp(i w n) for a bin-centred tone."""
    N, k = 2048, 37
    n = np.arange(N)
    y = fp.analytic_signal(np.cos(2 * np.pi * k * n / N))
    assert np.abs(y - np.exp(2j * np.pi * k * n / N)).max() < 1e-12
