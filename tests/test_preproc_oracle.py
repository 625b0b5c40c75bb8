"""Pins of the oracle's preprocessing restatement (clipping, medians, trace rms,
outlier rule, AveWhite, polarity), each by an independent computation: sorting
instead of selection, hand-computed values, a closed form for whitening an
impulse, and invariants."""
import numpy as np
import pytest

import oracle


def _sorted_median(a):
    s = np.sort(np.asarray(a, dtype=np.float32))
    k = s.size // 2
    if s.size % 2:
        return np.float32(s[k])
    return np.float32((s[k] + s[k - 1]) / np.float32(2))


@pytest.mark.parametrize("n", [1, 2, 3, 10, 11, 1000, 1001, 3600])
def test_medians_equal_sorting(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32)
    x[: n // 3] = np.round(x[: n // 3])  # ties
    assert np.float32(oracle.qabsmedian(x)) == _sorted_median(np.abs(x))
    assert np.float32(oracle.qmedian(x)) == _sorted_median(x)


def test_clipping_known_answer():
    x = np.array([1.0, -2.0, 3.0, -100.0], dtype=np.float32)
    thr = np.float32(4 * np.float32(2.5) / 0.6745)
    assert oracle.clip_threshold(x) == thr
    assert np.array_equal(oracle.clipping(x), np.array([1.0, -2.0, 3.0, -thr], dtype=np.float32))


def test_clipping_properties():
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(5001) * np.where(rng.random(5001) < 0.01, 50, 1)).astype(np.float32)
    thr = oracle.clip_threshold(x)
    assert thr == np.float32(4 * _sorted_median(np.abs(x)) / 0.6745)
    y = oracle.clipping(x)
    keep = np.abs(x) <= thr
    assert np.array_equal(y[keep], x[keep])
    assert np.array_equal(y[~keep], np.where(x[~keep] > 0, thr, -thr).astype(np.float32))


def test_trace_rms():
    rng = np.random.default_rng(4)
    for n in (1, 7, 3600, 86400):
        x = rng.standard_normal(n).astype(np.float32)
        ref = np.float32(np.sqrt(np.sum((x * x).astype(np.float64)) / n))
        got = oracle.trace_rms(x)
        assert abs(int(got.view(np.int32)) - int(ref.view(np.int32))) <= 1


def _filter_numpy(H, Nz, N, n1, n2):
    """The whitening filter built with numpy primitives: piecewise inverse, clamp,
    an 11-tap normalised Blackman window (length parameter N) applied as a
    correlation over the spectrum mirrored at DC and Nyquist, shifted by 5."""
    H = np.asarray(H, dtype=np.float64)
    Nh = H.size
    G = np.empty(Nh)
    G[:n1] = 1 / H[n1]
    G[n1:n2] = 1 / H[n1:n2]
    G[n2:] = 1 / H[n2]
    mn = G[n1:max(n2, n1 + 1)].min()
    mx = min(G[n1:max(n2, n1 + 1)].max(), 1000 * mn)
    G = np.minimum(G, mx) / (mn * Nz)
    m = np.arange(11)
    win = 0.42 - 0.5 * np.cos(2 * np.pi * m / (N - 1)) + 0.08 * np.cos(4 * np.pi * m / (N - 1))
    win /= win.sum()
    ext = np.concatenate([G[10:0:-1], G, G[-2:-12:-1]])  # mirrored without repeating DC / Nyquist
    HS = np.array([np.dot(win, ext[10 + c + 5 - np.arange(11)]) for c in range(Nh)])
    HS[5] = 0.0  # the reference's smoothing loops never write bin 5
    return G, HS


@pytest.mark.parametrize("case", ["ramp", "clamped", "noise"])
def test_white_filter_against_numpy(case):
    Nz, N, n1, n2 = 64, 60, 4, 20
    Nh = Nz // 2 + 1
    if case == "ramp":
        H = np.arange(1, Nh + 1, dtype=np.float64)
    elif case == "clamped":
        H = np.ones(Nh)
        H[9] = 1e-5  # 1/H far above 1000 x the minimum gain
    else:
        H = np.random.default_rng(5).random(Nh) + 0.1
    G, HS = oracle.white_filter(H, Nz, N, n1, n2)
    G2, HS2 = _filter_numpy(H, Nz, N, n1, n2)
    np.testing.assert_allclose(G, G2, rtol=1e-13)
    np.testing.assert_allclose(HS, HS2, rtol=1e-12, atol=1e-300)
    assert HS[5] == 0.0
    if case == "clamped":
        assert G.max() == pytest.approx(1000 * G[n1:n2].min(), rel=1e-13)


@pytest.mark.parametrize("N", [1024, 1000])
def test_avewhite_impulse_closed_form(N):
    """An impulse has a flat amplitude spectrum, so the filter is 1/Nz in every
    bin but the unwritten bin 5: the output is the impulse over Nz minus bin 5's
    component, y[n] = d[n-n0]/Nz - 2 cos(2 pi 5 (n-n0)/Nz) / Nz^2."""
    Nz, n0 = 1024, 300
    x = np.zeros((1, N), dtype=np.float32)
    x[0, n0] = 1.0
    y = oracle.avewhite(x, (0.05, 0.3), 1.0)[0]
    n = np.arange(N)
    ref = (n == n0) / Nz - 2 * np.cos(2 * np.pi * 5 * (n - n0) / Nz) / Nz ** 2
    np.testing.assert_allclose(y, ref, atol=1e-9)


def test_avewhite_homogeneous_and_flattening():
    """The filter is normalised by its minimum gain, so it does not depend on the
    input's scale and the output scales with the input (exactly, for a power of
    two); a filter missing that normalisation would scale quadratically."""
    rng = np.random.default_rng(6)
    b = np.convolve(rng.standard_normal(4 * 3000), np.hanning(9), mode="same").astype(np.float32).reshape(4, 3000)
    y1 = oracle.avewhite(b, (0.02, 0.4), 1.0)
    y2 = oracle.avewhite(b * np.float32(64.0), (0.02, 0.4), 1.0)
    np.testing.assert_array_equal(y2, y1 * np.float32(64.0))
    # the averaged amplitude spectrum inside the band is much flatter after whitening
    Nz = 4096
    A0 = np.abs(np.fft.rfft(b.astype(np.float64), Nz)).mean(0)[100:1500]
    A1 = np.abs(np.fft.rfft(y1.astype(np.float64), Nz)).mean(0)[100:1500]
    assert A1.std() / A1.mean() < 0.5 * A0.std() / A0.mean()


def test_preprocess_rules():
    rng = np.random.default_rng(7)
    S, W, N = 3, 6, 500
    x = rng.standard_normal((S, W * N)).astype(np.float32)
    x[0, 2 * N:3 * N] = 0.0                 # a zero window
    x[1, 4 * N:5 * N] *= 20.0               # a loud window
    x[2, 1 * N + 7] = 30.0                  # a spike, clipped (rms 1.7: kept)
    sign = np.ones((S, W), dtype=np.float32)
    sign[2, 3] = -1.0
    y, valid, rms = oracle.preprocess(x, W, N, clip=True, std=3.0, sign=sign)
    assert not valid[0, 2] and not valid[1, 4] and valid.sum() == S * W - 2
    assert np.all(y[0, 2 * N:3 * N] == 0) and np.all(y[1, 4 * N:5 * N] == 0)
    # rms before clipping, median rule of the remaining windows
    assert rms[2, 1] == oracle.trace_rms(x[2, N:2 * N])
    assert rms[1, 4] > 3.0 * oracle.qmedian(rms[1][valid[1] | (np.arange(W) == 4)])
    assert np.abs(y[2, N:2 * N]).max() == oracle.clip_threshold(x[2, N:2 * N])
    # polarity: the flipped window is the negated clipped window
    assert np.array_equal(y[2, 3 * N:4 * N], -oracle.clipping(x[2, 3 * N:4 * N]))
    yc, vc, _ = oracle.preprocess(x, W, N, clip=True, std=3.0, common=True)
    assert np.array_equal(vc, np.repeat(valid.all(0, keepdims=True), S, 0))
    assert np.all(yc[2, 2 * N:3 * N] == 0)


def test_preprocess_whitening_uses_kept_windows():
    rng = np.random.default_rng(8)
    S, W, N = 2, 5, 800
    x = rng.standard_normal((S, W * N)).astype(np.float32)
    x[1, 0:N] = 0.0
    y, valid, _ = oracle.preprocess(x, W, N, white=(0.05, 0.3), dt=1.0)
    for s in range(S):
        kept = [w for w in range(W) if valid[s, w]]
        ref = oracle.avewhite(x[s].reshape(W, N)[kept], (0.05, 0.3), 1.0)
        np.testing.assert_array_equal(y[s].reshape(W, N)[kept], ref)
