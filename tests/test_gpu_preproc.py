"""GPU per-station preprocessing (fpcc_preprocess) against the oracle: zero and
outlier windows, rms, clipping and polarity bit-exact; FP64 whitening within a
float ulp-level tolerance; preprocessing followed by all-pairs correlation."""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu


def _network(S, W, N, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((S, W * N)).astype(np.float32)
    spikes = rng.random((S, W * N)) < 0.002
    x[spikes] *= 40.0                                   # clipped by the median rule
    x[0, 1 * N:2 * N] = 0.0                             # a zero window
    x[S - 1, (W - 1) * N:W * N] *= 25.0                 # an outlier window
    x[1 % S, (W - 1) * N + min(5, N - 1)] = -0.0        # a signed zero
    sign = np.where(rng.random((S, W)) < 0.2, -1.0, 1.0).astype(np.float32)
    return x, sign


def _ulps(a, b):
    return np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))


@pytest.mark.parametrize("N", [3600, 3601, 10000, 2])
def test_clip_outliers_exact(fp, orc, N):
    S, W = 4, 6
    print(f"preprocess clip+std N={N}", flush=True)
    x, sign = _network(S, W, N, N)
    y, valid, rms = fp.preprocess(x, W, N, clip=True, std=3.0, sign=sign)
    yr, vr, rr = orc.preprocess(x, W, N, clip=True, std=3.0, sign=sign)
    assert np.array_equal(valid, vr)
    assert _ulps(rms, rr).max() <= 1
    assert np.array_equal(y.view(np.int32), yr.view(np.int32))


def test_common_windows_and_no_clip(fp, orc):
    S, W, N = 3, 8, 1000
    x, sign = _network(S, W, N, 5)
    y, valid, _ = fp.preprocess(x, W, N, std=2.5, common=True, sign=sign)
    yr, vr, _ = orc.preprocess(x, W, N, std=2.5, common=True, sign=sign)
    assert np.array_equal(valid, vr) and np.all(valid == valid[:1])
    assert np.array_equal(y.view(np.int32), yr.view(np.int32))


# 2049..4096 samples take the compile-time 2048-point FP64 kernels; the others the runtime plans
@pytest.mark.parametrize("N,band", [(3600, (0.02, 0.2)), (2049, (0.02, 0.2)), (4096, (0.05, 0.3)), (1000, (0.05, 0.45)),
                                    (8192, (0.01, 0.3)), (64, (0.1, 0.4))])
def test_whitening_matches_oracle(fp, orc, N, band):
    S, W = 3, 5
    print(f"whitening N={N} band={band}", flush=True)
    x, sign = _network(S, W, N, 11 + N)
    y, valid, _ = fp.preprocess(x, W, N, clip=True, std=3.0, white=band, dt=1.0, sign=sign)
    yr, vr, _ = orc.preprocess(x, W, N, clip=True, std=3.0, white=band, dt=1.0, sign=sign)
    assert np.array_equal(valid, vr)
    scale = np.abs(yr).max()
    err = np.abs(y - yr).max()
    print(f"  max|gpu-oracle|/max|oracle| = {err / scale:.2e}")
    assert err <= 1e-6 * scale
    assert np.all(y[~np.repeat(valid, N, axis=1)] == 0)


@pytest.mark.parametrize("N,W", [(9000, 3), (21600, 3), (86400, 2)])
def test_whitening_long_windows(fp, orc, N, W):
    """Nz 16384 to 131072: the FP64 four-step path (21600 samples: the reference's
    day files at 4 s, examples/G)."""
    S = 3
    print(f"whitening four-step N={N}", flush=True)
    x, sign = _network(S, W, N, 3 + N)
    y, valid, _ = fp.preprocess(x, W, N, clip=True, std=3.0, white=(0.02, 0.2), dt=1.0, sign=sign)
    yr, vr, _ = orc.preprocess(x, W, N, clip=True, std=3.0, white=(0.02, 0.2), dt=1.0, sign=sign)
    assert np.array_equal(valid, vr)
    err = np.abs(y - yr).max() / np.abs(yr).max()
    print(f"  max|gpu-oracle|/max|oracle| = {err:.2e}")
    assert err <= 1e-6


def test_whitening_limits(fp):
    x = np.zeros((2, 2 * 600000), dtype=np.float32)
    with pytest.raises(fp.FastPCCError) as e:
        fp.preprocess(x, 2, 600000, white=(0.01, 0.1))   # Nz 2^20: beyond the four-step limit
    assert e.value.code == -2
    with pytest.raises(fp.FastPCCError):
        fp.preprocess(x[:, :2000], 2, 1000, white=(0.1, 0.6))   # band past Nyquist


def test_preprocess_then_correlate(fp, orc):
    """Preprocessed stations correlate as the oracle's preprocessed stations do;
    the per-window outputs of dropped windows are zero for PCC2."""
    S, W, N, lag = 5, 4, 3600, 500
    x, _ = synth.make_stations(S, W * N, 200, seed=77)
    x[2, N:2 * N] = 0.0
    y, valid, _ = fp.preprocess(x, W, N, clip=True, std=3.0, white=(0.02, 0.3), dt=1.0)
    yr, vr, _ = orc.preprocess(x, W, N, clip=True, std=3.0, white=(0.02, 0.3), dt=1.0)
    assert np.array_equal(valid, vr)
    c = fp.allpairs(y, W, N, -lag, lag, kind="pcc2", stack="none")
    cr = orc.allpairs("pcc2", yr, W, N, -lag, lag, stacked=False)
    assert np.abs(c - cr).max() <= 1e-5
    P = S * (S - 1) // 2
    c = c.reshape(P, W, -1)
    p = 0
    for i in range(S):
        for j in range(i + 1, S):
            for w in range(W):
                if not (valid[i, w] and valid[j, w]):
                    assert np.all(c[p, w] == 0)
            p += 1
