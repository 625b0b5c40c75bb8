"""The benchmark configurations at their own sizes, against the oracle.

- 64 stations x 86400 samples in 24 windows of 3600, nfft 8192, lags +-1000, all
  2016 pairs: GNCC (mean stack) and 1-bit GNCC (mean stack; FP32 and split-FP16
  Gram stacks) against the reference-precision oracle (src/FFTapps.c:883-1073) on
  pair ranges at 64-station block boundaries, the stacks against each other over
  every pair.
- WPCC on 86400 samples with the literal 30 log-spaced Morlet scales between
  periods 2 and 200 samples (ratio 100^(1/29)), a filter bank the caller builds
  and hands to cuda_tspcc2_set_freq (src/ccs_cuda.cu:898-1036), against the FP64
  oracle with the same bank.
- Server requests of the serving mix's largest sizes (128 stations, one window of
  86400 samples; PCC2, GNCC, 1-bit; WPCC with 8 voices over 2-200 samples at
  nfft 131072), submitted together: every served result equals the same job run
  alone bit for bit, and a few pairs match the oracle.
"""
import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu
TOL = 1e-5
TOL16 = 1e-4


def _p(S, i, j):
    return i * S - i * (i + 1) // 2 + (j - i - 1)


def _ranges(S):
    P = S * (S - 1) // 2
    return [(0, 6), (_p(S, 0, 62), _p(S, 0, 63) + 1), (_p(S, 30, 63) - 2, _p(S, 30, 63) + 1), (P - 6, P)]


@pytest.fixture(scope="module")
def c2_network():
    x, sh = synth.make_stations(64, 24 * 3600, 1000, seed=2016)
    return x, sh


def _oracle_ranges(orc, kind, x, W, N, lag, y, f64=False):
    worst = 0.0
    for b, e in _ranges(x.shape[0]):
        yr = orc.allpairs(kind, x, W, N, -lag, lag, stacked=True, f64=f64, pair_begin=b, pair_end=e)
        worst = max(worst, float(np.abs(y[b:e] - yr).max()))
    return worst


def test_c2_gncc_full_size(fp, orc, c2_network):
    x, _ = c2_network
    print("GNCC 64 x 24 x 3600, lags +-1000, mean stack", flush=True)
    y = fp.allpairs(x, 24, 3600, -1000, 1000, kind="gncc", stack="plain")
    assert y.shape == (2016, 2001)
    e_ref = _oracle_ranges(orc, "gncc", x, 24, 3600, 1000, y)
    e64 = _oracle_ranges(orc, "gncc", x, 24, 3600, 1000, y, f64=True)
    print(f"  max|gpu-oracle| = {e_ref:.3e} (reference precision), {e64:.3e} (FP64)", flush=True)
    assert e_ref <= TOL and e64 <= TOL


def test_c2_cc1b_full_size(fp, orc, c2_network):
    x, _ = c2_network
    print("1-bit 64 x 24 x 3600, lags +-1000: plain, gram, gram16", flush=True)
    ys = {st: fp.allpairs(x, 24, 3600, -1000, 1000, kind="cc1b", stack=st) for st in ("plain", "gram", "gram16")}
    e_ref = _oracle_ranges(orc, "cc1b", x, 24, 3600, 1000, ys["plain"])
    e64 = {st: _oracle_ranges(orc, "cc1b", x, 24, 3600, 1000, y, f64=True) for st, y in ys.items()}
    d_gram = float(np.abs(ys["gram"] - ys["plain"]).max())
    d_16 = float(np.abs(ys["gram16"] - ys["plain"]).max())
    print(f"  plain: max|gpu-oracle| = {e_ref:.3e}; vs FP64 {e64}; over all 2016 pairs "
          f"|gram-plain| = {d_gram:.3e}, |gram16-plain| = {d_16:.3e}", flush=True)
    assert e_ref <= TOL and e64["plain"] <= TOL and e64["gram"] <= TOL and e64["gram16"] <= TOL16
    assert d_gram <= TOL and d_16 <= TOL16


def test_c4_literal_30_scales(fp, orc):
    N, lag = 86400, 2000
    w0 = np.pi * np.sqrt(2 / np.log(2))
    s0 = 2.0 * w0 / (2 * np.pi)  # period 2 samples, as tspcc2_set's Morlet s0 (src/FFTapps.c:1100-1106)
    scales = s0 * (100.0 ** (np.arange(30) / 29))  # 30 scales, periods 2 .. 200 samples
    Nz = 131072  # 2^ceil(log2(N + M)); Nz - (N + M) exceeds the longest support, so no doubling
    bank = synth.morlet_bank(scales, Nz)
    assert Nz - (N + lag) >= int(bank["Ls"][-1])
    x, sh = synth.make_stations(2, N, 1000, seed=30)
    C = np.float32(1.0 / (bank["K0"] * N))
    print(f"cuda_tspcc2_set_freq 30 scales {scales[0]:.3f}..{scales[-1]:.2f}, N={N}, Nz={Nz}", flush=True)
    y = fp.cuda_tspcc2_set_freq(x[:1], x[1:], Nz, -lag, lag, bank["scale"], bank["wc"], bank["Ls"], bank["center"], C)
    yr = orc.wpcc_bank(x[:1], x[1:], Nz, -lag, lag, bank["scale"], bank["wc"], bank["Ls"], bank["center"], float(C))
    err = float(np.abs(y - yr).max())
    print(f"  max|gpu-fp64 oracle| = {err:.3e}, peak {np.abs(yr).max():.3e} at lag {np.argmax(yr[0]) - lag}, "
          f"true delay {sh[1] - sh[0]}", flush=True)
    assert err <= TOL
    assert np.argmax(y[0]) - lag == sh[1] - sh[0]


def test_c5_largest_requests_served(fp, orc):
    """One window of 86400 samples at 128 stations per request (the serving mix's
    upper bounds), lags at the mix's extremes; all submitted at once."""
    N = 86400
    wkw = dict(pmin=2.0, pmax=200.0, V=8, wtype=-1)
    specs = [("pcc2", 128, 2000, "none", {}), ("gncc", 128, 1500, "none", {}), ("cc1b", 128, 200, "none", {}),
             ("wpcc", 128, 2000, "none", wkw), ("pcc2", 128, 200, "none", {})]
    jobs = []
    for k, (kind, S, lag, stack, kw) in enumerate(specs):
        x, _ = synth.make_stations(S, N, 1000, seed=500 + k)
        jobs.append((kind, x, lag, stack, kw))
    alone = []
    for kind, x, lag, stack, kw in jobs:
        print(f"alone: {kind} S={x.shape[0]} N={N} lags +-{lag}", flush=True)
        alone.append(fp.allpairs(x, 1, N, -lag, lag, kind=kind, stack=stack, **kw).ravel())
    with fp.Server(n_streams=4, max_inflight=8) as srv:
        ids, ys = [], []
        for kind, x, lag, stack, kw in jobs:
            job = fp.make_job(kind, x.shape[0], 1, N, -lag, lag, stack=stack, **kw)
            y = np.empty(int(np.prod(fp.output_shape(job))), np.float32)
            ys.append(y)
            ids.append(srv.submit(job, x, y))
        for rid in ids:
            assert srv.wait(rid)["state"] == 2
        st = srv.stats()
        print(f"served {st['completed']}: inflight peak {st['inflight_peak']}, device peak "
              f"{st['device_peak'] / 2**30:.1f} GiB, p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms", flush=True)
        assert st["device_used"] == 0
    for (kind, x, lag, stack, kw), a, y in zip(jobs, alone, ys):
        assert np.array_equal(a, y), kind
        S = x.shape[0]
        P = S * (S - 1) // 2
        L = (2 * lag + 1) if kind != "wpcc" else abs(2 * lag + 1)
        for b in (0, _p(S, 63, 64), P - 1):
            yr = orc.allpairs(kind, x, 1, N, -lag, lag, stacked=True, f64=(kind == "wpcc"), pair_begin=b,
                              pair_end=b + 1, **kw)
            err = float(np.abs(y.reshape(P, L)[b] - yr[0]).max())
            print(f"  {kind} lags +-{lag} pair {b}: served = alone, max|gpu-oracle| = {err:.3e}", flush=True)
            assert err <= TOL
