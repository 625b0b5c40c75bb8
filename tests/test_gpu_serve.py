"""Request server on the GPU: each result is the same alone and inside a mix of
concurrent requests (the same kernels, so bit-identical), admission by memory
footprint, cancellation, timeouts and injected failures, with no device memory
left reserved afterwards."""
import time

import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu


def _mix():
    """Uneven jobs: kinds, station counts, window lengths (7-smooth and not, LDS
    and two-pass paths) and lag windows."""
    rng = np.random.default_rng(2024)
    specs = [("pcc2", 6, 3600, 500, "none"), ("gncc", 5, 5000, 800, "none"), ("cc1b", 7, 4001, 700, "plain"),
             ("pcc2", 4, 20011, 1500, "none"), ("wpcc", 3, 3000, 400, "none"), ("pcc1", 4, 2500, 200, "none"),
             ("pcc2", 9, 2400, 300, "gram16"), ("gncc", 3, 30000, 2000, "none"), ("pcc2", 5, 1801, 600, "none"),
             ("cc1b", 4, 12000, 1000, "none")]
    out = []
    for k, (kind, S, N, lag, stack) in enumerate(specs):
        W = 2 if stack in ("plain", "gram16") else 1
        x, _ = synth.make_stations(S, W * N + 37, min(lag, 300), seed=100 + k)
        kw = dict(pmin=2.0, pmax=200.0, V=8, wtype=-1) if kind == "wpcc" else {}
        out.append((kind, x, W, N, lag, stack, kw))
    return out


def _job(fp, kind, x, W, N, lag, stack, kw):
    return fp.make_job(kind, x.shape[0], W, N, -lag, lag, stack=stack, station_stride=x.shape[1], **kw)


def test_mix_invariance(fp):
    mix = _mix()
    alone = [fp.allpairs(x, W, N, -lag, lag, kind=kind, stack=stack, **kw) for kind, x, W, N, lag, stack, kw in mix]
    with fp.Server(device_bytes=4 << 30, pinned_bytes=256 << 20, max_inflight=8, n_streams=4) as srv:
        ys, ids = [], []
        for r, (kind, x, W, N, lag, stack, kw) in enumerate(mix):
            job = _job(fp, kind, x, W, N, lag, stack, kw)
            n = int(np.prod(fp.output_shape(job)))
            if r % 2:  # pinned buffers take the DMA path, pageable ones the staging arena
                xp = fp.pinned_empty(x.shape)
                xp[:] = x
                y = fp.pinned_empty(n)
            else:
                xp, y = x, np.empty(n, dtype=np.float32)
            ys.append(y)
            ids.append(srv.submit(job, xp, y))
        for rid in ids:
            assert srv.wait(rid)["state"] == 2
        st = srv.stats()
        print(f"mix: {st['completed']} done, inflight peak {st['inflight_peak']}, p50 {st['p50_ms']:.2f} ms, "
              f"p99 {st['p99_ms']:.2f} ms", flush=True)
        assert st["completed"] == len(mix) and st["device_used"] == 0
        assert st["inflight_peak"] >= 2  # requests overlapped on the device
    for (kind, *_), a, y in zip(mix, alone, ys):
        d = np.abs(a.ravel() - y).max()
        print(f"  {kind}: max|alone - mix| = {d:.3e}")
        assert d == 0.0


def test_admission_by_footprint(fp):
    x, _ = synth.make_stations(8, 8000, 200, seed=3)
    job = fp.make_job("pcc2", 8, 1, 8000, -1000, 1000, stack="none", station_stride=8000)
    need = fp.api._lib.lib().fpcc_workspace_size(__import__("ctypes").byref(job))
    # room for one request at a time, not two
    budget = int(1.3 * (need + x.nbytes + 28 * 2001 * 4)) + (1 << 16)
    with fp.Server(device_bytes=budget, pinned_bytes=64 << 20, max_inflight=8) as srv:
        ys = [np.empty(28 * 2001, np.float32) for _ in range(4)]
        ids = [srv.submit(job, x, y) for y in ys]
        for rid in ids:
            assert srv.wait(rid)["state"] == 2
        st = srv.stats()
        assert st["inflight_peak"] == 1 and st["device_used"] == 0 and st["device_peak"] <= budget
        ref = fp.allpairs(x, 1, 8000, -1000, 1000, kind="pcc2", stack="none").ravel()
        assert all(np.array_equal(ref, y) for y in ys)
        # a job that can never fit is refused at submission
        x2, _ = synth.make_stations(64, 8000, 200, seed=4)
        job2 = fp.make_job("pcc2", 64, 1, 8000, -1000, 1000, stack="none", station_stride=8000)
        with pytest.raises(fp.FastPCCError) as e:
            srv.submit(job2, x2, np.empty(2016 * 2001, np.float32))
        assert e.value.code == -4 and srv.stats()["rejected"] == 1


def test_cancel_timeout_and_failures(fp):
    x, _ = synth.make_stations(16, 16000, 200, seed=5)
    job = fp.make_job("gncc", 16, 1, 16000, -2000, 2000, stack="none", station_stride=16000)
    n = 120 * 4001
    with fp.Server(device_bytes=2 << 30, pinned_bytes=128 << 20, max_inflight=1, n_streams=1, fail_every=3) as srv:
        ys = [np.full(n, np.nan, np.float32) for _ in range(6)]
        ids = [srv.submit(job, x, y) for y in ys]
        assert srv.cancel(ids[5]) == 0  # still queued behind max_inflight = 1
        r4 = srv.wait(ids[4], timeout_ms=0)  # returns at once: None while still pending
        assert r4 is None or r4["state"] in (2, 3, 4)
        states = [srv.wait(rid)["state"] for rid in ids]
        # ids 3 and 6 hit the injected failure; id 6 was cancelled first
        assert states == [2, 2, 3, 2, 2, 4], states
        assert srv.poll(ids[2])["rc"] == -1
        ref = fp.allpairs(x, 1, 16000, -2000, 2000, kind="gncc", stack="none").ravel()
        for k in (0, 1, 3, 4):
            assert np.array_equal(ys[k], ref)
        assert np.isnan(ys[5]).all()
        st = srv.stats()
        assert (st["completed"], st["failed"], st["cancelled"]) == (4, 1, 1) and st["device_used"] == 0
        assert srv.forget(ids[0]) == 0 and srv.forget(ids[0]) != 0


# ---- resident station spectra ----

def _network(S, T, seed):
    x, _ = synth.make_stations(S, T, 300, seed=seed)
    return x


CACHE_CASES = [("pcc2", 3600, 500, "none", {}), ("pcc2", 2003, 500, "plain", {}), ("gncc", 5000, 800, "none", {}),
               ("cc1b", 12000, 1000, "none", {}), ("pcc1", 2500, 200, "none", {}),
               ("wpcc", 3000, 400, "none", dict(pmin=2.0, pmax=200.0, V=8, wtype=-1)),
               ("pcc2", 2400, 300, "gram16", {})]


@pytest.mark.parametrize("kind,N,lag,stack,kw", CACHE_CASES)
def test_cache_reuse_matches_uncached(fp, kind, N, lag, stack, kw):
    """Requests over overlapping station sets of one network: only unseen stations
    are computed, and every result equals the request run alone."""
    W = 2 if stack in ("plain", "gram16") else 1
    net = _network(16, W * N, 11)
    # overlap, a full repeat, single missing stations, and one station named twice
    sets = [list(range(0, 8)), list(range(4, 12)), list(range(0, 8)), [3, 12, 5, 6], [13, 13, 9, 2]]
    with fp.Server(device_bytes=4 << 30, pinned_bytes=256 << 20, max_inflight=4, n_streams=4,
                   cache_bytes=2 << 30) as srv:
        ys, ids, xs = [], [], []
        for st in sets:
            x = np.ascontiguousarray(net[st])
            job = fp.make_job(kind, len(st), W, N, -lag, lag, stack=stack, station_stride=x.shape[1], **kw)
            y = np.empty(int(np.prod(fp.output_shape(job))), np.float32)
            ids.append(srv.submit(job, x, y, keys=np.array(st, dtype=np.uint64)))
            ys.append(y)
            xs.append(x)
        for rid in ids:
            assert srv.wait(rid)["state"] == 2
        stat = srv.stats()
        # stations computed: 0-7, 8-11, 12, 13 (once, though named twice)
        print(f"{kind}/{stack}: hits {stat['cache_hits']} misses {stat['cache_misses']} "
              f"cache {stat['cache_used'] / 2**20:.1f} MiB", flush=True)
        assert stat["cache_misses"] == 14 and stat["cache_hits"] == 4 + 8 + 3 + 3
        assert stat["device_used"] == 0
    for st, x, y in zip(sets, xs, ys):
        ref = fp.allpairs(x, W, N, -lag, lag, kind=kind, stack=stack, **kw).ravel()
        assert np.array_equal(ref, y), f"{kind} stations {st}: max diff {np.abs(ref - y).max()}"


def test_cache_eviction_small_budget(fp):
    """A cache that holds about one request's stations: entries in use stay,
    others are evicted, results stay exact and the cache never exceeds its budget."""
    N, lag = 3600, 400
    net = _network(24, N, 12)
    entry = 8192 * 8  # one station's PCC2 spectrum row (Nz 8192, one window)
    budget = 10 * entry
    rng = np.random.default_rng(3)
    sets = [sorted(rng.choice(24, size=8, replace=False).tolist()) for _ in range(12)]
    with fp.Server(device_bytes=2 << 30, pinned_bytes=128 << 20, max_inflight=3, cache_bytes=budget) as srv:
        ys, ids = [], []
        for st in sets:
            x = np.ascontiguousarray(net[st])
            job = fp.make_job("pcc2", 8, 1, N, -lag, lag, stack="none", station_stride=N)
            y = np.empty(28 * (2 * lag + 1), np.float32)
            ids.append(srv.submit(job, x, y, keys=np.array(st, dtype=np.uint64)))
            ys.append(y)
        for rid in ids:
            assert srv.wait(rid)["state"] == 2
        stat = srv.stats()
        assert stat["cache_used"] <= budget and stat["device_used"] == 0
        print(f"eviction: hits {stat['cache_hits']} misses {stat['cache_misses']}")
    for st, y in zip(sets, ys):
        ref = fp.allpairs(np.ascontiguousarray(net[st]), 1, N, -lag, lag, kind="pcc2", stack="none").ravel()
        assert np.array_equal(ref, y)
