"""Serving workload: a Poisson stream of all-pairs jobs of uneven size run
through the request server (fpcc_server_*), reporting latency and throughput.

Jobs are cut from one synthetic network held in pinned host memory: a request
takes S consecutive stations and one window of N samples starting at a random
time (one window per job, N = trace length), so its input rows are read in place
(station_stride = the network's row length).  Outputs go to ordinary (pageable)
numpy arrays, as a Python client would pass them.
"""
import math
import threading
import time

import numpy as np

from . import synth
from .api import Server, make_job, output_shape, pinned_empty

# WPCC requests use the 2-200 s, 8 voices per octave Morlet bank (periods in
# samples at 1 Hz).
WPCC_PARAMS = dict(pmin=2.0, pmax=200.0, V=8, wtype=-1)
KINDS = ("pcc2", "gncc", "cc1b", "wpcc")
MIX = (0.5, 0.2, 0.1, 0.2)


def make_network(n_stations=128, length=86400, max_shift=1000, seed=2025):
    """[n_stations][length] float32 in pinned memory: noise plus a delayed common
    band-limited source (the same model as the fixed configurations)."""
    x, _ = synth.make_stations(n_stations, length, max_shift, seed=seed)
    xp = pinned_empty(x.shape)
    xp[:] = x
    return xp


def make_requests(n, rate=50.0, n_net=128, length=86400, seed=7, s_range=(4, 128), n_range=(1800, 86400),
                  lag_range=(200, 2000), mix=MIX):
    """n request descriptors with exponential inter-arrival gaps (mean 1/rate s),
    stations uniform in s_range, window length log-uniform in n_range, lag half
    width uniform in lag_range (kept below N) and the kind drawn from mix."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for k in range(n):
        S = int(rng.integers(s_range[0], s_range[1] + 1))
        N = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        N = min(max(N, n_range[0]), n_range[1], length)
        lag = min(int(rng.integers(lag_range[0], lag_range[1] + 1)), N - 1)
        kind = KINDS[int(rng.choice(len(KINDS), p=mix))]
        s0 = int(rng.integers(0, n_net - S + 1))
        t0 = int(rng.integers(0, length - N + 1))
        out.append(dict(t=float(t[k]), kind=kind, S=S, N=N, lag=lag, s0=s0, t0=t0))
    return out


def make_network_day_requests(n, rate=200.0, n_net=128, N=21600, lag=1000, seed=11, s_range=(8, 64),
                              mix=(("pcc2", 0.5), ("wpcc", 0.5))):
    """Requests that share one window of the network (the same day, t0 = 0):
    random station subsets, one kind each; their station spectra repeat across
    requests, which the server's resident-spectra cache reuses."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    kinds = [k for k, _ in mix]
    p = [w for _, w in mix]
    out = []
    for k in range(n):
        S = int(rng.integers(s_range[0], s_range[1] + 1))
        st = sorted(rng.choice(n_net, size=S, replace=False).tolist())
        out.append(dict(t=float(t[k]), kind=kinds[int(rng.choice(len(kinds), p=p))], S=S, N=N, lag=lag,
                        stations=st, t0=0))
    return out


def job_for(req, row_len):
    kw = WPCC_PARAMS if req["kind"] == "wpcc" else {}
    return make_job(req["kind"], req["S"], 1, req["N"], -req["lag"], req["lag"], stack="none",
                    station_stride=row_len, **kw)


def request_input(net, req):
    """The request's input as a view into the network: rows of the network's
    length starting at (s0, t0); flattened from its first sample to the last
    station's window end.  Requests naming scattered stations get a gathered copy
    (one row per station)."""
    if "stations" in req:
        return np.ascontiguousarray(net[req["stations"], req["t0"]:req["t0"] + req["N"]]).reshape(-1)
    flat = net.reshape(-1)
    start = req["s0"] * net.shape[1] + req["t0"]
    end = (req["s0"] + req["S"] - 1) * net.shape[1] + req["t0"] + req["N"]
    return flat[start:end]


def station_keys(req, row_len):
    """64-bit identity of each station row of a request: station index, window
    start and length (equal keys mean equal samples)."""
    if "stations" in req:
        st = np.asarray(req["stations"], dtype=np.uint64)
    else:
        st = np.arange(req["s0"], req["s0"] + req["S"], dtype=np.uint64)
    return (st << np.uint64(40)) | (np.uint64(req["t0"]) << np.uint64(20)) | np.uint64(req["N"] & 0xFFFFF)


class OutputPool:
    """Pinned output buffers of one size, handed out per request and returned once
    the client has consumed the result (results then arrive by DMA, without the
    server's staging copy); None when all are in use."""

    def __init__(self, n, floats):
        self.floats = floats
        self.free = [pinned_empty(floats) for _ in range(n)]
        self.lock = threading.Lock()

    def get(self, n):
        if n > self.floats:
            return None
        with self.lock:
            return self.free.pop() if self.free else None

    def put(self, buf):
        with self.lock:
            self.free.append(buf)


def out_len(req):
    P = req["S"] * (req["S"] - 1) // 2
    return P * (2 * req["lag"] + 1)


def run(reqs, net, server=None, realtime=True, keys=False, pool=None, keep=True, keep_idx=(), **server_kw):
    """Submits every request at its arrival time (realtime) or all at once, waits
    for all, and returns (server stats, per-request records, outputs).  keys:
    name each request's stations for the resident-spectra cache; pool: an
    OutputPool for outputs (a pageable array when it is empty); keep: return the
    outputs (pool buffers are recycled when False; the requests listed in
    keep_idx keep theirs either way)."""
    keep_idx = set(keep_idx)
    own = server is None
    srv = server or Server(**server_kw)
    recs = [None] * len(reqs)
    ys = [None] * len(reqs)
    ids = [None] * len(reqs)
    pooled = [None] * len(reqs)
    done = threading.Event()

    def waiter():
        for k in range(len(reqs)):
            while ids[k] is None:
                time.sleep(0.0002)
            st = srv.wait(ids[k])
            recs[k] = st
            if pooled[k] is not None:
                if keep or k in keep_idx:
                    ys[k] = ys[k].copy()
                else:
                    ys[k] = None
                pool.put(pooled[k])
            elif not keep and k not in keep_idx:
                ys[k] = None
        done.set()

    th = threading.Thread(target=waiter, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for k, rq in enumerate(reqs):
        if realtime:
            dt = rq["t"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
        row_len = rq["N"] if "stations" in rq else net.shape[1]
        job = job_for(rq, row_len)
        n = int(np.prod(output_shape(job)))
        buf = pool.get(n) if pool is not None else None
        if buf is not None:
            pooled[k] = buf
            y = buf[:n]
        else:
            y = np.empty(n, dtype=np.float32)
        ys[k] = y
        ids[k] = srv.submit(job, request_input(net, rq), y, keys=station_keys(rq, row_len) if keys else None)
    done.wait()
    wall = time.perf_counter() - t0
    stats = srv.stats()
    stats["wall_s"] = wall
    if own:
        srv.close()
    return stats, recs, ys


def summary(stats, reqs, rate):
    """JSON-ready summary of a serving run."""
    lat = stats
    busy = max(lat.get("busy_s", 0.0), 1e-9)
    kinds = {k: sum(1 for r in reqs if r["kind"] == k) for k in sorted({r["kind"] for r in reqs})}
    return {
        "requests": len(reqs), "arrival_rate_per_s": rate, "completed": int(lat["completed"]),
        "failed": int(lat["failed"]), "p50_ms": round(lat["p50_ms"], 3), "p99_ms": round(lat["p99_ms"], 3),
        "mean_ms": round(lat["mean_ms"], 3), "max_ms": round(lat["max_ms"], 3),
        "pair_correlations_per_s": round(lat["pair_correlations"] / max(lat["wall_s"], 1e-9), 1),
        "pair_correlations_per_busy_s": round(lat["pair_correlations"] / busy, 1),
        "completed_per_s": round(lat["completed"] / max(lat["wall_s"], 1e-9), 2),
        "pair_correlations": int(lat["pair_correlations"]), "wall_s": round(lat["wall_s"], 3),
        "inflight_peak": int(lat["inflight_peak"]), "device_peak_GB": round(lat["device_peak"] / 1e9, 3),
        "cache_hits": int(lat["cache_hits"]), "cache_misses": int(lat["cache_misses"]),
        "mix": {k: v for k, v in kinds.items() if v},
    }
