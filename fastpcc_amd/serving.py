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


def job_for(req, row_len):
    kw = WPCC_PARAMS if req["kind"] == "wpcc" else {}
    return make_job(req["kind"], req["S"], 1, req["N"], -req["lag"], req["lag"], stack="none",
                    station_stride=row_len, **kw)


def request_input(net, req):
    """The request's input as a view into the network: rows of the network's
    length starting at (s0, t0); flattened from its first sample to the last
    station's window end."""
    flat = net.reshape(-1)
    start = req["s0"] * net.shape[1] + req["t0"]
    end = (req["s0"] + req["S"] - 1) * net.shape[1] + req["t0"] + req["N"]
    return flat[start:end]


def run(reqs, net, server=None, realtime=True, **server_kw):
    """Submits every request at its arrival time (realtime) or all at once, waits
    for all, and returns (server stats, per-request records, outputs)."""
    own = server is None
    srv = server or Server(**server_kw)
    recs = [None] * len(reqs)
    ys = [None] * len(reqs)
    ids = [None] * len(reqs)
    done = threading.Event()

    def waiter():
        for k in range(len(reqs)):
            while ids[k] is None:
                time.sleep(0.0002)
            st = srv.wait(ids[k])
            recs[k] = st
        done.set()

    th = threading.Thread(target=waiter, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for k, rq in enumerate(reqs):
        if realtime:
            dt = rq["t"] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
        job = job_for(rq, net.shape[1])
        y = np.empty(int(np.prod(output_shape(job))), dtype=np.float32)
        ys[k] = y
        ids[k] = srv.submit(job, request_input(net, rq), y)
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
    kinds = {k: sum(1 for r in reqs if r["kind"] == k) for k in KINDS}
    return {
        "requests": len(reqs), "arrival_rate_per_s": rate, "completed": int(lat["completed"]),
        "failed": int(lat["failed"]), "p50_ms": round(lat["p50_ms"], 3), "p99_ms": round(lat["p99_ms"], 3),
        "mean_ms": round(lat["mean_ms"], 3), "max_ms": round(lat["max_ms"], 3),
        "pair_correlations_per_s": round(lat["pair_correlations"] / max(lat["wall_s"], 1e-9), 1),
        "pair_correlations": int(lat["pair_correlations"]), "wall_s": round(lat["wall_s"], 3),
        "inflight_peak": int(lat["inflight_peak"]), "device_peak_GB": round(lat["device_peak"] / 1e9, 3),
        "mix": kinds,
    }
