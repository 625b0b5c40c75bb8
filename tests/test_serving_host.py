"""Serving workload generator (host logic, no GPU): reproducible, in range, with
the stated arrival rate and type mix, and request inputs that are in-place views
of the network rows."""
import math

import numpy as np

from fastpcc_amd import serving


def test_requests_reproducible_and_in_range():
    a = serving.make_requests(3000, 50.0, seed=7)
    b = serving.make_requests(3000, 50.0, seed=7)
    assert a == b
    for r in a:
        assert 4 <= r["S"] <= 128 and 1800 <= r["N"] <= 86400
        assert 200 <= r["lag"] <= 2000 or r["lag"] == r["N"] - 1
        assert r["lag"] < r["N"] and 0 <= r["s0"] <= 128 - r["S"] and 0 <= r["t0"] <= 86400 - r["N"]
    t = np.array([r["t"] for r in a])
    assert np.all(np.diff(t) > 0)
    assert abs(t[-1] / len(a) - 1 / 50.0) < 0.002  # mean gap 20 ms
    frac = [sum(r["kind"] == k for r in a) / len(a) for k in serving.KINDS]
    assert np.allclose(frac, serving.MIX, atol=0.03)
    logn = np.log([r["N"] for r in a])  # log-uniform lengths
    assert abs(np.median(logn) - 0.5 * (math.log(1800) + math.log(86400))) < 0.1


def test_request_input_is_a_view_of_the_network_rows():
    net = np.arange(10 * 50, dtype=np.float32).reshape(10, 50)
    r = dict(S=3, N=7, s0=4, t0=11, kind="gncc", lag=3)
    x = serving.request_input(net, r)
    job = serving.job_for(r, net.shape[1])
    assert job.station_stride == 50 and job.n_stations == 3 and job.win_len == 7 and job.n_windows == 1
    assert np.shares_memory(x, net) and x.size == 2 * 50 + 7
    for s in range(3):
        assert np.array_equal(x[s * 50: s * 50 + 7], net[4 + s, 11:18])
