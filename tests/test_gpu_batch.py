"""Uneven jobs in one launch: fpcc_station_stage_batch_device runs the station
stages of GNCC, 1-bit and PCC2 jobs (long windows too) with different station
counts, window counts, window lengths and strides (one Nz per group) as shared
launches; each job's pair stage then gives exactly its standalone result, also
when fpcc_pair_stage_batch_device runs the pair stages of PCC2 jobs at Nz 8192 and
of GNCC and 1-bit jobs in shared launches.  Device memory comes from the HIP runtime the library itself links."""
import ctypes

import numpy as np
import pytest

from fastpcc_amd import synth

pytestmark = pytest.mark.gpu


# (kind, S, W, N, lag, stack): equal Nz within each group
GROUPS = [
    [("gncc", 5, 1, 3000, 800, "none"), ("gncc", 9, 2, 2500, 900, "plain"), ("gncc", 3, 1, 3900, 300, "none")],
    [("cc1b", 4, 1, 20000, 2000, "none"), ("cc1b", 7, 2, 12000, 1500, "plain"), ("cc1b", 2, 1, 16000, 900, "none")],
    [("gncc", 6, 1, 30000, 1000, "none"), ("gncc", 4, 3, 20000, 2000, "plain"), ("pcc2", 5, 1, 3600, 500, "none")],
    # PCC2, Nz 4096: window lengths 3000, 2500, 3840, 2800 (each its own FFT plan) and all stacks
    [("pcc2", 5, 1, 3000, 800, "none"), ("pcc2", 9, 2, 2500, 900, "plain"), ("pcc2", 3, 1, 3840, 200, "gram"),
     ("pcc2", 4, 3, 2800, 1000, "gram16")],
    # PCC2, Nz 8192: the specialised 3600-sample job runs alone, the others share a launch
    [("pcc2", 4, 2, 3600, 1000, "none"), ("pcc2", 6, 1, 6000, 2000, "plain"), ("pcc2", 3, 2, 5000, 1500, "zsum")],
    # long-window PCC2, Nz 32768: analytic signals on four plans (10007 samples: Bluestein), the zero-padded
    # transforms shared
    [("pcc2", 4, 1, 9000, 2000, "none"), ("pcc2", 3, 2, 12000, 3000, "plain"), ("pcc2", 5, 1, 10007, 1000, "none"),
     ("pcc2", 2, 2, 16000, 500, "zsum")],
]


@pytest.mark.parametrize("group", range(len(GROUPS)))
def test_station_stage_batch_exact(fp, hip, group):
    L = fp.api._lib.lib()
    if True:
        specs = GROUPS[group]
        jobs = (fp.api._lib.FpccJob * len(specs))()
        xs, dx, dws, wsb, refs = [], [], [], [], []
        for k, (kind, S, W, N, lag, stack) in enumerate(specs):
            x, _ = synth.make_stations(S, W * N + 11 * k, 200, seed=50 + k)
            job = fp.make_job(kind, S, W, N, -lag, lag, stack=stack, station_stride=x.shape[1])
            jobs[k] = job
            xs.append(x)
            d = hip.alloc(x.nbytes)
            hip.h2d(d, x)
            dx.append(d)
            wsb.append(L.fpcc_workspace_size(ctypes.byref(job)))
            dws.append(hip.alloc(wsb[-1]))
            refs.append(fp.allpairs(x, W, N, -lag, lag, kind=kind, stack=stack))
        n = len(specs)
        r = L.fpcc_station_stage_batch_device(jobs, n, (ctypes.c_void_p * n)(*dx), (ctypes.c_void_p * n)(*dws),
                                              (ctypes.c_size_t * n)(*wsb), None)
        assert r == 0
        for k in range(n):
            y = np.empty(refs[k].shape, np.float32)
            dy = hip.alloc(y.nbytes)
            assert L.fpcc_pair_stage_device(ctypes.byref(jobs[k]), dws[k], wsb[k], dy, None) == 0
            hip.d2h(y, dy)
            print(f"{specs[k]}: max|batch - alone| = {np.abs(y - refs[k]).max():.3e}")
            assert np.array_equal(y, refs[k])


# PCC2 at Nz 8192 (2048 < N <= 4096): uneven stations, windows, window lengths and lags;
# stacks none, plain and zsum share a launch per stack, gram and the others run alone
PAIR_SPECS = [("pcc2", 5, 2, 3000, 800, "none"), ("pcc2", 9, 1, 2500, 900, "none"), ("pcc2", 4, 3, 3600, 1000, "plain"),
              ("pcc2", 6, 2, 4000, 2000, "plain"), ("pcc2", 3, 2, 2100, 300, "zsum"), ("pcc2", 7, 1, 3300, 3000, "none"),
              ("pcc2", 4, 2, 3600, 700, "gram16"), ("pcc2", 2, 4, 2900, 1000, "zsum")]


# GNCC and 1-bit: Nz 8192 (compile-time inverse) and 4096 (runtime plan), stacks none
# and plain, uneven lags (1-bit up to the window edges, from the exact edge sums)
GN_PAIR_SPECS = [("gncc", 5, 2, 3000, 800, "none"), ("gncc", 4, 1, 2500, 2400, "none"), ("gncc", 6, 2, 3900, 300, "plain"),
                 ("gncc", 3, 3, 2200, 1000, "plain"), ("gncc", 4, 2, 1500, 700, "none"), ("gncc", 5, 1, 1200, 1100, "none"),
                 ("cc1b", 4, 2, 3000, 2999, "none"), ("cc1b", 6, 1, 2600, 500, "none"), ("cc1b", 3, 2, 3500, 900, "plain"),
                 ("cc1b", 5, 2, 2300, 1500, "plain"), ("gncc", 4, 1, 3000, 600, "pws")]


@pytest.mark.parametrize("specs", [PAIR_SPECS, GN_PAIR_SPECS], ids=["pcc2", "gncc_1bit"])
def test_pair_stage_batch_exact(fp, hip, specs):
    """fpcc_pair_stage_batch_device: every job's output equals its standalone run,
    bit for bit, whether its pair stage shared a launch or not."""
    L = fp.api._lib.lib()
    n = len(specs)
    jobs = (fp.api._lib.FpccJob * n)()
    dws, wsb, dys, refs = [], [], [], []
    for k, (kind, S, W, N, lag, stack) in enumerate(specs):
        x, _ = synth.make_stations(S, W * N + 7 * k, 200, seed=110 + k)
        job = fp.make_job(kind, S, W, N, -lag, lag, stack=stack, station_stride=x.shape[1])
        jobs[k] = job
        dx = hip.alloc(x.nbytes)
        hip.h2d(dx, x)
        wsb.append(L.fpcc_workspace_size(ctypes.byref(job)))
        dws.append(hip.alloc(wsb[-1]))
        assert L.fpcc_station_stage_device(ctypes.byref(jobs[k]), dx, dws[-1], wsb[-1], None) == 0
        refs.append(fp.allpairs(x, W, N, -lag, lag, kind=kind, stack=stack))
        dys.append(hip.alloc(refs[-1].nbytes))
    r = L.fpcc_pair_stage_batch_device(jobs, n, (ctypes.c_void_p * n)(*dws), (ctypes.c_size_t * n)(*wsb),
                                       (ctypes.c_void_p * n)(*dys), None)
    assert r == 0
    for k in range(n):
        y = np.empty(refs[k].shape, np.float32)
        hip.d2h(y, dys[k])
        print(f"{specs[k]}: max|batch - alone| = {np.abs(y - refs[k]).max():.3e}")
        assert np.array_equal(y, refs[k])


def test_server_groups_uneven_jobs(fp):
    """A burst of uneven GNCC requests with one Nz, queued together, runs as one
    group; every result equals the standalone run."""
    specs = [(5, 3000, 800), (9, 2500, 900), (3, 3900, 300), (6, 3500, 1000), (4, 2000, 1500)]
    items, refs = [], []
    for k, (S, N, lag) in enumerate(specs):
        x, _ = synth.make_stations(S, N, 200, seed=70 + k)
        job = fp.make_job("gncc", S, 1, N, -lag, lag, stack="none", station_stride=N)
        y = np.empty(S * (S - 1) // 2 * (2 * lag + 1), np.float32)
        items.append((job, x, y))
        refs.append(fp.allpairs(x, 1, N, -lag, lag, kind="gncc", stack="none").ravel())
    with fp.Server(device_bytes=2 << 30, pinned_bytes=128 << 20, max_inflight=16) as srv:
        ids = srv.submit_many(items)
        for rid in ids:
            assert srv.wait(rid)["state"] == 2
        st = srv.stats()
        print(f"batched {st['batched_requests']} in {st['batch_launches']} groups")
        assert st["batched_requests"] >= 2 and st["batch_launches"] >= 1 and st["device_used"] == 0
    for (job, x, y), ref in zip(items, refs):
        assert np.array_equal(y, ref)


def test_server_groups_uneven_pcc2(fp):
    """The same for a burst of PCC2 requests of different window lengths."""
    specs = [(5, 3000, 800), (9, 2500, 900), (3, 3840, 200), (6, 2800, 1000), (4, 2000, 1500)]
    items, refs = [], []
    for k, (S, N, lag) in enumerate(specs):
        x, _ = synth.make_stations(S, N, 200, seed=90 + k)
        job = fp.make_job("pcc2", S, 1, N, -lag, lag, stack="none", station_stride=N)
        y = np.empty(S * (S - 1) // 2 * (2 * lag + 1), np.float32)
        items.append((job, x, y))
        refs.append(fp.allpairs(x, 1, N, -lag, lag, kind="pcc2", stack="none").ravel())
    with fp.Server(device_bytes=2 << 30, pinned_bytes=128 << 20, max_inflight=16) as srv:
        ids = srv.submit_many(items)
        for rid in ids:
            assert srv.wait(rid)["state"] == 2
        st = srv.stats()
        print(f"batched {st['batched_requests']} in {st['batch_launches']} groups")
        assert st["batched_requests"] >= 2 and st["batch_launches"] >= 1
    for (job, x, y), ref in zip(items, refs):
        assert np.array_equal(y, ref)


# WPCC jobs: uneven station counts, window counts and lengths; the first three share
# a filter bank (Morlet, periods 4-40 samples, 3 voices, Nz 4096), the last one's
# window needs Nz 8192 (its own bank: it runs alone)
WPCC_KW = dict(pmin=4.0, pmax=40.0, V=3, wtype=-1)
WPCC_GROUP = [(5, 1, 3000, 400, "none"), (7, 2, 2500, 300, "plain"), (3, 1, 3200, 150, "none"),
              (4, 1, 3500, 200, "none")]


def test_station_stage_batch_wpcc(fp, hip):
    L = fp.api._lib.lib()
    n = len(WPCC_GROUP)
    jobs = (fp.api._lib.FpccJob * n)()
    dx, dws, wsb, refs = [], [], [], []
    for k, (S, W, N, lag, stack) in enumerate(WPCC_GROUP):
        x, _ = synth.make_stations(S, W * N + 7 * k, 100, seed=90 + k)
        job = fp.make_job("wpcc", S, W, N, -lag, lag, stack=stack, station_stride=x.shape[1], **WPCC_KW)
        jobs[k] = job
        d = hip.alloc(x.nbytes)
        hip.h2d(d, x)
        dx.append(d)
        wsb.append(L.fpcc_workspace_size(ctypes.byref(job)))
        dws.append(hip.alloc(wsb[-1]))
        refs.append(fp.allpairs(x, W, N, -lag, lag, kind="wpcc", stack=stack, **WPCC_KW))
    print("wpcc group: station stages of", [s[:3] for s in WPCC_GROUP], flush=True)
    r = L.fpcc_station_stage_batch_device(jobs, n, (ctypes.c_void_p * n)(*dx), (ctypes.c_void_p * n)(*dws),
                                          (ctypes.c_size_t * n)(*wsb), None)
    assert r == 0
    for k in range(n):
        y = np.empty(refs[k].shape, np.float32)
        dy = hip.alloc(y.nbytes)
        assert L.fpcc_pair_stage_device(ctypes.byref(jobs[k]), dws[k], wsb[k], dy, None) == 0
        hip.d2h(y, dy)
        print(f"  {WPCC_GROUP[k]}: max|batch - alone| = {np.abs(y - refs[k]).max():.3e}", flush=True)
        assert np.array_equal(y, refs[k])


def test_server_groups_wpcc(fp):
    """A burst of WPCC requests with one filter bank runs as a group; every result
    equals the standalone run."""
    items, refs = [], []
    for k, (S, W, N, lag, stack) in enumerate(WPCC_GROUP[:3]):
        x, _ = synth.make_stations(S, W * N, 100, seed=110 + k)
        job = fp.make_job("wpcc", S, W, N, -lag, lag, stack=stack, station_stride=W * N, **WPCC_KW)
        y = np.empty(int(np.prod(fp.output_shape(job))), np.float32)
        items.append((job, x, y))
        refs.append(fp.allpairs(x, W, N, -lag, lag, kind="wpcc", stack=stack, **WPCC_KW).ravel())
    with fp.Server(device_bytes=2 << 30, pinned_bytes=128 << 20, max_inflight=16) as srv:
        ids = srv.submit_many(items)
        for rid in ids:
            assert srv.wait(rid)["state"] == 2
        st = srv.stats()
        print(f"wpcc: batched {st['batched_requests']} in {st['batch_launches']} groups", flush=True)
        assert st["batched_requests"] >= 2 and st["batch_launches"] >= 1 and st["device_used"] == 0
    for (job, x, y), ref in zip(items, refs):
        assert np.array_equal(y, ref)
