#!/usr/bin/env python3
"""Throughput benchmark of the all-pairs PCC2 path on one MI355X.

Workload (default): 256 stations x 86400 float32 samples (one day at 1 Hz),
24 non-overlapping windows of 3600 samples, nfft 8192, all 32640 station pairs,
lags +-1000, linear stack (mean) over windows.  Inputs are seeded synthetic
records (noise + delayed band-limited common source, SNR 0.5) resident in HBM
before the timed region.  A step is one full pass: station stage (phasor spectra
of all 6144 station windows) + pair stage (all pair correlations, stacked).

Prints one JSON line: pair-correlations/s (one (pair, window) correlation of L
lags = one unit) with the stack on split-FP16 MFMA (FP32 accumulation), the
FP32 Gram and the plain FP32 path (per pair-window IFFT) timed beside it,
per-kernel roofline of the dominant kernel from HIP events recorded inside the
timed region, the oracle's CPU rate on a bounded sample, and a serving run: a
Poisson stream of uneven all-pairs jobs (4-128 stations, 1800-86400 samples,
lags +-200..2000, PCC2/GNCC/1-bit/WPCC 50/20/10/20) through the request server,
with p50/p99 latency from submission to the result in host memory.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CONFIGS = {
    # name: stations, windows, samples per window, max lag, max shift
    "c1": dict(S=8, W=1, N=3600, lag=500, shift=200),
    "c2": dict(S=64, W=24, N=3600, lag=1000, shift=1000),
    "c3": dict(S=256, W=24, N=3600, lag=1000, shift=1000),
}
HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"gram_f32": 157.3, "gram_split16": 2500.0}  # dense FP32 / FP16 MFMA


def nz_rule(N, M):
    n = 1
    while n < N + M:
        n <<= 1
    return n


def kernel_flops(name, cfg, Nz):
    """Algorithmic flops of one Gram launch: P pairs x (Nz/2+1) bins x 2W complex
    multiply-adds (8 flops each) of the Hermitian-folded stack."""
    if not name.startswith("gram"):
        return None
    S, W = cfg["S"], cfg["W"]
    P = S * (S - 1) // 2
    return P * (Nz // 2 + 1) * 2 * W * 8


# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by
# the same counters' readings for a 1 GiB copy), read from the committed summary
# of the counter runs for the headline workload; the summary names the commit it
# was measured at, which the JSON line repeats.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pcc2_allpairs_256st_gram16_hbm_traffic.md")
TRAFFIC_KERNELS = {"k_station_pcc_3600": "station_pcc", "k_spectra_to_u": "spectra_to_u",
                   "k_gram16_dma": "gram_split16", "k_pair_c2r_4096": "pair_c2r", "k_station_u_3600": "station_u"}


def measured_traffic():
    """{bench kernel name: traffic bytes per launch}, and the commit line of the summary."""
    out, commit = {}, None
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return out, None
    for ln in lines:
        if ln.lower().startswith("measured at commit"):
            commit = ln.split()[-1].strip("`.")
        cells = [c.strip() for c in ln.split("|")]
        if len(cells) < 8 or not cells[1].startswith("fpcc::"):
            continue
        sym = cells[1][len("fpcc::"):].split("<")[0]
        if sym in TRAFFIC_KERNELS:
            try:
                out[TRAFFIC_KERNELS[sym]] = float(cells[6]) * 1e6
            except ValueError:
                pass
    return out, commit


def kernel_bytes(name, cfg, Nz):
    """Algorithmic HBM bytes of one launch of `name`."""
    S, W, N, lag = cfg["S"], cfg["W"], cfg["N"], cfg["lag"]
    P, L = S * (S - 1) // 2, 2 * lag + 1
    spectra = S * W * Nz * 8
    if name.startswith("station"):
        return S * W * N * 4 + spectra            # read records, write spectra
    if name.startswith("pair_pcc2"):
        return spectra + P * L * 4                # read spectra once, write stacked lags
    nk = Nz // 2 + 1
    if name.startswith("gram"):
        # read the operands (2W rows a bin: the W spectra and the W conjugated mirror
        # bins, 8 bytes each: FP32 re/im, or FP16 hi+lo re/im), write the folded
        # cross spectra
        return S * 2 * W * nk * 8 + P * nk * 8
    if name.startswith("spectra_to_u"):
        return spectra + S * 2 * W * nk * 8       # read spectra, write the Gram operands
    if name.startswith("pair_c2r"):
        return P * (Nz // 2 + 1) * 8 + P * L * 4  # read folded cross spectra, write lags
    return None


def cpu_baseline(x, cfg, seconds):
    import oracle
    S, W, N, lag = cfg["S"], cfg["W"], cfg["N"], cfg["lag"]
    P = S * (S - 1) // 2
    done, t0 = 0, time.perf_counter()
    chunk = 4
    while done < P:
        n = min(chunk, P - done)
        oracle.allpairs("pcc2", x, W, N, -lag, lag, stacked=True, pair_begin=done, pair_end=done + n)
        done += n
        el = time.perf_counter() - t0
        if el >= seconds:
            break
        chunk = max(4, min(256, int(done / el * 0.25)))
    el = time.perf_counter() - t0
    return {"value": done * W / el, "unit": "pair-correlations/s", "cores": oracle.max_threads(), "kind": "port",
            "sample": f"first {done} of {P} pairs x {W} windows of the same workload ({el:.1f} s), "
                      f"oracle pcc2_set restatement, OpenMP over windows of one pair per call"}


def cpu_rate(kind, x, W, N, lag, budget_s, v=2.0):
    """The oracle's reference-precision restatement of the reference's *_set call for
    `kind` on this host's cores: one station pair per call with OpenMP over its
    windows (as the reference's one-pair-per-run driver), pairs in order until all
    are done or budget_s has passed."""
    import oracle
    S = x.shape[0]
    P = S * (S - 1) // 2
    done, t0, chunk = 0, time.perf_counter(), 1
    while done < P:
        n = min(chunk, P - done)
        oracle.allpairs(kind, x, W, N, -lag, lag, stacked=True, pair_begin=done, pair_end=done + n, v=v)
        done += n
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
        chunk = max(1, min(256, int(done / el * 0.25)))
    el = time.perf_counter() - t0
    return {"value": done * W / el, "unit": "pair-correlations/s", "cores": oracle.max_threads(), "kind": "port",
            "sample": (f"all {P} pairs" if done == P else f"first {done} of {P} pairs") +
                      f" x {W} windows ({el:.1f} s)"}


def _check_served(reqs, ys, idx, net):
    """Served outputs of requests idx against the same job run alone (fpcc_allpairs,
    bit for bit) and against the oracle on the first and last pair (reference
    precision; WPCC against its FP64 restatement, as the reference computes it)."""
    import numpy as np
    import oracle
    from fastpcc_amd import serving, api
    out = {"requests": len(idx), "alone_bit_exact": True, "max_abs_err_vs_oracle": 0.0, "checked": []}
    for k in idx:
        rq = reqs[k]
        job = serving.job_for(rq, net.shape[1])
        x = np.ascontiguousarray(net[rq["s0"]:rq["s0"] + rq["S"], rq["t0"]:rq["t0"] + rq["N"]])
        kw = serving.WPCC_PARAMS if rq["kind"] == "wpcc" else {}
        alone = api.allpairs(x, 1, rq["N"], -rq["lag"], rq["lag"], kind=rq["kind"], stack="none", **kw).ravel()
        y = np.asarray(ys[k])
        same = bool(np.array_equal(alone, y))
        out["alone_bit_exact"] &= same
        P = rq["S"] * (rq["S"] - 1) // 2
        L = 2 * rq["lag"] + 1
        err = 0.0
        for pb in (0, P - 1):
            yr = oracle.allpairs(rq["kind"], x, 1, rq["N"], -rq["lag"], rq["lag"], stacked=True,
                                 f64=rq["kind"] == "wpcc", pair_begin=pb, pair_end=pb + 1, **kw)
            err = max(err, float(np.abs(y.reshape(P, L)[pb] - yr[0]).max()))
        out["max_abs_err_vs_oracle"] = max(out["max_abs_err_vs_oracle"], err)
        out["checked"].append({"kind": rq["kind"], "stations": rq["S"], "N": rq["N"], "lags": rq["lag"],
                               "alone_bit_exact": same, "max_abs_err_vs_oracle": err})
        del job
    return out


def serve_run(n_requests, rate, sweep_rates, sweep_n, x3=None, y3_ref=None, check=True):
    """The serving workload through the request server: after a warm-up (plans and
    filter banks of every zero-pad length are built on first use), the Poisson
    stream at `rate` with a sample of its outputs checked; an offered-load sweep
    (the same mix at increasing arrival rates, up to and past saturation); the
    256-station headline job submitted through the server; and the network-day
    stream (station subsets of one shared window) with the resident spectra cache
    on and off."""
    import numpy as np
    from fastpcc_amd import serving, Server, make_job, pinned_empty
    net = serving.make_network()
    kw = dict(device_bytes=96 << 30, pinned_bytes=4 << 30, max_inflight=16, n_streams=4)
    warm = serving.make_requests(48, rate, seed=99)
    reqs = serving.make_requests(n_requests, rate, seed=7)
    # two requests of each kind (WPCC: windows up to 30000 samples, for the FP64 oracle's time)
    idx = []
    for kind in serving.KINDS:
        cand = [k for k, r in enumerate(reqs) if r["kind"] == kind and (kind != "wpcc" or r["N"] <= 30000)]
        idx += cand[:2]
    pool = serving.OutputPool(24, max(serving.out_len(r) for r in reqs + warm))
    with Server(**kw) as srv:
        serving.run(warm, net, server=srv, realtime=False, pool=pool, keep=False)
    with Server(**kw) as srv:
        stats, _, ys = serving.run(reqs, net, server=srv, realtime=True, pool=pool, keep=False,
                                   keep_idx=idx if check else ())
    out = serving.summary(stats, reqs, rate)
    out["workload"] = ("Poisson stream, stations U{4..128}, one window of N log-uniform in [1800, 86400], "
                       "lags +-U{200..2000} (< N), PCC2/GNCC/1-bit/WPCC 50/20/10/20 (WPCC: Morlet, periods "
                       "2-200 samples, 8 voices), inputs from a pinned 128 x 86400 network, outputs into a "
                       "pool of 24 pinned buffers (pageable when all are busy); no station keys (every "
                       "request starts at its own time, so no station window repeats)")
    if check:
        out["served_output_check"] = _check_served(reqs, ys, idx, net)
    del ys
    sweep = []
    for i, r in enumerate(sweep_rates):
        rq = serving.make_requests(sweep_n, r, seed=300 + i)
        with Server(**kw) as srv:
            st, _, _ = serving.run(rq, net, server=srv, realtime=True, pool=pool, keep=False)
        e = serving.summary(st, rq, r)
        sweep.append({k: e[k] for k in ("arrival_rate_per_s", "requests", "completed", "completed_per_s",
                                        "pair_correlations_per_s", "p50_ms", "p99_ms", "max_ms",
                                        "inflight_peak", "wall_s")})
    out["load_sweep"] = {"requests_per_rate": sweep_n, "rates": sweep,
                         "note": "the same mix at increasing Poisson arrival rates; past saturation the "
                                 "completion rate levels off and latency grows with the queue"}
    del pool
    if x3 is not None:
        # the headline job (256 stations x 24 windows of 3600, split-FP16 stack) through
        # the server from pinned host memory: four requests in a row, overlapping the
        # input and output copies of one with the kernels of another
        S3 = x3.shape[0]
        P3 = S3 * (S3 - 1) // 2
        job = make_job("pcc2", S3, 24, 3600, -1000, 1000, stack="gram16")
        xp = pinned_empty(x3.shape)
        xp[:] = x3
        ys3 = [pinned_empty(P3 * 2001) for _ in range(4)]
        with Server(**kw) as srv:
            srv.wait(srv.submit(job, xp, ys3[0]))  # plans
            t0 = time.perf_counter()
            ids = [srv.submit(job, xp, y) for y in ys3]
            lat = []
            for rid in ids:
                st = srv.wait(rid)
                lat.append((st["t_done"] - st["t_submit"]) * 1e3)
            wall = time.perf_counter() - t0
        c3 = {"requests": 4, "wall_ms": round(wall * 1e3, 2), "latency_ms": [round(v, 2) for v in lat],
              "pair_correlations_per_s": 4 * P3 * 24 / wall,
              "note": "PCIe-inclusive: 88.5 MB in and 261 MB out per request, pinned host buffers"}
        if y3_ref is not None:
            c3["equal_to_device_path"] = bool(all(np.array_equal(np.asarray(y), y3_ref) for y in ys3))
        out["headline_job_served"] = c3
        del ys3, xp
    day_rate = 200.0
    day = serving.make_network_day_requests(max(100, n_requests // 12), day_rate)
    out["network_day"] = {"workload": "Poisson 200/s, stations: random subsets of 8-64 of the 128, the same "
                                      "21600-sample window (lags +-1000), PCC2/WPCC 50/50"}
    for name, cb in (("cache", 16 << 30), ("no_cache", 0)):
        with Server(cache_bytes=cb, **kw) as srv:
            serving.run(day[:16], net, server=srv, realtime=False, keys=cb > 0)  # plans, filter banks
        with Server(cache_bytes=cb, **kw) as srv:
            st, _, _ = serving.run(day, net, server=srv, realtime=True, keys=cb > 0)
        out["network_day"][name] = serving.summary(st, day, day_rate)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--stack", default="gram16", choices=["plain", "zsum", "gram", "gram16"])
    ap.add_argument("--variants", default="plain,gram",
                    help="other stack modes timed briefly for comparison (comma list, '' for none)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=3000,
                    help="requests of the serving run (0: skip); 3000 is the full stream")
    ap.add_argument("--serve-rate", type=float, default=50.0, help="mean arrivals per second")
    ap.add_argument("--sweep-rates", default="100,200,400,800,1600,3200",
                    help="arrival rates of the offered-load sweep ('' for none)")
    ap.add_argument("--sweep-requests", type=int, default=400, help="requests per rate of the sweep")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph replay of the headline step")
    ap.add_argument("--no-host", action="store_true", help="skip the PCIe-inclusive host-buffer timing")
    ap.add_argument("--no-pre", action="store_true", help="skip the preprocessing timing")
    ap.add_argument("--no-wpcc", action="store_true", help="skip the WPCC all-pairs timing (configuration 4)")
    ap.add_argument("--no-kinds", action="store_true",
                    help="skip the other correlation kinds at configuration 2 (64 stations)")
    ap.add_argument("--no-paper", action="store_true",
                    help="skip the reference's published workload (one pair, 649 day windows of 21600)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="before the warm-up, one fill and one copy of this many MiB (the known-size kernels "
                         "that scale FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the path shards over station-pair tiles; "
              "multi-GPU runs are not part of it); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import fastpcc_amd
    from fastpcc_amd import synth, _lib

    cfg = CONFIGS[args.config]
    S, W, N, lag = cfg["S"], cfg["W"], cfg["N"], cfg["lag"]
    P, L = S * (S - 1) // 2, 2 * lag + 1
    Nz = nz_rule(N, lag)
    x, _ = synth.make_stations(S, W * N, cfg["shift"], seed=1234)
    dev = torch.device("cuda:0")
    x_dev = torch.from_numpy(x).to(dev)
    job = fastpcc_amd.make_job("pcc2", S, W, N, -lag, lag, stack=args.stack)
    runner = fastpcc_amd.DeviceRunner(job, dev)
    stream = torch.cuda.current_stream(dev)
    lib = _lib.lib()
    lib.fpcc_profile_report.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]

    if args.calib_copy_mb > 0:
        ca = torch.empty(args.calib_copy_mb << 20, dtype=torch.uint8, device=dev)
        cb = torch.empty_like(ca)
        ca.fill_(1)
        cb.copy_(ca)
        torch.cuda.synchronize()
        del ca, cb
    for _ in range(args.warmup):
        runner.run(x_dev, stream)
    torch.cuda.synchronize()

    check = None
    if not args.no_check:
        import oracle
        y = runner.y.view(P, L).cpu().numpy()
        npair = 2
        yr = oracle.allpairs("pcc2", x, W, N, -lag, lag, stacked=True, pair_end=npair)
        y64 = oracle.allpairs("pcc2", x, W, N, -lag, lag, stacked=True, pair_end=npair, f64=True)
        check = {"pairs": npair, "max_abs_err_vs_oracle": float(np.abs(y[:npair] - yr).max()),
                 "max_abs_err_vs_fp64": float(np.abs(y[:npair] - y64).max())}

    # the host-side check above leaves the GPU idle for seconds: warm it up again so
    # the timed steps do not start at idle clocks
    for _ in range(args.warmup):
        runner.run(x_dev, stream)
    # the headline: K steps with no per-kernel timing inside the timed loop
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        runner.run(x_dev, stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # the same K steps again with a HIP event pair around every launch (on the
    # stream it runs on), for the per-kernel durations and the roofline
    lib.fpcc_profile_reset()
    lib.fpcc_profile_enable(1)
    torch.cuda.synchronize()
    tp0 = time.perf_counter()
    for _ in range(args.steps):
        runner.run(x_dev, stream)
    torch.cuda.synchronize()
    tp1 = time.perf_counter()
    lib.fpcc_profile_enable(0)
    names = ctypes.create_string_buffer(4096)
    tot = (ctypes.c_double * 16)()
    cnt = (ctypes.c_longlong * 16)()
    nk = lib.fpcc_profile_report(names, 4096, tot, cnt, 16)
    kernels = {}
    for k, nm in enumerate(names.value.decode().split("\n")[:nk]):
        kernels[nm] = {"avg_ms": tot[k] / max(cnt[k], 1), "launches": int(cnt[k])}

    ms_per_step = (t1 - t0) * 1e3 / args.steps
    ms_per_step_profiled = (tp1 - tp0) * 1e3 / args.steps
    value = P * W / (ms_per_step / 1e3)
    traffic, traffic_commit = measured_traffic() if (args.config, args.stack) == ("c3", "gram16") else ({}, None)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"]) if kernels else None
    roof = None
    for k, v in kernels.items():
        b = kernel_bytes(k, cfg, Nz)
        if b:
            v["algorithmic_GBps"] = round(b / (v["avg_ms"] * 1e-3) / 1e9, 1)
        f = kernel_flops(k, cfg, Nz)
        if f:
            v["algorithmic_TFLOPs"] = round(f / (v["avg_ms"] * 1e-3) / 1e12, 2)
    for k, v in kernels.items():
        if k in traffic:
            v["traffic_MB"] = round(traffic[k] / 1e6, 1)
    if dom:
        f = kernel_flops(dom, cfg, Nz)
        b = kernel_bytes(dom, cfg, Nz)
        # the binding roof: MFMA time at peak (the split path issues 3 FP16 products
        # per real product) against HBM time of the algorithmic bytes at peak
        mfma_bound = False
        if f and dom in MFMA_PEAK_TFLOPS:
            issued = f * (3 if dom == "gram_split16" else 1)
            mfma_bound = issued / (MFMA_PEAK_TFLOPS[dom] * 1e12) >= b / (HBM_PEAK_GBS * 1e9)
        if mfma_bound:
            ach = f / (kernels[dom]["avg_ms"] * 1e-3) / 1e12
            peak = MFMA_PEAK_TFLOPS[dom]
            roof = {"kernel": dom, "bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                    "frac": round(ach / peak, 4), "traffic": traffic.get(dom),
                    "algorithmic_flops": f}
        else:
            b = kernel_bytes(dom, cfg, Nz)
            ach = b / (kernels[dom]["avg_ms"] * 1e-3) / 1e9
            roof = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic.get(dom),
                    "algorithmic_bytes": b}
        if roof["traffic"] is not None:
            roof["traffic_source"] = (os.path.relpath(TRAFFIC_FILE, ROOT) +
                                      (f" (measured at commit {traffic_commit})" if traffic_commit else ""))

    variants = {}
    if not args.no_graph:
        # the same step replayed as one HIP graph (captured on its own stream)
        gstream = torch.cuda.Stream(dev)
        runner.capture(x_dev, gstream)
        for _ in range(2):
            runner.replay(gstream)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(args.steps):
            runner.replay(gstream)
        torch.cuda.synchronize()
        gms = (time.perf_counter() - ta) * 1e3 / args.steps
        variants["graph_" + args.stack] = {"ms_per_step": gms, "value": P * W / (gms / 1e3), "steps": args.steps,
                                           "note": "the headline step replayed as one captured HIP graph"}
        runner.release_graph()
    for vs in [v for v in args.variants.split(",") if v and v != args.stack]:
        vjob = fastpcc_amd.make_job("pcc2", S, W, N, -lag, lag, stack=vs)
        vr = fastpcc_amd.DeviceRunner(vjob, dev)
        vr.run(x_dev, stream)
        torch.cuda.synchronize()
        nst = 3 if vs == "plain" else 10
        ta = time.perf_counter()
        for _ in range(nst):
            vr.run(x_dev, stream)
        torch.cuda.synchronize()
        vms = (time.perf_counter() - ta) * 1e3 / nst
        ent = {"ms_per_step": vms, "value": P * W / (vms / 1e3), "steps": nst}
        if not args.no_check:
            import oracle
            yv = vr.y.view(P, L).cpu().numpy()[:2]
            y64 = oracle.allpairs("pcc2", x, W, N, -lag, lag, stacked=True, pair_end=2, f64=True)
            ent["max_abs_err_vs_fp64"] = float(np.abs(yv - y64).max())
        variants[vs] = ent
        del vr
    if not args.no_pre:
        # per-station preprocessing of the same network: clipping, outlier windows
        # (rms > 3 x median), FP64 whitening over 0.02-0.2 Hz
        pp = fastpcc_amd.make_preproc(clip=True, std=3.0, white=(0.02, 0.2), dt=1.0)
        pjob = fastpcc_amd.make_job("pcc2", S, W, N, 0, 0, station_stride=W * N)
        pws = lib.fpcc_preprocess_workspace_size(ctypes.byref(pjob), ctypes.byref(pp))
        xp = torch.empty_like(x_dev)
        pvalid = torch.empty(S * W, dtype=torch.uint8, device=dev)
        prms = torch.empty(S * W, dtype=torch.float32, device=dev)
        pwsb = torch.empty(pws, dtype=torch.uint8, device=dev)

        def pre_once():
            xp.copy_(x_dev)
            r = lib.fpcc_preprocess_device(ctypes.byref(pjob), ctypes.byref(pp), ctypes.c_void_p(xp.data_ptr()), None,
                                           ctypes.c_void_p(pvalid.data_ptr()), ctypes.c_void_p(prms.data_ptr()),
                                           ctypes.c_void_p(pwsb.data_ptr()), pws, ctypes.c_void_p(stream.cuda_stream))
            if r:
                raise RuntimeError(f"fpcc_preprocess_device returned {r}")

        pre_once()
        torch.cuda.synchronize()
        lib.fpcc_profile_reset()
        lib.fpcc_profile_enable(1)
        for _ in range(5):
            pre_once()
        torch.cuda.synchronize()
        lib.fpcc_profile_enable(0)
        nk = lib.fpcc_profile_report(names, 4096, tot, cnt, 16)
        pk = {nm: round(tot[k] / max(cnt[k], 1), 4) for k, nm in enumerate(names.value.decode().split("\n")[:nk])}
        pms = sum(pk.values())
        variants["preprocess"] = {
            "ms_per_step": pms, "kernels_avg_ms": pk, "steps": 5, "kept_windows": int(pvalid.sum().item()),
            "x_GBps": round(5 * x.nbytes / (pms * 1e-3) / 1e9, 1),
            "note": "fpcc_preprocess_device on the same network: clipping, outlier windows (rms > 3 x median), "
                    "FP64 whitening 0.02-0.2 Hz; ms = sum of kernel averages; x_GBps counts 5 passes over x"}
        del xp, pwsb
    if not args.no_wpcc:
        # configuration 4: WPCC of all 496 pairs of 32 stations x 86400 samples (one window,
        # nfft 131072), Morlet filter bank, periods 2-200 samples, 8 voices (56 scales), lags +-2000
        Sw, Nw, lw = 32, 86400, 2000
        kw = dict(pmin=2.0, pmax=200.0, V=8, wtype=-1)
        xw, _ = synth.make_stations(Sw, Nw, 1000, seed=4321)
        xwd = torch.from_numpy(xw).to(dev)
        wjob = fastpcc_amd.make_job("wpcc", Sw, 1, Nw, -lw, lw, stack="plain", **kw)
        wr = fastpcc_amd.DeviceRunner(wjob, dev)
        wr.run(xwd, stream)
        torch.cuda.synchronize()
        lib.fpcc_profile_reset()
        lib.fpcc_profile_enable(1)
        nst = 5
        ta = time.perf_counter()
        for _ in range(nst):
            wr.run(xwd, stream)
        torch.cuda.synchronize()
        wms = (time.perf_counter() - ta) * 1e3 / nst
        lib.fpcc_profile_enable(0)
        nk = lib.fpcc_profile_report(names, 4096, tot, cnt, 16)
        wk = {nm: round(tot[k] / max(cnt[k], 1), 4) for k, nm in enumerate(names.value.decode().split("\n")[:nk])}
        Pw = Sw * (Sw - 1) // 2
        ent = {"ms_per_step": wms, "value": Pw / (wms / 1e3), "unit": "pair-correlations/s", "steps": nst,
               "kernels_avg_ms": wk, "workspace_GB": round(wr.ws_bytes / 1e9, 2),
               "workload": "WPCC all 496 pairs, 32 stations x 86400 samples, nfft 131072, Morlet, periods "
                           "2-200 samples, 8 voices (56 scales), lags +-2000, per-window output"}
        if not args.no_cpu or not args.no_check:
            import oracle
            yw = wr.y.view(Pw, 2 * lw + 1).cpu().numpy()[0]
            ta = time.perf_counter()
            yr = oracle.run("wpcc", xw[:1], xw[1:2], -lw, lw, **kw)[0]
            cpu_s = time.perf_counter() - ta
            ent["max_abs_err_vs_oracle"] = float(np.abs(yw - yr).max())
            ent["cpu_baseline"] = {"value": 1.0 / cpu_s, "unit": "pair-correlations/s", "cores": 1, "kind": "port",
                                   "sample": "pair (0, 1): the oracle's tspcc2_set restatement (FP64, as the "
                                             "reference), one core"}
        variants["wpcc_c4"] = ent
        del wr, xwd
    if not args.no_kinds:
        # configuration 2: 64 stations x 24 windows of 3600, lags +-1000, every kind, mean stack
        # (PCC-v with v = 1.5: the reference routes v = 1 and 2 to PCC1 and PCC2)
        S2, W2c, N2c, l2 = 64, 24, 3600, 1000
        x2, _ = synth.make_stations(S2, W2c * N2c, 1000, seed=2024)
        x2d = torch.from_numpy(x2).to(dev)
        P2 = S2 * (S2 - 1) // 2
        kinds = {}
        for name, kind, kst, kv, nst in (("pcc2", "pcc2", "gram16", 2.0, 10), ("pcc2_plain", "pcc2", "plain", 2.0, 5),
                                         ("pcc2_pws", "pcc2", "pws", 2.0, 5), ("gncc", "gncc", "plain", 2.0, 10),
                                         ("cc1b", "cc1b", "gram16", 2.0, 10), ("cc1b_pws", "cc1b", "pws", 2.0, 5),
                                         ("pcc1", "pcc1", "plain", 2.0, 2), ("pccv", "pccv", "plain", 1.5, 1)):
            kj = fastpcc_amd.make_job(kind, S2, W2c, N2c, -l2, l2, stack=kst, v=kv)
            kr = fastpcc_amd.DeviceRunner(kj, dev)
            kr.run(x2d, stream)
            torch.cuda.synchronize()
            ta = time.perf_counter()
            for _ in range(nst):
                kr.run(x2d, stream)
            torch.cuda.synchronize()
            kms = (time.perf_counter() - ta) * 1e3 / nst
            ent = {"stack": kst, "ms_per_step": round(kms, 4), "value": P2 * W2c / (kms / 1e3), "steps": nst}
            if kind == "pccv":
                ent["v"] = kv
            if not args.no_check:
                import oracle
                yk = kr.y.view(P2, 2 * l2 + 1).cpu().numpy()[:2]
                if kst == "pws":
                    y64 = oracle.pws(kind, x2, W2c, N2c, -l2, l2, 2.0, pair_end=2)
                else:
                    y64 = oracle.allpairs(kind, x2, W2c, N2c, -l2, l2, stacked=True, pair_end=2, f64=True, v=kv)
                ent["max_abs_err_vs_fp64"] = float(np.abs(yk - y64).max())
            if not args.no_cpu and name in ("pcc2", "gncc", "cc1b", "pcc1", "pccv"):
                # the reference's CPU path for this kind on the same network (all pairs
                # within the budget, else the first pairs)
                ent["cpu_baseline"] = cpu_rate(kind, x2, W2c, N2c, l2, 3.0 if kind in ("pcc1", "pccv") else 12.0, kv)
            kinds[name] = ent
            del kr
        variants["kinds_c2"] = {"unit": "pair-correlations/s", "kinds": kinds,
                                "workload": "64 stations x 86400 samples, 24 windows of 3600, nfft 8192, lags "
                                            "+-1000, all 2016 pairs, mean stack (pws: phase-weighted stack, "
                                            "power 2)"}
        del x2d
        # configuration 1: 8 stations x 3600 samples, all 28 pairs, lags +-500, PCC2
        x1c, _ = synth.make_stations(8, 3600, 200, seed=1)
        x1d = torch.from_numpy(x1c).to(dev)
        j1 = fastpcc_amd.make_job("pcc2", 8, 1, 3600, -500, 500, stack="plain")
        r1 = fastpcc_amd.DeviceRunner(j1, dev)
        r1.run(x1d, stream)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        for _ in range(50):
            r1.run(x1d, stream)
        torch.cuda.synchronize()
        c1ms = (time.perf_counter() - ta) * 1e3 / 50
        c1 = {"ms_per_step": round(c1ms, 4), "value": 28 / (c1ms / 1e3), "unit": "pair-correlations/s",
              "steps": 50, "workload": "PCC2, 8 stations x 3600 samples, one window, all 28 pairs, lags +-500"}
        if not args.no_cpu:
            c1["cpu_baseline"] = cpu_rate("pcc2", x1c, 1, 3600, 500, 10.0)
        variants["c1"] = c1
        del r1, x1d
    if not args.no_paper:
        # the reference's published benchmark (SRL 2019, doc/Ventosa_S_2019_j-srl_tow_pldvPCC.pdf p.4): 649
        # day-long windows of one station pair, N = 21600 (dt = 4 s), lags +-12000 s = +-3000 samples, per-window
        # output; PCC1 took 71.65 s on a 4-core i5-4690K and 0.70 s GPU-enhanced (GTX 1060), PCC2 and 1-bit
        # GNCC on the CPU 1.79x and 4.17x faster than that GPU run.  Here through the reference's own entry
        # points (host buffers in and out; the library keeps their staging between calls), and
        # device-resident (inputs already in HBM, the pair's 649 windows as one job).
        Tw, Nw, lp = 649, 21600, 3000
        xa, _ = synth.make_stations(2, Tw * Nw, 1000, seed=649)
        x1p, x2p = xa[0].reshape(Tw, Nw), xa[1].reshape(Tw, Nw)
        published = {"pcc1": {"gpu_gtx1060_s": 0.70, "cpu_i5_4690k_s": 71.65},
                     "pcc2": {"cpu_i5_4690k_s": round(0.70 / 1.79, 3)}, "cc1b": {"cpu_i5_4690k_s": round(0.70 / 4.17, 3)}}
        paper = {}
        for kind, fn in (("pcc1", fastpcc_amd.pcc1_set), ("pcc2", fastpcc_amd.pcc2_set), ("cc1b", fastpcc_amd.cc1b_set)):
            fn(x1p[:8], x2p[:8], -lp, lp)  # plans and code objects
            torch.cuda.synchronize()
            ta = time.perf_counter()
            fn(x1p, x2p, -lp, lp)  # first call at this size: grows the entry points' kept buffers
            first_s = time.perf_counter() - ta
            ta = time.perf_counter()
            yp = fn(x1p, x2p, -lp, lp)
            host_s = time.perf_counter() - ta
            pj = fastpcc_amd.make_job(kind, 2, Tw, Nw, -lp, lp, stack="none")
            pr = fastpcc_amd.DeviceRunner(pj, dev)
            xpd = torch.from_numpy(xa).to(dev)
            pr.run(xpd, stream)
            torch.cuda.synchronize()
            ta = time.perf_counter()
            for _ in range(3):
                pr.run(xpd, stream)
            torch.cuda.synchronize()
            dev_s = (time.perf_counter() - ta) / 3
            ent = {"entry_point_s": round(host_s, 4), "first_call_s": round(first_s, 4),
                   "device_resident_s": round(dev_s, 5), "published": published[kind]}
            for k, ref_s in published[kind].items():
                ent["speedup_vs_" + k.rsplit("_", 1)[0]] = round(ref_s / host_s, 1)
            if not args.no_check:
                import oracle
                yr = oracle.run(kind, x1p[:2], x2p[:2], -lp, lp)
                ent["max_abs_err_vs_oracle"] = float(np.abs(np.asarray(yp)[:2] - yr).max())
                ent["max_abs_diff_vs_device_resident"] = float(
                    np.abs(np.asarray(yp) - pr.y.view(Tw, 2 * lp + 1).cpu().numpy()).max())
            paper[kind] = ent
            del pr, xpd
        variants["paper_649_windows"] = {
            "unit": "s", "paper": paper,
            "workload": "one station pair, 649 windows of 21600 samples (synthetic, the published run's shape), "
                        "lags +-3000, per-window output; entry_point_s = pcc1_set / pcc2_set / cc1b_set from host "
                        "buffers to host output, a repeated call; first_call_s = the first call at this size, which "
                        "also allocates the pinned and device buffers the entry points keep"}
    if not args.no_host:
        # PCIe-inclusive: host input in, host output out, through fpcc_allpairs
        # (pageable host arrays)
        fastpcc_amd.allpairs(x, W, N, -lag, lag, kind="pcc2", stack=args.stack)
        ta = time.perf_counter()
        for _ in range(3):
            fastpcc_amd.allpairs(x, W, N, -lag, lag, kind="pcc2", stack=args.stack)
        hms = (time.perf_counter() - ta) * 1e3 / 3
        variants["host_buffers_" + args.stack] = {
            "ms_per_step": hms, "value": P * W / (hms / 1e3), "steps": 3,
            "note": "PCIe-inclusive: fpcc_allpairs from pageable host input to host output "
                    f"({x.nbytes / 1e6:.1f} MB in, {P * L * 4 / 1e6:.1f} MB out)"}
    speedup = None
    if "plain" in variants:
        speedup = round(value / variants["plain"]["value"], 2)
        for name, ent in variants.items():
            if name != "plain" and "value" in ent and "workload" not in ent:  # same workload only
                ent["speedup_vs_plain"] = round(ent["value"] / variants["plain"]["value"], 2)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(x, cfg, args.cpu_seconds)

    serving = None
    if args.serve_requests > 0:
        rates = [float(r) for r in args.sweep_rates.split(",") if r]
        y_ref = runner.y.view(P, L).cpu().numpy().ravel() if args.config == "c3" and args.stack == "gram16" else None
        serving = serve_run(args.serve_requests, args.serve_rate, rates, args.sweep_requests,
                            x3=x if y_ref is not None else None, y3_ref=y_ref, check=not args.no_check)

    out = {
        "metric": "PCC2 pair-correlations/s, 1 GPU: FP32 vs split-FP16 MFMA stack; p99 latency",
        "value": value,
        "unit": "pair-correlations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "ms_per_step_profiled": ms_per_step_profiled,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.stack != "gram16" else "f32 (stack: fp16 hi+lo split operands, f32 accumulation)",
        "data": "synthetic (seeded noise + delayed band-limited common source, SNR 0.5)",
        "config": {"workload": f"PCC2 all pairs, {S} stations x {W * N} samples, {W} windows of {N}, "
                               f"nfft {Nz}, lags +-{lag}, mean stack over windows",
                   "stations": S, "pairs": P, "windows": W, "win_len": N, "nfft": Nz, "lags": L,
                   "stack": args.stack, "parallelism": "1 GPU"},
        "kernels": kernels,
        "variants": variants,
        "speedup_vs_plain": speedup,
        "roofline": roof,
        "cpu_baseline": cpu,
        "check": check,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
