"""Loads libfastpcc.so (the HIP library) and declares its C ABI for ctypes.

There is no fallback: if the library is missing or fails to load, every entry
point raises.  Build it with ``python -m fastpcc_amd.build`` (or
``__graft_entry__.build()``).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastpcc.so")

FPCC_PCC2, FPCC_PCC1, FPCC_PCCV, FPCC_GNCC, FPCC_CC1B, FPCC_WPCC = range(6)
STACK_NONE, STACK_PLAIN, STACK_ZSUM, STACK_GRAM, STACK_GRAM16, STACK_PWS = range(6)


class FpccJob(ctypes.Structure):
    """Mirror of fpcc_job in include/fastpcc.h."""
    _fields_ = [
        ("kind", ctypes.c_int),
        ("stack", ctypes.c_int),
        ("n_stations", ctypes.c_uint),
        ("n_windows", ctypes.c_uint),
        ("win_len", ctypes.c_uint),
        ("station_stride", ctypes.c_ulonglong),
        ("lag1", ctypes.c_int),
        ("lag2", ctypes.c_int),
        ("v", ctypes.c_double),
        ("pmin", ctypes.c_double),
        ("pmax", ctypes.c_double),
        ("V", ctypes.c_uint),
        ("wtype", ctypes.c_int),
        ("op1", ctypes.c_double),
        ("pws_power", ctypes.c_double),
    ]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = range(5)


class Preproc(ctypes.Structure):
    """Mirror of fpcc_preproc."""
    _fields_ = [
        ("clip", ctypes.c_int),
        ("std_max", ctypes.c_float),
        ("band", ctypes.c_double * 2),
        ("dt", ctypes.c_double),
        ("common_windows", ctypes.c_int),
    ]


class ServerConfig(ctypes.Structure):
    """Mirror of fpcc_server_config."""
    _fields_ = [
        ("device_bytes", ctypes.c_size_t),
        ("pinned_bytes", ctypes.c_size_t),
        ("max_inflight", ctypes.c_int),
        ("n_streams", ctypes.c_int),
        ("max_bypass", ctypes.c_int),
        ("fail_every", ctypes.c_int),
        ("cache_bytes", ctypes.c_size_t),
    ]


class Request(ctypes.Structure):
    """Mirror of fpcc_request."""
    _fields_ = [
        ("job", FpccJob),
        ("x", ctypes.POINTER(ctypes.c_float)),
        ("y", ctypes.POINTER(ctypes.c_float)),
        ("station_keys", ctypes.POINTER(ctypes.c_ulonglong)),
    ]


class RequestStatus(ctypes.Structure):
    """Mirror of fpcc_request_status."""
    _fields_ = [
        ("state", ctypes.c_int),
        ("rc", ctypes.c_int),
        ("t_submit", ctypes.c_double),
        ("t_start", ctypes.c_double),
        ("t_done", ctypes.c_double),
    ]


class ServingStats(ctypes.Structure):
    """Mirror of fpcc_serving_stats."""
    _fields_ = [(n, ctypes.c_ulonglong) for n in ("submitted", "completed", "failed", "cancelled", "rejected",
                                                  "cache_hits", "cache_misses")] + \
               [("cache_used", ctypes.c_size_t)] + \
               [(n, ctypes.c_ulonglong) for n in ("batched_requests", "batch_launches")] + \
               [(n, ctypes.c_ulonglong) for n in ("queued", "inflight", "inflight_peak")] + \
               [(n, ctypes.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "pair_correlations",
                                               "busy_s")] + \
               [(n, ctypes.c_size_t) for n in ("device_bytes", "device_used", "device_peak")]


_lib = None


def lib():
    """The loaded library; raises RuntimeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"fastpcc: HIP library not built ({LIB_PATH}); run python -m fastpcc_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    fpp = ctypes.POINTER(ctypes.POINTER(ctypes.c_float))
    u, i, d = ctypes.c_uint, ctypes.c_int, ctypes.c_double
    for name in ("pcc2_set", "ccgn_set", "cc1b_set", "pcc1_set"):
        getattr(L, name).argtypes = [fpp, fpp, fpp, u, u, i, i]
        getattr(L, name).restype = i
    L.pcc_set.argtypes = [fpp, fpp, fpp, i, u, d, i, i]
    L.pcc_set.restype = i
    L.pcc1_highlevel.argtypes = [fpp, fpp, fpp, i, u, i, i, ctypes.c_void_p]
    L.pcc1_highlevel.restype = i
    L.pcc_highlevel.argtypes = [fpp, fpp, fpp, i, u, ctypes.c_float, i, i, ctypes.c_void_p]
    L.pcc_highlevel.restype = i
    L.pcc1_highlevel2.argtypes = [fpp, fpp, fpp, i, u, i, i]
    L.pcc1_highlevel2.restype = None
    L.tspcc2_set.argtypes = [fpp, fpp, fpp, u, u, i, i, d, d, u, i, d]
    L.tspcc2_set.restype = i
    L.cuda_tspcc2_set_freq.argtypes = [fpp, fpp, fpp, u, u, u, i, i, u, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint),
                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint), ctypes.c_float]
    L.cuda_tspcc2_set_freq.restype = i
    L.AnalyticSignal.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), u]
    L.AnalyticSignal.restype = i
    jp = ctypes.POINTER(FpccJob)
    L.fpcc_output_len.argtypes = [jp]
    L.fpcc_output_len.restype = ctypes.c_size_t
    L.fpcc_workspace_size.argtypes = [jp]
    L.fpcc_workspace_size.restype = ctypes.c_size_t
    L.fpcc_allpairs_device.argtypes = [jp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_void_p]
    L.fpcc_allpairs_device.restype = i
    L.fpcc_allpairs.argtypes = [jp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.fpcc_allpairs.restype = i
    L.fpcc_station_stage_batch_device.argtypes = [jp, i, ctypes.POINTER(ctypes.c_void_p),
                                                  ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_size_t),
                                                  ctypes.c_void_p]
    L.fpcc_station_stage_batch_device.restype = i
    L.fpcc_station_stage_device.argtypes = [jp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    L.fpcc_station_stage_device.restype = i
    L.fpcc_pair_stage_device.argtypes = [jp, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    L.fpcc_pair_stage_device.restype = i
    vp, ull = ctypes.c_void_p, ctypes.c_ulonglong
    L.fpcc_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    L.fpcc_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(ull)]
    L.fpcc_server_submit_many.argtypes = [vp, ctypes.POINTER(Request), i, ctypes.POINTER(ull)]
    L.fpcc_server_poll.argtypes = [vp, ull, ctypes.POINTER(RequestStatus)]
    L.fpcc_server_wait.argtypes = [vp, ull, i, ctypes.POINTER(RequestStatus)]
    L.fpcc_server_cancel.argtypes = [vp, ull]
    L.fpcc_server_forget.argtypes = [vp, ull]
    L.fpcc_server_stats.argtypes = [vp, ctypes.POINTER(ServingStats)]
    L.fpcc_server_destroy.argtypes = [vp]
    L.fpcc_graph_create.argtypes = [jp, vp, vp, vp, ctypes.c_size_t, vp, ctypes.POINTER(vp)]
    L.fpcc_graph_create.restype = i
    L.fpcc_graph_launch.argtypes = [vp, vp]
    L.fpcc_graph_launch.restype = i
    L.fpcc_graph_destroy.argtypes = [vp]
    L.fpcc_graph_destroy.restype = i
    pp = ctypes.POINTER(Preproc)
    L.fpcc_preprocess_workspace_size.argtypes = [jp, pp]
    L.fpcc_preprocess_workspace_size.restype = ctypes.c_size_t
    L.fpcc_preprocess_device.argtypes = [jp, pp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp]
    L.fpcc_preprocess_device.restype = i
    L.fpcc_preprocess.argtypes = [jp, pp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                  ctypes.POINTER(ctypes.c_ubyte), ctypes.POINTER(ctypes.c_float)]
    L.fpcc_preprocess.restype = i
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    L.fpcc_autotune_fft.argtypes = [i, i, i, ip, dp, ip, dp]
    L.fpcc_autotune_fft.restype = i
    L.fpcc_fft_split.argtypes = [i]
    L.fpcc_fft_split.restype = i
    L.fpcc_autotune_radix.argtypes = [i, i, i, ip, ip, dp, dp]
    L.fpcc_autotune_radix.restype = i
    L.fpcc_fft_radices.argtypes = [i, ip]
    L.fpcc_fft_radices.restype = i
    L.fpcc_host_alloc.argtypes = [ctypes.c_size_t]
    L.fpcc_host_alloc.restype = vp
    L.fpcc_host_free.argtypes = [vp]
    L.fpcc_host_free.restype = None
    L.fpcc_release_buffers.argtypes = []
    L.fpcc_release_buffers.restype = i
    L.fpcc_set_kept_limits.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
    L.fpcc_set_kept_limits.restype = None
    for name in ("create", "submit", "submit_many", "poll", "wait", "cancel", "forget", "stats", "destroy"):
        getattr(L, "fpcc_server_" + name).restype = i
    _lib = L
    return L
