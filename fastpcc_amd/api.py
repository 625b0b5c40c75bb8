"""Host-side mirror of the reference's correlation interface (src/FFTapps.h).

Each function takes numpy arrays shaped like the reference's row-pointer arrays
([Tr][N] for x1 and x2) and returns y as [Tr][L], calling the HIP library
through its C ABI.  Return codes of the C layer are raised as exceptions with the
reference's meaning (-1 invalid/GPU failure, -2 unsupported option, -3 lags too
large).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import FpccJob

_ERR = {-1: "invalid argument or GPU failure", -2: "unsupported option", -3: "|lag| >= N or unsupported type"}


class FastPCCError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where} returned {code}: {_ERR.get(code, 'error')}")
        self.code = code


def _rows(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    ptrs = (ctypes.POINTER(ctypes.c_float) * a.shape[0])()
    for t in range(a.shape[0]):
        ptrs[t] = a[t].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    return a, ptrs


def _lag_len(lag1, lag2):
    return abs(lag2 - lag1) + 1


def _pair_set(name, x1, x2, lag1, lag2):
    x1 = np.atleast_2d(x1)
    x2 = np.atleast_2d(x2)
    if x1.shape != x2.shape:
        raise ValueError("x1 and x2 must have the same shape")
    Tr, N = x1.shape
    L = _lag_len(lag1, lag2)
    y = np.zeros((Tr, L), dtype=np.float32)
    _, py = _rows(y)
    a1, p1 = _rows(x1)
    a2, p2 = _rows(x2)
    r = getattr(_lib.lib(), name)(py, p1, p2, N, Tr, lag1, lag2)
    if r != 0:
        raise FastPCCError(r, name)
    return y


def analytic_signal(x):
    """FP64 analytic signal of one real trace (AnalyticSignal, src/FFTapps.c:45-80)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(x.size, dtype=np.complex128)
    r = _lib.lib().AnalyticSignal(ctypes.c_void_p(y.ctypes.data), x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  x.size)
    if r != 0:
        raise FastPCCError(r, "AnalyticSignal")
    return y


def pcc2_set(x1, x2, lag1, lag2):
    """PCC2 of one station pair over Tr windows (src/FFTapps.c:785-881)."""
    return _pair_set("pcc2_set", x1, x2, lag1, lag2)


def ccgn_set(x1, x2, lag1, lag2):
    """GNCC of one station pair over Tr windows (src/FFTapps.c:883-978)."""
    return _pair_set("ccgn_set", x1, x2, lag1, lag2)


def pcc1_set(x1, x2, lag1, lag2):
    """PCC1 of one station pair over Tr windows (src/FFTapps.c:611-783).  Rows
    hold Lag2-Lag1+1 values; the reference's lag rules apply (see include/fastpcc.h)."""
    return _time_set("pcc1_set", x1, x2, lag1, lag2)


def pcc_set(x1, x2, v, lag1, lag2):
    """PCC of power v of one station pair over Tr windows (src/FFTapps.c:475-608)."""
    return _time_set("pcc_set", x1, x2, lag1, lag2, v=v)


def pcc1_highlevel(z1, z2, lag1, lag2, y=None, anok=None):
    """Analytic-signal entry of the reference's GPU engine (src/ccs_cuda.h:10):
    z1, z2 are [Tr][N] complex64 analytic signals; returns y [Tr][L] (only the
    computable lags are written into y, which starts as zeros or the given array).
    anok: None, or the address of an array of Tr sem_t (row tr is read only after
    sem_wait(&anok[tr]) returns).  With anok, z1 and z2 must already be C-contiguous
    complex64 arrays: the library reads them in place while a producer fills them."""
    return _time_set("pcc1_highlevel", z1, z2, lag1, lag2, complex_in=True, y=y, anok=anok)


def pcc_highlevel(z1, z2, v, lag1, lag2, y=None, anok=None):
    """PCC-v analytic-signal entry (src/ccs_cuda.h:9); anok as pcc1_highlevel."""
    return _time_set("pcc_highlevel", z1, z2, lag1, lag2, v=v, complex_in=True, y=y, anok=anok)


def pcc1_highlevel2(x1, x2, lag1, lag2, y=None):
    """Real-input PCC1 entry of the reference's GPU engine (src/ccs_cuda.h:11,
    src/ccs_cuda.cu:796-870): returns nothing in C (errors are printed); only the
    computable lags of y are written."""
    return _time_set("pcc1_highlevel2", x1, x2, lag1, lag2, y=y)


def release_buffers():
    """Frees the buffers the host entries keep between calls; returns the number
    of calls still holding theirs."""
    return _lib.lib().fpcc_release_buffers()


def set_kept_limits(host_bytes, device_bytes):
    """Caps of the pinned host and device bytes kept per device between calls."""
    _lib.lib().fpcc_set_kept_limits(host_bytes, device_bytes)


def _time_set(name, x1, x2, lag1, lag2, v=None, complex_in=False, y=None, anok=None):
    dt = np.complex64 if complex_in else np.float32
    if anok is not None:
        for a in (x1, x2):
            if not (isinstance(a, np.ndarray) and a.dtype == dt and a.ndim == 2 and a.flags.c_contiguous):
                raise ValueError("with anok the inputs must be C-contiguous [Tr][N] arrays read in place")
    x1 = np.ascontiguousarray(np.atleast_2d(x1), dtype=dt)
    x2 = np.ascontiguousarray(np.atleast_2d(x2), dtype=dt)
    if x1.shape != x2.shape:
        raise ValueError("x1 and x2 must have the same shape")
    Tr, N = x1.shape
    L = lag2 - lag1 + 1
    if y is None:
        y = np.zeros((Tr, max(L, 0)), dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    _, py = _rows(y) if L > 0 else (None, (ctypes.POINTER(ctypes.c_float) * max(Tr, 1))())
    p1 = (ctypes.POINTER(ctypes.c_float) * Tr)(*[x1[t].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
                                                  for t in range(Tr)])
    p2 = (ctypes.POINTER(ctypes.c_float) * Tr)(*[x2[t].ctypes.data_as(ctypes.POINTER(ctypes.c_float))
                                                  for t in range(Tr)])
    fn = getattr(_lib.lib(), name)
    if name == "pcc_set":
        r = fn(py, p1, p2, N, Tr, v, lag1, lag2)
    elif name == "pcc_highlevel":
        r = fn(py, p1, p2, N, Tr, v, lag1, lag2, anok)
    elif name == "pcc1_highlevel":
        r = fn(py, p1, p2, N, Tr, lag1, lag2, anok)
    elif name == "pcc1_highlevel2":
        fn(py, p1, p2, N, Tr, lag1, lag2)
        r = 0
    else:
        r = fn(py, p1, p2, N, Tr, lag1, lag2)
    if r != 0:
        raise FastPCCError(r, name)
    return y


def tspcc2_set(x1, x2, lag1, lag2, pmin, pmax, V, wtype=-1, op1=0.0):
    """Wavelet PCC of one station pair over Tr windows (src/FFTapps.c:1076-1255);
    pmin/pmax in samples.  Rows hold abs(lag2-lag1+1) lags from min(lag1, lag2)."""
    x1 = np.atleast_2d(x1)
    x2 = np.atleast_2d(x2)
    Tr, N = x1.shape
    L = abs(lag2 - lag1 + 1)
    y = np.zeros((Tr, L), dtype=np.float32)
    _, py = _rows(y)
    a1, p1 = _rows(x1)
    a2, p2 = _rows(x2)
    r = _lib.lib().tspcc2_set(py, p1, p2, N, Tr, lag1, lag2, pmin, pmax, V, wtype, op1)
    if r != 0:
        raise FastPCCError(r, "tspcc2_set")
    return y


def cuda_tspcc2_set_freq(x1, x2, Nz, lag1, lag2, scale, wc, Ls, center, C, down_smp=None):
    """The reference GPU engine's WPCC entry (src/ccs_cuda.h:12) with a caller-built
    filter bank: scale[S], wavelets wc (list of complex128 arrays), Ls[S], center[S]."""
    x1 = np.atleast_2d(x1)
    x2 = np.atleast_2d(x2)
    Tr, N = x1.shape
    L = abs(lag2 - lag1) + 1
    y = np.zeros((Tr, L), dtype=np.float32)
    _, py = _rows(y)
    a1, p1 = _rows(x1)
    a2, p2 = _rows(x2)
    S = len(scale)
    sc = np.ascontiguousarray(scale, dtype=np.float64)
    waves = [np.ascontiguousarray(w, dtype=np.complex128) for w in wc]
    wp = (ctypes.c_void_p * S)(*[w.ctypes.data for w in waves])
    ls = np.ascontiguousarray(Ls, dtype=np.uint32)
    ce = np.ascontiguousarray(center, dtype=np.int32)
    ds = np.ascontiguousarray(np.ones(S) if down_smp is None else down_smp, dtype=np.uint32)
    r = _lib.lib().cuda_tspcc2_set_freq(py, p1, p2, N, Tr, Nz, lag1, lag2, S,
                                        sc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), wp,
                                        ls.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                                        ce.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                        ds.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), C)
    if r != 0:
        raise FastPCCError(r, "cuda_tspcc2_set_freq")
    return y


def cc1b_set(x1, x2, lag1, lag2):
    """1-bit GNCC of one station pair over Tr windows (src/FFTapps.c:980-1073)."""
    return _pair_set("cc1b_set", x1, x2, lag1, lag2)


KIND = {"pcc2": _lib.FPCC_PCC2, "pcc1": _lib.FPCC_PCC1, "pccv": _lib.FPCC_PCCV, "gncc": _lib.FPCC_GNCC,
        "cc1b": _lib.FPCC_CC1B, "wpcc": _lib.FPCC_WPCC}
STACK = {"none": _lib.STACK_NONE, "plain": _lib.STACK_PLAIN, "zsum": _lib.STACK_ZSUM, "gram": _lib.STACK_GRAM,
         "gram16": _lib.STACK_GRAM16, "pws": _lib.STACK_PWS}


def make_job(kind, n_stations, n_windows, win_len, lag1, lag2, stack="plain", station_stride=None, v=2.0,
             pmin=0.0, pmax=0.0, V=8, wtype=-1, op1=0.0, pws_power=2.0):
    job = FpccJob()
    job.kind = KIND.get(kind, kind)
    job.stack = STACK.get(stack, stack)
    job.n_stations = n_stations
    job.n_windows = n_windows
    job.win_len = win_len
    job.station_stride = station_stride if station_stride is not None else n_windows * win_len
    job.lag1, job.lag2 = lag1, lag2
    job.v, job.pmin, job.pmax, job.V, job.wtype, job.op1 = v, pmin, pmax, V, wtype, op1
    job.pws_power = pws_power
    return job


def output_shape(job):
    P = job.n_stations * (job.n_stations - 1) // 2
    L = _lag_len(job.lag1, job.lag2)
    return (P, job.n_windows, L) if job.stack == _lib.STACK_NONE else (P, L)


def allpairs(x, n_windows, win_len, lag1, lag2, kind="pcc2", stack="plain", **kw):
    """All station pairs from host data x [S][>= W*N]; returns y on the host."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    job = make_job(kind, x.shape[0], n_windows, win_len, lag1, lag2, stack, station_stride=x.shape[1], **kw)
    n = _lib.lib().fpcc_output_len(ctypes.byref(job))
    if n == 0:
        raise FastPCCError(-1, "fpcc_output_len")
    y = np.zeros(output_shape(job), dtype=np.float32)
    r = _lib.lib().fpcc_allpairs(ctypes.byref(job), x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                 y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    if r != 0:
        raise FastPCCError(r, "fpcc_allpairs")
    return y


def make_preproc(clip=False, std=0.0, white=None, dt=1.0, common=False):
    """fpcc_preproc: clip (bool), std (outlier factor, 0 off), white ((f1, f2) Hz
    or None), dt (s), common (keep windows kept at every station)."""
    p = _lib.Preproc()
    p.clip = int(bool(clip))
    p.std_max = float(std)
    if white:
        p.band[0], p.band[1] = float(white[0]), float(white[1])
    p.dt = float(dt)
    p.common_windows = int(bool(common))
    return p


def preprocess(x, n_windows, win_len, clip=False, std=0.0, white=None, dt=1.0, common=False, sign=None):
    """Per-station preprocessing of host data x [S][>= W*N] (fpcc_preprocess): returns
    (preprocessed copy of x, valid [S][W] bool, rms [S][W])."""
    y = np.array(x, dtype=np.float32, order="C", copy=True)
    S = y.shape[0]
    job = make_job("pcc2", S, n_windows, win_len, 0, 0, station_stride=y.shape[1])
    pp = make_preproc(clip, std, white, dt, common)
    valid = np.zeros((S, n_windows), dtype=np.uint8)
    rms = np.zeros((S, n_windows), dtype=np.float32)
    sg = None if sign is None else np.ascontiguousarray(sign, dtype=np.float32).reshape(S, n_windows)
    fp = ctypes.POINTER(ctypes.c_float)
    r = _lib.lib().fpcc_preprocess(ctypes.byref(job), ctypes.byref(pp), y.ctypes.data_as(fp),
                                   None if sg is None else sg.ctypes.data_as(fp),
                                   valid.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)), rms.ctypes.data_as(fp))
    if r != 0:
        raise FastPCCError(r, "fpcc_preprocess")
    return y, valid.astype(bool), rms


class DeviceRunner:
    """Device-resident all-pairs runs on torch-allocated HBM (torch is plumbing:
    allocation and the stream handle).  The workspace is allocated once."""

    def __init__(self, job, device="cuda:0"):
        import torch
        self.torch = torch
        self.job = job
        self.device = torch.device(device)
        L = _lib.lib()
        self.nout = L.fpcc_output_len(ctypes.byref(job))
        if self.nout == 0:
            raise FastPCCError(-1, "fpcc_output_len")
        self.ws_bytes = L.fpcc_workspace_size(ctypes.byref(job))
        self.ws = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=self.device)
        self.y = torch.empty(self.nout, dtype=torch.float32, device=self.device)

    def run(self, x_dev, stream=None):
        torch = self.torch
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        assert x_dev.dtype == torch.float32 and x_dev.is_contiguous() and x_dev.device == self.device
        need = (self.job.n_stations - 1) * self.job.station_stride + self.job.n_windows * self.job.win_len
        if x_dev.numel() < need:
            raise ValueError("input tensor too small for the job")
        r = _lib.lib().fpcc_allpairs_device(ctypes.byref(self.job), ctypes.c_void_p(x_dev.data_ptr()),
                                            ctypes.c_void_p(self.y.data_ptr()), ctypes.c_void_p(self.ws.data_ptr()),
                                            self.ws_bytes, ctypes.c_void_p(stream.cuda_stream))
        if r != 0:
            raise FastPCCError(r, "fpcc_allpairs_device")
        return self.y.view(output_shape(self.job))

    def capture(self, x_dev, stream):
        """Captures the job on `stream` (a non-default torch stream) as a HIP graph
        reading x_dev; replay() then recomputes y from x_dev's current contents."""
        self.release_graph()
        g = ctypes.c_void_p()
        r = _lib.lib().fpcc_graph_create(ctypes.byref(self.job), ctypes.c_void_p(x_dev.data_ptr()),
                                         ctypes.c_void_p(self.y.data_ptr()), ctypes.c_void_p(self.ws.data_ptr()),
                                         self.ws_bytes, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(g))
        if r != 0:
            raise FastPCCError(r, "fpcc_graph_create")
        self._graph, self._gx = g, x_dev
        return self

    def replay(self, stream):
        r = _lib.lib().fpcc_graph_launch(self._graph, ctypes.c_void_p(stream.cuda_stream))
        if r != 0:
            raise FastPCCError(r, "fpcc_graph_launch")
        return self.y.view(output_shape(self.job))

    def release_graph(self):
        if getattr(self, "_graph", None):
            _lib.lib().fpcc_graph_destroy(self._graph)
            self._graph = None

    def __del__(self):
        try:
            self.release_graph()
        except Exception:
            pass


_ERR[-4] = "request exceeds the server's memory budget"


class _PinnedBlock:
    def __init__(self, nbytes):
        self.p = _lib.lib().fpcc_host_alloc(nbytes)
        if not self.p:
            raise FastPCCError(-1, "fpcc_host_alloc")

    def __del__(self):
        if self.p:
            _lib.lib().fpcc_host_free(self.p)
            self.p = None


def pinned_empty(shape, dtype=np.float32):
    """numpy array in pinned host memory of the library's HIP runtime (freed with
    the last view), for Server buffers copied by DMA."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape))
    nbytes = max(n * dtype.itemsize, 1)
    blk = _PinnedBlock(nbytes)
    buf = (ctypes.c_byte * nbytes).from_address(blk.p)
    buf._owner = blk
    return np.frombuffer(buf, dtype=dtype, count=n).reshape(shape)


class Server:
    """Request server over fpcc_server_* (include/fastpcc.h): continuous batching
    of all-pairs jobs with admission by HBM footprint.  submit() keeps references
    to x and y until the request is finished; y is filled in place."""

    def __init__(self, device_bytes=0, pinned_bytes=0, max_inflight=0, n_streams=0, max_bypass=0, fail_every=0,
                 cache_bytes=0):
        cfg = _lib.ServerConfig(device_bytes, pinned_bytes, max_inflight, n_streams, max_bypass, fail_every,
                                cache_bytes)
        h = ctypes.c_void_p()
        r = _lib.lib().fpcc_server_create(ctypes.byref(cfg), ctypes.byref(h))
        if r != 0:
            raise FastPCCError(r, "fpcc_server_create")
        self._h = h
        self._keep = {}

    def submit(self, job, x, y, keys=None):
        """x: float32 array whose rows (station_stride apart) hold the job's input;
        y: float32 array of fpcc_output_len(job) elements; keys: optional per-station
        64-bit identities for the resident-spectra cache.  Returns the request id."""
        assert x.dtype == np.float32 and x.flags.c_contiguous and y.dtype == np.float32 and y.flags.c_contiguous
        need = (job.n_stations - 1) * job.station_stride + job.n_windows * job.win_len
        if x.size < need or y.size < _lib.lib().fpcc_output_len(ctypes.byref(job)):
            raise ValueError("x or y too small for the job")
        kp = None
        if keys is not None:
            keys = np.ascontiguousarray(keys, dtype=np.uint64)
            if keys.size != job.n_stations:
                raise ValueError("one key per station")
            kp = keys.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong))
        req = _lib.Request(job, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                           y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), kp)
        rid = ctypes.c_ulonglong()
        r = _lib.lib().fpcc_server_submit(self._h, ctypes.byref(req), ctypes.byref(rid))
        if r != 0:
            raise FastPCCError(r, "fpcc_server_submit")
        self._keep[rid.value] = (x, y, keys)
        return rid.value

    def submit_many(self, items):
        """items: (job, x, y) or (job, x, y, keys) tuples entering the queue together.
        Returns the request ids."""
        reqs = (_lib.Request * len(items))()
        keep = []
        for k, it in enumerate(items):
            job, x, y = it[:3]
            keys = it[3] if len(it) > 3 else None
            assert x.dtype == np.float32 and x.flags.c_contiguous and y.dtype == np.float32 and y.flags.c_contiguous
            kp = None
            if keys is not None:
                keys = np.ascontiguousarray(keys, dtype=np.uint64)
                kp = keys.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong))
            reqs[k] = _lib.Request(job, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                   y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), kp)
            keep.append((x, y, keys))
        ids = (ctypes.c_ulonglong * len(items))()
        r = _lib.lib().fpcc_server_submit_many(self._h, reqs, len(items), ids)
        if r != 0:
            raise FastPCCError(r, "fpcc_server_submit_many")
        for k in range(len(items)):
            self._keep[ids[k]] = keep[k]
        return list(ids)

    def _status(self, st):
        return {"state": st.state, "rc": st.rc, "t_submit": st.t_submit, "t_start": st.t_start,
                "t_done": st.t_done}

    def poll(self, rid):
        st = _lib.RequestStatus()
        r = _lib.lib().fpcc_server_poll(self._h, rid, ctypes.byref(st))
        if r != 0:
            raise FastPCCError(r, "fpcc_server_poll")
        return self._status(st)

    def wait(self, rid, timeout_ms=-1):
        """Status dict once the request is final; None on timeout."""
        st = _lib.RequestStatus()
        r = _lib.lib().fpcc_server_wait(self._h, rid, timeout_ms, ctypes.byref(st))
        if r == 1:
            return None
        if r != 0:
            raise FastPCCError(r, "fpcc_server_wait")
        if st.state >= _lib.REQ_DONE:
            self._keep.pop(rid, None)
        return self._status(st)

    def cancel(self, rid):
        return _lib.lib().fpcc_server_cancel(self._h, rid)

    def forget(self, rid):
        return _lib.lib().fpcc_server_forget(self._h, rid)

    def stats(self):
        s = _lib.ServingStats()
        r = _lib.lib().fpcc_server_stats(self._h, ctypes.byref(s))
        if r != 0:
            raise FastPCCError(r, "fpcc_server_stats")
        return {n: getattr(s, n) for n, _ in s._fields_}

    def close(self):
        if self._h:
            _lib.lib().fpcc_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
