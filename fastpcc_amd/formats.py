"""Python mirror of include/fastpcc_io.h: SAC and MSACS1 input, SAC and ts-PWS
binary output, trace pairing by start time (host code in libfastpcc.so)."""
import ctypes
import os

import numpy as np

from . import _lib
from .api import FastPCCError


class TraceHeader(ctypes.Structure):
    """Mirror of fpcc_trace_header (the reference's t_HeaderInfo)."""
    _fields_ = [
        ("npts", ctypes.c_int32), ("dt", ctypes.c_float), ("b", ctypes.c_float),
        ("year", ctypes.c_int32), ("yday", ctypes.c_int32), ("hour", ctypes.c_int32), ("min", ctypes.c_int32),
        ("sec", ctypes.c_int32), ("msec", ctypes.c_int32), ("t", ctypes.c_int64),
        ("stla", ctypes.c_float), ("stlo", ctypes.c_float), ("stel", ctypes.c_float), ("stdp", ctypes.c_float),
        ("cmpaz", ctypes.c_float), ("cmpinc", ctypes.c_float),
        ("sta", ctypes.c_char * 9), ("net", ctypes.c_char * 9), ("chn", ctypes.c_char * 9), ("loc", ctypes.c_char * 9),
        ("nostloc", ctypes.c_int32), ("nocmp", ctypes.c_int32),
    ]


_bound = False


def _L():
    global _bound
    L = _lib.lib()
    if not _bound:
        hp = ctypes.POINTER(TraceHeader)
        fp = ctypes.POINTER(ctypes.c_float)
        up = ctypes.POINTER(ctypes.c_uint)
        u, i, c = ctypes.c_uint, ctypes.c_int, ctypes.c_char_p
        L.fpcc_sac_read.argtypes = [c, u, hp, fp]
        L.fpcc_sac_read_list.argtypes = [c, up, u, hp, fp, up, fp]
        L.fpcc_msacs_read.argtypes = [c, u, hp, fp, up, up]
        L.fpcc_msacs_write.argtypes = [c, hp, fp, u, u]
        L.fpcc_sac_write_cc.argtypes = [c, c, ctypes.c_float, ctypes.c_float, fp, i, hp, hp]
        L.fpcc_obin_write.argtypes = [c, c, fp, u, i, ctypes.c_float, hp, hp, up, u]
        L.fpcc_sort_traces.argtypes = [hp, u, up]
        L.fpcc_pair_traces.argtypes = [hp, u, hp, u, up, up, up]
        L.fpcc_check_pairs.argtypes = [hp, u, hp, u]
        L.fpcc_polarity.argtypes = [hp]
        for n in ("fpcc_sac_read", "fpcc_sac_read_list", "fpcc_msacs_read", "fpcc_msacs_write", "fpcc_sac_write_cc",
                  "fpcc_obin_write", "fpcc_sort_traces", "fpcc_pair_traces", "fpcc_check_pairs", "fpcc_polarity"):
            getattr(L, n).restype = i
        _bound = True
    return L


def _b(s):
    return os.fsencode(s) if not isinstance(s, bytes) else s


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _check(r, where):
    if r != 0:
        raise FastPCCError(r, where)


def read_sac(path, nmax=0):
    """(TraceHeader, samples) of one SAC file."""
    h = TraceHeader()
    _check(_L().fpcc_sac_read(_b(path), nmax, ctypes.byref(h), None), "fpcc_sac_read")
    n = h.npts if nmax == 0 else min(h.npts, nmax)
    x = np.zeros(max(n, 0), dtype=np.float32)
    _check(_L().fpcc_sac_read(_b(path), nmax, ctypes.byref(h), _fp(x)), "fpcc_sac_read")
    return h, x


def read_sac_list(filelist, N=0):
    """(headers, x [traces][N], dt) of the SAC files listed in filelist."""
    n, Nc, dt = ctypes.c_uint(0), ctypes.c_uint(N), ctypes.c_float(0)
    _check(_L().fpcc_sac_read_list(_b(filelist), ctypes.byref(Nc), 0, None, None, ctypes.byref(n), ctypes.byref(dt)),
           "fpcc_sac_read_list")
    hdrs = (TraceHeader * max(n.value, 1))()
    x = np.zeros((n.value, Nc.value), dtype=np.float32)
    _check(_L().fpcc_sac_read_list(_b(filelist), ctypes.byref(Nc), n.value, hdrs, _fp(x), ctypes.byref(n),
                                   ctypes.byref(dt)), "fpcc_sac_read_list")
    return list(hdrs)[:n.value], x, dt.value


def read_msacs(path):
    n, N = ctypes.c_uint(0), ctypes.c_uint(0)
    _check(_L().fpcc_msacs_read(_b(path), 0, None, None, ctypes.byref(n), ctypes.byref(N)), "fpcc_msacs_read")
    hdrs = (TraceHeader * max(n.value, 1))()
    x = np.zeros((n.value, N.value), dtype=np.float32)
    _check(_L().fpcc_msacs_read(_b(path), n.value, hdrs, _fp(x), ctypes.byref(n), ctypes.byref(N)),
           "fpcc_msacs_read")
    return list(hdrs)[:n.value], x


def write_msacs(path, hdrs, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    arr = (TraceHeader * max(len(hdrs), 1))(*hdrs)
    _check(_L().fpcc_msacs_write(_b(path), arr, _fp(x), x.shape[0], x.shape[1]), "fpcc_msacs_write")


def write_sac_cc(path, method, beg, dt, y, h1, h2):
    y = np.ascontiguousarray(y, dtype=np.float32)
    _check(_L().fpcc_sac_write_cc(_b(path), _b(method), beg, dt, _fp(y), y.size, ctypes.byref(h1), ctypes.byref(h2)),
           "fpcc_sac_write_cc")


def write_obin(path, method, y, lag1, dt, h1, h2, sel=None):
    """y [traces][L]; h1, h2 lists of headers; sel: the traces of this channel pair."""
    y = np.ascontiguousarray(y, dtype=np.float32)
    sel = np.arange(y.shape[0], dtype=np.uint32) if sel is None else np.asarray(sel, dtype=np.uint32)
    a1 = (TraceHeader * len(h1))(*h1)
    a2 = (TraceHeader * len(h2))(*h2)
    _check(_L().fpcc_obin_write(_b(path), _b(method), _fp(y), y.shape[1], lag1, dt, a1, a2,
                                sel.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)), sel.size), "fpcc_obin_write")


def sort_traces(hdrs):
    arr = (TraceHeader * max(len(hdrs), 1))(*hdrs)
    perm = (ctypes.c_uint * max(len(hdrs), 1))()
    _check(_L().fpcc_sort_traces(arr, len(hdrs), perm), "fpcc_sort_traces")
    return list(perm)[:len(hdrs)]


def pair_traces(h1, h2):
    a1 = (TraceHeader * max(len(h1), 1))(*h1)
    a2 = (TraceHeader * max(len(h2), 1))(*h2)
    i1 = (ctypes.c_uint * max(len(h1), 1))()
    i2 = (ctypes.c_uint * max(len(h2), 1))()
    n = ctypes.c_uint(0)
    _check(_L().fpcc_pair_traces(a1, len(h1), a2, len(h2), i1, i2, ctypes.byref(n)), "fpcc_pair_traces")
    return list(i1)[:n.value], list(i2)[:n.value]


def check_pairs(h1, h2):
    a1 = (TraceHeader * max(len(h1), 1))(*h1)
    a2 = (TraceHeader * max(len(h2), 1))(*h2)
    return _L().fpcc_check_pairs(a1, len(h1), a2, len(h2))


def polarity(h):
    """(sign, updated header) per the reference's polarity rule."""
    hc = TraceHeader.from_buffer_copy(h)
    s = _L().fpcc_polarity(ctypes.byref(hc))
    return s, hc


def binary_dir():
    """Directory of the built C drivers (pcc_fullpair, filelist2msacs)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin")
