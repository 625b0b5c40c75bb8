"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the CPU oracle.

The oracle restates the CPU branches of the reference's src/FFTapps.c and
src/FWTa/wavelet_def_v7.c; see oracle.h for the per-function citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

PCC2, PCC1, PCCV, GNCC, CC1B, WPCC = range(6)
KINDS = {"pcc2": PCC2, "pcc1": PCC1, "pccv": PCCV, "gncc": GNCC, "cc1b": CC1B, "wpcc": WPCC}

_lib = None


def build(force=False):
    """Compile liboracle.so with the oracle's Makefile (gcc)."""
    args = ["make", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        u, i, d = ctypes.c_uint, ctypes.c_int, ctypes.c_double
        L.orc_run_f32.argtypes = [i, fp, fp, fp, u, u, i, i, d, d, d, u, i, d]
        L.orc_run_f64.argtypes = [i, dp, fp, fp, u, u, i, i, d, d, d, u, i, d]
        L.orc_allpairs_f32.argtypes = [i, fp, fp, u, u, u, i, i, i, u, u, d, d, d, u, i, d]
        L.orc_allpairs_f64.argtypes = [i, dp, fp, u, u, u, i, i, i, u, u, d, d, d, u, i, d]
        L.orc_analytic_f.argtypes = [ctypes.c_void_p, fp, u]
        L.orc_analytic_d.argtypes = [ctypes.c_void_p, dp, u]
        L.orc_ampnormf.argtypes = [ctypes.c_void_p, u]
        L.orc_ampnorm.argtypes = [ctypes.c_void_p, u]
        L.orc_max_threads.restype = i
        L.orc_erfi.argtypes = [d]
        L.orc_erfi.restype = d
        L.orc_wpcc_family_export.argtypes = [u, i, i, d, d, u, i, d, ctypes.POINTER(ctypes.c_uint),
                                             dp, dp, ctypes.POINTER(ctypes.c_uint),
                                             ctypes.POINTER(ctypes.c_int), dp, u]
        L.orc_wpcc_bank_f64.argtypes = [dp, fp, fp, u, u, u, i, i, u, dp, dp, ctypes.POINTER(ctypes.c_uint),
                                         ctypes.POINTER(ctypes.c_int), d]
        L.orc_pws_f64.argtypes = [i, dp, fp, u, u, u, i, i, u, u, d]
        L.orc_qabsmedian.argtypes = [fp, i]
        L.orc_qabsmedian.restype = ctypes.c_float
        L.orc_qmedian.argtypes = [fp, i]
        L.orc_qmedian.restype = ctypes.c_float
        L.orc_clip_threshold.argtypes = [fp, u]
        L.orc_clip_threshold.restype = ctypes.c_float
        L.orc_clipping.argtypes = [fp, u]
        L.orc_trace_rms.argtypes = [fp, u]
        L.orc_trace_rms.restype = ctypes.c_float
        L.orc_white_nz.argtypes = [u]
        L.orc_white_nz.restype = u
        L.orc_white_filter.argtypes = [dp, dp, u, u, u, u, u]
        L.orc_avewhite.argtypes = [ctypes.POINTER(fp), u, u, dp, d]
        L.orc_preprocess.argtypes = [fp, u, u, u, ctypes.c_ulonglong, i, ctypes.c_float, d, d, d, i, fp,
                                     ctypes.POINTER(ctypes.c_ubyte), fp]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def out_len(kind, lag1, lag2):
    if kind == WPCC:
        return abs(lag2 - lag1 + 1)
    return abs(lag2 - lag1) + 1


def run(kind, x1, x2, lag1, lag2, f64=False, v=2.0, pmin=0.0, pmax=0.0, V=8, wtype=-1, op1=0.0):
    """One pair over Tr windows, like the reference's *_set: x1, x2 are [Tr][N]."""
    kind = KINDS.get(kind, kind)
    x1 = _f32(np.atleast_2d(x1))
    x2 = _f32(np.atleast_2d(x2))
    Tr, N = x1.shape
    L = out_len(kind, lag1, lag2)
    y = np.zeros((Tr, L), dtype=np.float64 if f64 else np.float32)
    fn = lib().orc_run_f64 if f64 else lib().orc_run_f32
    r = fn(kind, _ptr(y, ctypes.c_double if f64 else ctypes.c_float), _ptr(x1, ctypes.c_float),
           _ptr(x2, ctypes.c_float), N, Tr, lag1, lag2, v, pmin, pmax, V, wtype, op1)
    if r != 0:
        raise ValueError(f"oracle returned {r}")
    return y


def allpairs(kind, x, W, N, lag1, lag2, stacked=True, f64=False, pair_begin=0, pair_end=None,
             v=2.0, pmin=0.0, pmax=0.0, V=8, wtype=-1, op1=0.0):
    """All station pairs (i<j, row-major upper triangle) over W windows of N samples."""
    kind = KINDS.get(kind, kind)
    x = _f32(x)
    S = x.shape[0]
    P = S * (S - 1) // 2
    if pair_end is None or pair_end > P:
        pair_end = P
    L = out_len(kind, lag1, lag2)
    shape = (pair_end - pair_begin, L) if stacked else (pair_end - pair_begin, W, L)
    y = np.zeros(shape, dtype=np.float64 if f64 else np.float32)
    fn = lib().orc_allpairs_f64 if f64 else lib().orc_allpairs_f32
    r = fn(kind, _ptr(y, ctypes.c_double if f64 else ctypes.c_float), _ptr(x, ctypes.c_float), S, W, N,
           lag1, lag2, 1 if stacked else 0, pair_begin, pair_end, v, pmin, pmax, V, wtype, op1)
    if r != 0:
        raise ValueError(f"oracle returned {r}")
    return y


def pws(kind, x, W, N, lag1, lag2, power, pair_begin=0, pair_end=None):
    """FP64 phase-weighted stack over W windows of pairs [pair_begin, pair_end): [pairs][L]."""
    kind = KINDS.get(kind, kind)
    x = _f32(x)
    S = x.shape[0]
    P = S * (S - 1) // 2
    if pair_end is None or pair_end > P:
        pair_end = P
    y = np.zeros((pair_end - pair_begin, out_len(kind, lag1, lag2)))
    r = lib().orc_pws_f64(kind, _ptr(y, ctypes.c_double), _ptr(x, ctypes.c_float), S, W, N, lag1, lag2,
                          pair_begin, pair_end, power)
    if r != 0:
        raise ValueError(f"oracle returned {r}")
    return y


def max_threads():
    return lib().orc_max_threads()


def analytic(x, f64=False):
    x = np.ascontiguousarray(x, dtype=np.float64 if f64 else np.float32)
    y = np.zeros(x.shape[0], dtype=np.complex128 if f64 else np.complex64)
    if f64:
        lib().orc_analytic_d(y.ctypes.data, _ptr(x, ctypes.c_double), x.shape[0])
    else:
        lib().orc_analytic_f(y.ctypes.data, _ptr(x, ctypes.c_float), x.shape[0])
    return y


def ampnormf(z):
    z = np.array(z, dtype=np.complex64)
    lib().orc_ampnormf(z.ctypes.data, z.shape[0])
    return z


def ampnorm(z):
    z = np.array(z, dtype=np.complex128)
    lib().orc_ampnorm(z.ctypes.data, z.shape[0])
    return z


def erfi(z):
    return lib().orc_erfi(float(z))


def wpcc_family(N, lag1, lag2, pmin, pmax, V, wtype=-1, op1=0.0):
    """The WPCC filter bank as tspcc2_set builds it: dict of S, J, V, Nz, Ls0, s0, K0, op1,
    scale[S], Ls[S], center[S], and the complex wavelets (list of arrays)."""
    meta = (ctypes.c_uint * 5)()
    fl = (ctypes.c_double * 3)()
    r = lib().orc_wpcc_family_export(N, lag1, lag2, pmin, pmax, V, wtype, op1, meta, fl,
                                     None, None, None, None, 0)
    if r < 0:
        raise ValueError(f"oracle returned {r}")
    S = meta[0]
    scale = np.zeros(S)
    Ls = np.zeros(S, dtype=np.uint32)
    center = np.zeros(S, dtype=np.int32)
    lib().orc_wpcc_family_export(N, lag1, lag2, pmin, pmax, V, wtype, op1, meta, fl,
                                 _ptr(scale, ctypes.c_double), _ptr(Ls, ctypes.c_uint),
                                 _ptr(center, ctypes.c_int), None, 0)
    tot = int(Ls.sum())
    wbuf = np.zeros(2 * tot)
    lib().orc_wpcc_family_export(N, lag1, lag2, pmin, pmax, V, wtype, op1, meta, fl,
                                 _ptr(scale, ctypes.c_double), _ptr(Ls, ctypes.c_uint),
                                 _ptr(center, ctypes.c_int), _ptr(wbuf, ctypes.c_double), 2 * tot)
    wc = wbuf.view(np.complex128)
    waves, off = [], 0
    for s in range(S):
        waves.append(wc[off:off + Ls[s]].copy())
        off += int(Ls[s])
    return dict(S=S, J=meta[1], V=meta[2], Nz=meta[3], Ls0=meta[4], s0=fl[0], K0=fl[1], op1=fl[2],
                scale=scale, Ls=Ls, center=center, wc=waves)


def wpcc_bank(x1, x2, Nz, lag1, lag2, scale, wc, Ls, center, C):
    """FP64 WPCC of one pair over Tr windows with a caller-built filter bank (the
    semantics of cuda_tspcc2_set_freq with the CPU phasor): y [Tr][|lag2-lag1|+1]."""
    x1 = _f32(np.atleast_2d(x1))
    x2 = _f32(np.atleast_2d(x2))
    Tr, N = x1.shape
    S = len(scale)
    L = abs(lag2 - lag1) + 1
    y = np.zeros((Tr, L))
    scale = np.ascontiguousarray(scale, dtype=np.float64)
    Ls = np.ascontiguousarray(Ls, dtype=np.uint32)
    center = np.ascontiguousarray(center, dtype=np.int32)
    wflat = np.ascontiguousarray(np.concatenate([np.asarray(w, dtype=np.complex128) for w in wc])).view(np.float64)
    r = lib().orc_wpcc_bank_f64(_ptr(y, ctypes.c_double), _ptr(x1, ctypes.c_float), _ptr(x2, ctypes.c_float), N, Tr,
                                Nz, lag1, lag2, S, _ptr(scale, ctypes.c_double), _ptr(wflat, ctypes.c_double),
                                _ptr(Ls, ctypes.c_uint), _ptr(center, ctypes.c_int), float(C))
    if r != 0:
        raise ValueError(f"oracle returned {r}")
    return y


# ---- per-station preprocessing (src/PCC_fullpair_1b.c:881-925, 1065-1251) ----

def qabsmedian(x):
    x = _f32(x)
    return float(lib().orc_qabsmedian(_ptr(x, ctypes.c_float), x.size))


def qmedian(x):
    x = _f32(x)
    return float(lib().orc_qmedian(_ptr(x, ctypes.c_float), x.size))


def clip_threshold(x):
    x = _f32(x)
    return np.float32(lib().orc_clip_threshold(_ptr(x, ctypes.c_float), x.size))


def clipping(x):
    y = _f32(x).copy()
    lib().orc_clipping(_ptr(y, ctypes.c_float), y.size)
    return y


def trace_rms(x):
    x = _f32(x)
    return np.float32(lib().orc_trace_rms(_ptr(x, ctypes.c_float), x.size))


def white_filter(H, Nz, N, n1, n2):
    """(H after the filter design, HS) for an averaged amplitude spectrum H[Nz/2+1]."""
    H = np.ascontiguousarray(H, dtype=np.float64).copy()
    HS = np.zeros_like(H)
    lib().orc_white_filter(_ptr(H, ctypes.c_double), _ptr(HS, ctypes.c_double), H.size, Nz, N, n1, n2)
    return H, HS


def avewhite(x, band, dt):
    """AveWhite over the rows of x [Tr][N] (a whitened copy)."""
    y = _f32(x).copy()
    rows = (ctypes.POINTER(ctypes.c_float) * y.shape[0])(
        *[y[t].ctypes.data_as(ctypes.POINTER(ctypes.c_float)) for t in range(y.shape[0])])
    b = np.asarray(band, dtype=np.float64)
    r = lib().orc_avewhite(rows, y.shape[1], y.shape[0], _ptr(b, ctypes.c_double), dt)
    if r:
        raise ValueError(f"oracle avewhite returned {r}")
    return y


def preprocess(x, W, N, clip=False, std=0.0, white=None, dt=1.0, common=False, sign=None):
    """The station preprocessing of x [S][>= W*N] (a copy): returns (x, valid [S][W],
    rms [S][W])."""
    y = _f32(x).copy()
    S = y.shape[0]
    valid = np.zeros((S, W), dtype=np.uint8)
    rms = np.zeros((S, W), dtype=np.float32)
    sg = None if sign is None else _f32(sign).reshape(S, W)
    f1, f2 = white if white else (0.0, 0.0)
    r = lib().orc_preprocess(_ptr(y, ctypes.c_float), S, W, N, y.shape[1], int(clip), std, f1, f2, dt, int(common),
                             None if sg is None else _ptr(sg, ctypes.c_float), _ptr(valid, ctypes.c_ubyte),
                             _ptr(rms, ctypes.c_float))
    if r:
        raise ValueError(f"oracle preprocess returned {r}")
    return y, valid.astype(bool), rms
