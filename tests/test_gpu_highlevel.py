"""The reference GPU-engine entries under the reference's threading contract
(src/ccs_cuda.h:8-19).

pcc_highlevel / pcc1_highlevel run in the reference beside a CPU producer team
that computes each trace's analytic signal and posts anok[tr] when it is ready
(src/FFTapps.c:636-680, 518-543); the consumer waits on anok[tr] before it reads
row tr (src/ccs_cuda.cu:575, 693).  Here a Python producer thread fills the rows
in place, late and in trace order, and posts each row's semaphore; the library
must give bit for bit the result of the same call on the finished rows with
anok = NULL (a row read before its post would still hold zeros).

pcc1_highlevel2 (src/ccs_cuda.cu:796-870) takes real traces; it is checked
against the oracle's PCC1 (src/FFTapps.c:611-783) on the lags it writes.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
TOL = 1e-5
SEM_T_BYTES = 32  # sizeof(sem_t) on x86-64 glibc

_libc = ctypes.CDLL("libc.so.6", use_errno=True)
_libc.sem_init.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint]
_libc.sem_post.argtypes = [ctypes.c_void_p]
_libc.sem_destroy.argtypes = [ctypes.c_void_p]


class SemArray:
    """Tr process-private POSIX semaphores, initially 0."""

    def __init__(self, n):
        self.n = n
        self.buf = ctypes.create_string_buffer(SEM_T_BYTES * max(n, 1))
        self.base = ctypes.addressof(self.buf)
        assert self.base % 8 == 0
        for k in range(n):
            assert _libc.sem_init(self.base + SEM_T_BYTES * k, 0, 0) == 0

    def post(self, k):
        assert _libc.sem_post(self.base + SEM_T_BYTES * k) == 0

    def close(self):
        for k in range(self.n):
            _libc.sem_destroy(self.base + SEM_T_BYTES * k)


def _analytic_rows(orc, x):
    return np.array([orc.analytic(r) for r in x], dtype=np.complex64)


def _producer(z1, z2, src1, src2, sems, delay_s):
    """Fills row tr of z1 and z2 and posts anok[tr], in trace order."""
    for tr in range(z1.shape[0]):
        time.sleep(delay_s)
        z1[tr] = src1[tr]
        z2[tr] = src2[tr]
        sems.post(tr)


@pytest.mark.parametrize("entry,Tr,N,lag1,lag2", [
    ("pcc1", 100, 21600, -500, 500),   # 34.6 MB of complex rows: several 8 MB staging pieces
    ("pccv", 24, 3600, -1000, 1000),
])
def test_highlevel_waits_for_late_rows(fp, orc, entry, Tr, N, lag1, lag2):
    print(f"{entry}_highlevel with anok: Tr={Tr} N={N} lags=({lag1},{lag2})", flush=True)
    rng = np.random.default_rng(Tr * 7 + N)
    x1 = rng.standard_normal((Tr, N)).astype(np.float32)
    x2 = rng.standard_normal((Tr, N)).astype(np.float32)
    a1, a2 = _analytic_rows(orc, x1), _analytic_rows(orc, x2)
    z1 = np.zeros((Tr, N), np.complex64)
    z2 = np.zeros((Tr, N), np.complex64)
    sems = SemArray(Tr)
    prod = threading.Thread(target=_producer, args=(z1, z2, a1, a2, sems, 0.002))
    prod.start()
    try:
        if entry == "pcc1":
            y_wait = fp.pcc1_highlevel(z1, z2, lag1, lag2, anok=sems.base)
        else:
            y_wait = fp.pcc_highlevel(z1, z2, 1.5, lag1, lag2, anok=sems.base)
    finally:
        prod.join()
        sems.close()
    assert np.array_equal(z1, a1) and np.array_equal(z2, a2)
    if entry == "pcc1":
        y_now = fp.pcc1_highlevel(a1, a2, lag1, lag2)
    else:
        y_now = fp.pcc_highlevel(a1, a2, 1.5, lag1, lag2)
    d = np.abs(y_wait - y_now).max()
    print(f"  max|anok - NULL| = {d:.3e}", flush=True)
    assert np.array_equal(y_wait, y_now)
    rows = [0, Tr // 2, Tr - 1]
    if entry == "pcc1":
        yr = orc.run("pcc1", x1[rows], x2[rows], lag1, lag2)
    else:
        yr = orc.run("pccv", x1[rows], x2[rows], lag1, lag2, v=1.5)
    err = np.abs(y_wait[rows] - yr).max()
    print(f"  max|gpu-oracle| (rows {rows}) = {err:.3e}", flush=True)
    assert err <= TOL


def test_highlevel_anok_already_posted(fp, orc):
    """Semaphores posted before the call: the waits return at once."""
    Tr, N = 5, 900
    rng = np.random.default_rng(5)
    x1 = rng.standard_normal((Tr, N)).astype(np.float32)
    x2 = rng.standard_normal((Tr, N)).astype(np.float32)
    a1, a2 = _analytic_rows(orc, x1), _analytic_rows(orc, x2)
    sems = SemArray(Tr)
    for k in range(Tr):
        sems.post(k)
    try:
        y = fp.pcc1_highlevel(a1, a2, -200, 200, anok=sems.base)
    finally:
        sems.close()
    assert np.array_equal(y, fp.pcc1_highlevel(a1, a2, -200, 200))


@pytest.mark.parametrize("N,Tr,lag1,lag2", [(1000, 3, -300, 300), (3600, 2, -1000, 1000), (100, 2, -120, 130),
                                            (64, 2, -80, 10)])
def test_pcc1_highlevel2_matches_oracle(fp, orc, N, Tr, lag1, lag2):
    """Real-input PCC1: the lags l in [l1, L') are written (l1 > 0 when lag1 < -N,
    L' < L when lag2 > N, src/ccs_cuda.cu:807-808); the rest of y is untouched."""
    print(f"pcc1_highlevel2 N={N} Tr={Tr} lags=({lag1},{lag2})", flush=True)
    rng = np.random.default_rng(N + Tr)
    x1 = rng.standard_normal((Tr, N)).astype(np.float32)
    x2 = rng.standard_normal((Tr, N)).astype(np.float32)
    L = lag2 - lag1 + 1
    sentinel = np.float32(7.0)
    y = np.full((Tr, L), sentinel, np.float32)
    y = fp.pcc1_highlevel2(x1, x2, lag1, lag2, y=y)
    l1 = 0 if lag1 >= -N else -(lag1 + N)
    Lp = L - (lag2 - N if lag2 > N else 0)
    assert np.all(y[:, :l1] == sentinel) and np.all(y[:, Lp:] == sentinel)
    yr = orc.run("pcc1", x1, x2, lag1, lag2)
    a, b = y[:, l1:Lp], yr[:, l1:Lp]
    na, nb = np.isnan(a), np.isnan(b)
    assert np.array_equal(na, nb)
    err = np.abs(a[~na] - b[~nb]).max()
    print(f"  max|gpu-oracle| = {err:.3e}", flush=True)
    assert err <= TOL
