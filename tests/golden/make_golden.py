"""Writes the config-1 golden fixtures (8 stations x 3600 samples, 28 pairs,
lags -500..500) with an FP64 numpy route independent of the oracle's C code:
analytic signals from scipy.signal.hilbert, phasors z/(|z| + 1e-6 max|z|), and
direct (non-FFT) correlation sums from numpy.correlate.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.signal as ss

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from fastpcc_amd import synth  # noqa: E402

S, N, LAG = 8, 3600, 500


def phasor(x):
    z = ss.hilbert(x.astype(np.float64))
    return z / (np.abs(z) + 1e-6 * np.abs(z).max())


def main():
    x, shifts = synth.make_stations(S, N, 200, seed=20190401)
    lags = np.arange(-LAG, LAG + 1)
    cnt = N - np.abs(lags)
    pcc2, gncc, cc1b = [], [], []
    for i, j in synth.pair_list(S):
        a, b = x[i].astype(np.float64), x[j].astype(np.float64)
        # numpy.correlate(b, a, 'full')[k + N - 1] = sum_n b[n + k] conj(a[n])
        full = np.correlate(phasor(b), phasor(a), "full")
        pcc2.append(np.real(full[lags + N - 1]) / N)
        c = np.correlate(b, a, "full")[lags + N - 1]
        e1 = np.array([np.sum(a[max(0, -l):min(N, N - l)] ** 2) for l in lags])
        e2 = np.array([np.sum(b[max(0, l):min(N, N + l)] ** 2) for l in lags])
        gncc.append(c / np.sqrt(e1 * e2))
        sa, sb = np.where(a >= 0, 1.0, -1.0), np.where(b > 0, 1.0, -1.0)
        cc1b.append(np.correlate(sb, sa, "full")[lags + N - 1] / cnt)
    np.save(os.path.join(HERE, "c1_x.npy"), x)
    np.save(os.path.join(HERE, "c1_shifts.npy"), shifts.astype(np.int64))
    np.save(os.path.join(HERE, "c1_pcc2_fp64.npy"), np.array(pcc2))
    np.save(os.path.join(HERE, "c1_gncc_fp64.npy"), np.array(gncc))
    np.save(os.path.join(HERE, "c1_cc1b_fp64.npy"), np.array(cc1b))


if __name__ == "__main__":
    main()
