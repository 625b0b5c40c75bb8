"""Seeded synthetic station records: white noise plus a delayed common source.

Each station s records x_s[n] = noise_s[n] + snr * src[n - shift_s], where noise
is N(0,1), src is N(0,1) noise band-passed by a windowed-sinc FIR (Blackman,
default 0.05-0.2 Hz at 1 Hz sampling) and scaled to unit standard deviation,
and shift_s is a uniform integer delay in [-max_shift, max_shift].  The source is
drawn over n_samples + 2*max_shift samples and sliced, so delays are true delays:
the correlation of stations i and j peaks at lag shift_j - shift_i.
"""
import numpy as np


def bandpass_fir(fs=1.0, band=(0.05, 0.2), ntaps=201):
    """Windowed-sinc band-pass FIR (Blackman window), unit gain at the band centre."""
    n = np.arange(ntaps) - (ntaps - 1) / 2
    f1, f2 = band[0] / fs, band[1] / fs
    h = 2 * f2 * np.sinc(2 * f2 * n) - 2 * f1 * np.sinc(2 * f1 * n)
    h *= np.blackman(ntaps)
    fc = 0.5 * (f1 + f2)
    gain = np.abs(np.sum(h * np.exp(-2j * np.pi * fc * np.arange(ntaps))))
    return h / gain


def make_stations(n_stations, n_samples, max_shift, snr=0.5, seed=0, fs=1.0, band=(0.05, 0.2), ntaps=201):
    """Returns (x float32 [S][n_samples], shifts int [S])."""
    rng = np.random.default_rng(seed)
    shifts = rng.integers(-max_shift, max_shift + 1, n_stations)
    src_len = n_samples + 2 * max_shift
    raw = rng.standard_normal(src_len + ntaps - 1)
    src = np.convolve(raw, bandpass_fir(fs, band, ntaps), mode="valid")
    src /= src.std()
    x = np.empty((n_stations, n_samples), dtype=np.float32)
    for s in range(n_stations):
        start = max_shift - int(shifts[s])
        x[s] = rng.standard_normal(n_samples) + snr * src[start:start + n_samples]
    return x, shifts


def pair_index(i, j, S):
    """Row-major upper-triangle index of station pair (i<j)."""
    return i * S - i * (i + 1) // 2 + (j - i - 1)


def pair_list(S):
    """(i, j) for p = 0..P-1 in the library's pair order."""
    return [(i, j) for i in range(S) for j in range(i + 1, S)]


def morlet_bank(scales, Nz, w0=0.0):
    """A caller-built Morlet filter bank for cuda_tspcc2_set_freq, from any list of
    scales (in samples): per scale s the support Ls = min(2 ceil(5 s) + 1, Nz) and
    origin center = Ls // 2 (src/FWTa/wavelet_def_v7.c:312-322), samples
    w[u] = (sqrt(pi) s)^(-1/2) exp(i w0 t) exp(-t^2/2), t = (u - center)/s (the
    Morlet of :38-52; w0 = pi sqrt(2/ln 2) when 0).  Returns dict(scale, wc, Ls,
    center, K0) with K0 = sum 1/s, the normalisation tspcc2_set uses (C = 1/(K0 N))."""
    if w0 == 0.0:
        w0 = np.pi * np.sqrt(2 / np.log(2))
    scales = np.asarray(scales, dtype=np.float64)
    wc, Ls, center = [], [], []
    for s in scales:
        n = min(2 * int(np.ceil(5 * s)) + 1, Nz)
        c = n // 2
        t = (np.arange(n) - c) / s
        k = 1 / np.sqrt(np.sqrt(np.pi) * s)
        wc.append(k * np.exp(1j * w0 * t) * np.exp(-0.5 * t * t))
        Ls.append(n)
        center.append(c)
    return dict(scale=scales, wc=wc, Ls=np.array(Ls, dtype=np.uint32), center=np.array(center, dtype=np.int32),
                K0=float(np.sum(1 / scales)))
