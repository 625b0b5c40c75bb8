"""Autotuner for the FFT shapes (run on the GPU it tunes for).

Radix order: every single-workgroup LDS FFT runs its passes in a radix order;
for each length this times every order (largest-first and smallest-first of every
multiset of supported radices with that product, and the default) on a batch of
forward transforms (fpcc_autotune_radix) and writes the fastest to
tuning/fft_radix_gfx950.txt ("n r1 r2 ... # ms_best ms_default").

Two-pass split: transforms longer than one workgroup's LDS run as a column pass of
n1 points and a row pass of n2 = n / n1 points; for each length this times every
admissible split (fpcc_autotune_fft, with the radix orders just tuned) and writes
the fastest to tuning/fft_splits_gfx950.txt ("n n1 ms_best heuristic_n1
ms_heuristic").  The library reads both tables at its first plan of each length:

    python -m fastpcc_amd.autotune

Split lengths: the zero-padded sizes of long windows (powers of two 2^13..2^20,
also the Bluestein sizes) and the 7-smooth window lengths of hourly-to-daily traces
at 1 Hz.  Radix lengths: the column and row lengths those splits can use, powers
of two 64..2048 (half spectra, pair transforms) and common window lengths.
"""
import argparse
import ctypes
import os
import sys
import time

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning", "fft_splits_gfx950.txt")
RADIX_TABLE = os.path.join(HERE, "tuning", "fft_radix_gfx950.txt")
FIXED_RADIX = (3600, 4096, 8192)  # compile-time kernels rely on their default order


def _smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


DEFAULT_SIZES = sorted(set([1 << k for k in range(13, 21)] +
                           [m for m in range(7200, 86401, 1800) if _smooth7(m)]))


def _splits(n):
    return [d for d in range(4, 513) if n % d == 0 and n // d <= 2048 and _smooth7(d) and _smooth7(n // d)]


def radix_sizes():
    """Column / row lengths of every admissible split of the split lengths, powers
    of two 64..2048 and common window lengths (minus the fixed ones)."""
    out = {1 << k for k in range(6, 12)} | {1800, 2700, 3000, 4800, 5400, 6000, 7200}
    for n in DEFAULT_SIZES:
        for d in _splits(n):
            out.update((d, n // d))
    return sorted(m for m in out if m not in FIXED_RADIX and 2 <= m <= 8192)


def tune_radix(n, reps=10, batch=None):
    """(best radices, best ms, default ms) for a batch of FP32 LDS FFTs of n."""
    b = batch or max(256, (64 << 20) // (8 * n))
    r = (ctypes.c_int * 16)()
    np_ = ctypes.c_int(0)
    ms, dms = ctypes.c_double(0), ctypes.c_double(0)
    rc = _lib.lib().fpcc_autotune_radix(n, b, reps, r, ctypes.byref(np_), ctypes.byref(ms), ctypes.byref(dms))
    if rc:
        raise RuntimeError(f"fpcc_autotune_radix({n}) returned {rc}")
    return list(r)[:np_.value], ms.value, dms.value


def read_radix_table(path=RADIX_TABLE):
    """{n: [radices]} from a radix table."""
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            line = line.split("#")[0].split()
            if len(line) >= 2:
                out[int(line[0])] = [int(v) for v in line[1:]]
    return out


def batch_for(n, budget_bytes=128 << 20):
    """Transforms per timing launch: about budget_bytes of complex64 data."""
    return max(8, budget_bytes // (8 * n))


def tune(n, reps=10, batch=None):
    """(best n1, best ms, heuristic n1, heuristic ms) per batch of transforms of n."""
    b = batch or batch_for(n)
    n1, h1 = ctypes.c_int(0), ctypes.c_int(0)
    ms, hms = ctypes.c_double(0), ctypes.c_double(0)
    r = _lib.lib().fpcc_autotune_fft(n, b, reps, ctypes.byref(n1), ctypes.byref(ms), ctypes.byref(h1),
                                     ctypes.byref(hms))
    if r:
        raise RuntimeError(f"fpcc_autotune_fft({n}) returned {r}")
    return n1.value, ms.value, h1.value, hms.value


def read_table(path=DEFAULT_TABLE):
    """{n: n1} from a tuning table."""
    out = {}
    if not os.path.exists(path):
        return out
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            parts = line.split()
            out[int(parts[0])] = int(parts[1])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--radix-out", default=RADIX_TABLE)
    ap.add_argument("--sizes", default=None, help="comma-separated split lengths (default: the built-in list)")
    ap.add_argument("--no-radix", action="store_true", help="tune the splits only")
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args(argv)
    sizes = [int(s) for s in a.sizes.split(",")] if a.sizes else DEFAULT_SIZES
    if not a.no_radix:
        rrows = []
        for n in radix_sizes():
            t = time.perf_counter()
            r, ms, dms = tune_radix(n, a.reps)
            rrows.append((n, r, ms, dms))
            print(f"radix n={n:5d} best {'x'.join(map(str, r)):>16s} {ms:.4f} ms | default {dms:.4f} ms "
                  f"({dms / ms:.2f}x) [{time.perf_counter() - t:.1f} s]", flush=True)
        os.makedirs(os.path.dirname(os.path.abspath(a.radix_out)), exist_ok=True)
        with open(a.radix_out, "w") as f:
            f.write("# FP32 LDS FFT radix order per length on gfx950: n r1 r2 ... # ms_best ms_default\n")
            f.write(f"# python -m fastpcc_amd.autotune --reps {a.reps}; batches of max(256, 2^23/n) transforms\n")
            for n, r, ms, dms in rrows:
                f.write(f"{n} {' '.join(map(str, r))} # {ms:.5f} {dms:.5f}\n")
        print(a.radix_out)
    rows = []
    for n in sizes:
        t = time.perf_counter()
        n1, ms, h1, hms = tune(n, a.reps)
        rows.append((n, n1, ms, h1, hms))
        print(f"n={n:8d} batch={batch_for(n):5d} best n1={n1:4d} {ms:.4f} ms | heuristic n1={h1:4d} {hms:.4f} ms "
              f"({hms / ms:.2f}x) [{time.perf_counter() - t:.1f} s]", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        f.write("# two-pass FFT split per length on gfx950: n n1 ms_best heuristic_n1 ms_heuristic\n")
        f.write(f"# python -m fastpcc_amd.autotune --reps {a.reps}; ms per forward batch of max(8, 2^24/n)\n")
        for n, n1, ms, h1, hms in rows:
            f.write(f"{n} {n1} {ms:.5f} {h1} {hms:.5f}\n")
    print(a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
