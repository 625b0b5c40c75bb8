"""Autotuner for the two-pass FFT split (run on the GPU it tunes for).

Transforms longer than one workgroup's LDS run as a column pass of n1 points and
a row pass of n2 = n / n1 points.  For each length this times every admissible
split on a batch of forward transforms (fpcc_autotune_fft, HIP events) and
writes the fastest to a table "n n1 ms_best heuristic_n1 ms_heuristic", which
the library reads at its first plan of each length:

    python -m fastpcc_amd.autotune --out fastpcc_amd/tuning/fft_splits_gfx950.txt

Lengths: the zero-padded sizes of long windows (powers of two 2^13..2^20, also the
Bluestein sizes) and the 7-smooth window lengths of hourly-to-daily traces at 1 Hz
that the analytic signal transforms at their own length.
"""
import argparse
import ctypes
import os
import sys
import time

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning", "fft_splits_gfx950.txt")


def _smooth7(n):
    for p in (2, 3, 5, 7):
        while n % p == 0:
            n //= p
    return n == 1


DEFAULT_SIZES = sorted(set([1 << k for k in range(13, 21)] +
                           [m for m in range(7200, 86401, 1800) if _smooth7(m)]))


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
    ap.add_argument("--sizes", default=None, help="comma-separated lengths (default: the built-in list)")
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args(argv)
    sizes = [int(s) for s in a.sizes.split(",")] if a.sizes else DEFAULT_SIZES
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
