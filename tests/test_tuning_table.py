"""The committed FFT tuning tables: the split table lists only admissible splits
(n = n1 * n2, n1 <= 512, n2 <= 2048, both 7-smooth), one per length, each no
slower than the heuristic's split when it was measured; the radix table lists
orders of supported radices whose product is the length, none for the lengths
the compile-time kernels fix, each no slower than the default order."""
import os

from fastpcc_amd import autotune


def test_table_entries_admissible():
    path = autotune.DEFAULT_TABLE
    assert os.path.exists(path), "no tuning table committed"
    seen = set()
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            n, n1, ms, h1, hms = line.split()
            n, n1, h1 = int(n), int(n1), int(h1)
            assert n not in seen
            seen.add(n)
            for d in (n1, h1):
                assert n % d == 0 and 4 <= d <= 512 and n // d <= 2048
                assert autotune._smooth7(d) and autotune._smooth7(n // d)
            assert 0 < float(ms) <= float(hms)
    assert seen == set(autotune.DEFAULT_SIZES)


def test_radix_table_entries_valid():
    path = autotune.RADIX_TABLE
    assert os.path.exists(path), "no radix table committed"
    seen = set()
    with open(path) as f:
        for line in f:
            body, _, times = line.partition("#")
            parts = body.split()
            if line.startswith("#") or not parts:
                continue
            n, radices = int(parts[0]), [int(v) for v in parts[1:]]
            assert n not in seen and n not in autotune.FIXED_RADIX
            seen.add(n)
            assert radices and all(r in (2, 3, 4, 5, 7, 8, 9, 16, 25) for r in radices) and len(radices) <= 10
            prod = 1
            for r in radices:
                prod *= r
            assert prod == n
            best, default = (float(v) for v in times.split())
            assert 0 < best <= default
    assert seen == set(autotune.radix_sizes())
