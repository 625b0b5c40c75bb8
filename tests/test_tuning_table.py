"""The committed FFT split table lists only admissible splits (n = n1 * n2,
n1 <= 512, n2 <= 2048, both 7-smooth), one per length, each no slower than the
heuristic's split when it was measured."""
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
