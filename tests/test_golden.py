"""The oracle against the committed config-1 golden fixtures (tests/golden,
written by make_golden.py with an FP64 numpy route independent of the oracle)."""
import os

import numpy as np
import pytest

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


@pytest.mark.parametrize("kind", ["pcc2", "gncc", "cc1b"])
def test_oracle_matches_golden(orc, kind):
    x = _load("c1_x.npy")
    gold = _load(f"c1_{kind}_fp64.npy")
    y64 = orc.allpairs(kind, x, 1, 3600, -500, 500, stacked=True, f64=True)
    y32 = orc.allpairs(kind, x, 1, 3600, -500, 500, stacked=True)
    assert np.abs(y64 - gold).max() < 1e-12
    assert np.abs(y32 - gold).max() < 1e-6


def test_golden_peaks_at_true_delays():
    x = _load("c1_x.npy")
    sh = _load("c1_shifts.npy")
    gold = _load("c1_pcc2_fp64.npy")
    S = x.shape[0]
    p = 0
    for i in range(S):
        for j in range(i + 1, S):
            assert np.argmax(gold[p]) - 500 == sh[j] - sh[i]
            p += 1
