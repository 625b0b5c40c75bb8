"""FFT autotuners: the split tuner returns an admissible split no slower than the
heuristic's and later plans use it; the radix tuner returns an order of the
length no slower than the default; correlations computed with forced extreme
splits and unusual radix orders (read from tuning tables) match the oracle."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from fastpcc_amd import autotune

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-5


def _admissible(n, n1):
    return 4 <= n1 <= 512 and n % n1 == 0 and n // n1 <= 2048 and autotune._smooth7(n1) and \
        autotune._smooth7(n // n1)


def test_tuner_picks_admissible_split_and_uses_it(fp):
    n = 98304  # 3 * 2^15: no other test builds a plan of this length
    n1, ms, h1, hms = autotune.tune(n, reps=3, batch=64)
    print(f"n={n} best n1={n1} {ms:.4f} ms heuristic n1={h1} {hms:.4f} ms", flush=True)
    assert _admissible(n, n1) and _admissible(n, h1)
    assert 0 < ms <= hms
    assert fp._lib.lib().fpcc_fft_split(n) == n1
    assert fp._lib.lib().fpcc_autotune_fft(12289, 8, 1, None, None, None, None) == -1  # prime: no split


_CHILD = r"""
import sys
import numpy as np
import fastpcc_amd as fp
x1, x2 = np.load(sys.argv[1]), np.load(sys.argv[2])
out = dict(pcc2=fp.pcc2_set(x1, x2, -3000, 3000), gncc=fp.ccgn_set(x1, x2, -3000, 3000),
           cc1b=fp.cc1b_set(x1, x2, -3000, 3000), split=np.array(fp._lib.lib().fpcc_fft_split(32768)))
np.savez(sys.argv[3], **out)
"""


@pytest.mark.parametrize("n1", [16, 512])
def test_forced_split_matches_oracle(orc, tmp_path, n1):
    """N = 21600, lags +-3000 give Nz = 32768; the table forces 16 x 2048 or
    512 x 64 (the two ends of the admissible range)."""
    rng = np.random.default_rng(n1)
    x1 = rng.standard_normal((2, 21600)).astype(np.float32)
    x2 = rng.standard_normal((2, 21600)).astype(np.float32)
    np.save(tmp_path / "x1.npy", x1)
    np.save(tmp_path / "x2.npy", x2)
    table = tmp_path / "splits.txt"
    table.write_text(f"# forced\n32768 {n1}\n")
    env = dict(os.environ, FASTPCC_FFT_TUNING=str(table), PYTHONPATH=ROOT)
    print(f"child run: Nz=32768 split {n1} x {32768 // n1}", flush=True)
    r = subprocess.run([sys.executable, "-c", _CHILD, str(tmp_path / "x1.npy"), str(tmp_path / "x2.npy"),
                        str(tmp_path / "y.npz")], env=env, cwd=ROOT, timeout=300, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    y = np.load(tmp_path / "y.npz")
    assert int(y["split"]) == n1
    for kind in ("pcc2", "gncc", "cc1b"):
        err = np.abs(y[kind] - orc.run(kind, x1, x2, -3000, 3000)).max()
        print(f"  {kind}: max|gpu-oracle|={err:.3e}")
        assert err <= TOL


def test_radix_tuner_picks_valid_order(fp):
    n = 1400  # 2^3 5^2 7: no other test builds a plan of this length
    r, ms, dms = autotune.tune_radix(n, reps=3, batch=512)
    print(f"n={n} best radices {r} {ms:.4f} ms default {dms:.4f} ms", flush=True)
    assert int(np.prod(r)) == n and 0 < ms <= dms
    out = (ctypes.c_int * 16)()
    k = fp._lib.lib().fpcc_fft_radices(n, out)
    assert list(out)[:k] == r
    assert fp._lib.lib().fpcc_autotune_radix(4096, 8, 1, None, None, None, None) == -1  # fixed length


_CHILD_RADIX = r"""
import sys, ctypes
import numpy as np
import fastpcc_amd as fp
x = np.load(sys.argv[1])
out = dict(pcc2=fp.allpairs(x, 2, 3000, -800, 800, kind="pcc2", stack="none"),
           gncc=fp.allpairs(x, 2, 3000, -800, 800, kind="gncc", stack="plain"),
           gram=fp.allpairs(x, 2, 3000, -800, 800, kind="pcc2", stack="gram"))
r = (ctypes.c_int * 16)()
k = fp._lib.lib().fpcc_fft_radices(3000, r)
out["radices3000"] = np.array(list(r)[:k])
np.savez(sys.argv[2], **out)
"""


def test_forced_radix_orders_match_oracle(orc, tmp_path):
    """Orders forced through a table: the window transform of N = 3000 as
    3 x 5 x 8 x 25 (smallest first) and the half-length transforms of
    Nz = 4096 (GNCC, the Gram stack) as 2 x 4 x 16 x 16: the same correlations
    as the oracle."""
    rng = np.random.default_rng(4)
    x = rng.standard_normal((4, 6000)).astype(np.float32)
    np.save(tmp_path / "x.npy", x)
    table = tmp_path / "radix.txt"
    table.write_text("3000 3 5 8 25\n2048 2 4 16 16\n")
    env = dict(os.environ, FASTPCC_RADIX_TUNING=str(table), PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-c", _CHILD_RADIX, str(tmp_path / "x.npy"), str(tmp_path / "y.npz")],
                       env=env, cwd=ROOT, timeout=300, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    y = np.load(tmp_path / "y.npz")
    assert list(y["radices3000"]) == [3, 5, 8, 25]
    for name, kind, stacked in (("pcc2", "pcc2", False), ("gncc", "gncc", True), ("gram", "pcc2", True)):
        yr = orc.allpairs(kind, x, 2, 3000, -800, 800, stacked=stacked)
        err = np.abs(y[name].reshape(yr.shape) - yr).max()
        print(f"  {name}: max|gpu-oracle|={err:.3e}")
        assert err <= TOL
