"""The C correlation driver (pcc_fullpair) end to end on the reference's example
SAC files: read, pair by start time, preprocess and correlate on the GPU, write
ts-PWS binary or SAC files; the written correlations are compared with the
oracle run on the same samples (read here with numpy, paired by date)."""
import glob
import os
import subprocess

import numpy as np
import pytest

from fastpcc_amd import formats as fio

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
SAC = os.path.join(HERE, "golden", "sac")
CAN = sorted(glob.glob(os.path.join(SAC, "*.G.CAN.*")))   # days 002-004
ECH = sorted(glob.glob(os.path.join(SAC, "*.G.ECH.*")))   # days 001-003
TOL = 1e-5


def _samples(paths):
    return np.stack([np.frombuffer(open(p, "rb").read()[632:], "<f4") for p in paths]).astype(np.float32)


def _paired():
    """Days 002 and 003: CAN files 0, 1 and ECH files 1, 2."""
    return _samples(CAN[:2]), _samples(ECH[1:3])


def _run(tmp_path, args):
    (tmp_path / "can.txt").write_text("".join(p + "\n" for p in CAN))
    (tmp_path / "ech.txt").write_text("".join(p + "\n" for p in ECH))
    cmd = [os.path.join(fio.binary_dir(), "pcc_fullpair"), *args]
    print("run:", " ".join(os.path.basename(a) for a in cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr


def _read_bin(path):
    raw = open(path, "rb").read()
    L, nseq = np.frombuffer(raw[96:104], "<u4")
    data = np.frombuffer(raw[116 + 12 * nseq:], "<f4").reshape(nseq, L)
    return raw[:8].rstrip(b"\0").decode(), np.frombuffer(raw[116:116 + 8 * nseq], "<i8"), data


def test_cli_obin_pcc2_gncc_1bit(tmp_path, orc):
    _run(tmp_path, ["can.txt", "ech.txt", "pcc", "ccgn", "cc1b", "tl1=-12000", "tl2=12000", "obin"])
    x1, x2 = _paired()
    for name, kind in (("pcc2", "pcc2"), ("ccgn", "gncc"), ("cc1b", "cc1b")):
        method, times, y = _read_bin(tmp_path / f"CAN.00Z.ECH.00Z_{name}.bin")
        assert method == name and y.shape == (2, 6001)
        assert list(times) == [fio.read_sac(p)[0].t for p in CAN[:2]]
        err = np.abs(y - orc.run(kind, x1, x2, -3000, 3000)).max()
        print(f"{name}: max|cli-oracle| = {err:.2e}")
        assert err <= TOL


def test_cli_msacs_osac_pcc1(tmp_path, orc):
    b = fio.binary_dir()
    for lst, out in (("can.txt", "can.msacs"), ("ech.txt", "ech.msacs")):
        (tmp_path / "can.txt").write_text("".join(p + "\n" for p in CAN))
        (tmp_path / "ech.txt").write_text("".join(p + "\n" for p in ECH))
        subprocess.run([os.path.join(b, "filelist2msacs"), lst, out], cwd=tmp_path, check=True, capture_output=True)
    _run(tmp_path, ["can.msacs", "ech.msacs", "imsacs", "pcc", "v=1", "tl1=-2000", "tl2=2000", "osac"])
    x1, x2 = _paired()
    yr = orc.run("pcc1", x1, x2, -500, 500)
    for k, day in enumerate(("002", "003")):
        p = tmp_path / f"G.CAN.00.LHZ.G.ECH.00.LHZ_pcc1_2017.{day}.00.00.00.sac"
        raw = open(p, "rb").read()
        y = np.frombuffer(raw[632:], "<f4")
        assert np.frombuffer(raw[:4], "<f4")[0] == 4.0 and np.frombuffer(raw[20:24], "<f4")[0] == -2000.0
        err = np.abs(y - yr[k]).max()
        print(f"pcc1 day {day}: max|cli-oracle| = {err:.2e}")
        assert err <= TOL


def test_cli_preprocessing(tmp_path, orc):
    _run(tmp_path, ["can.txt", "ech.txt", "pcc", "clip", "std=3", "awhite=0.005,0.05", "tl1=-12000", "tl2=12000",
                    "obin"])
    N = 21600
    c, e = _samples(CAN), _samples(ECH)
    # per station over its own traces: zero traces, rms, clipping, outliers
    cp, cv, _ = orc.preprocess(c.reshape(1, -1), 3, N, clip=True, std=3.0)
    ep, ev, _ = orc.preprocess(e.reshape(1, -1), 3, N, clip=True, std=3.0)
    # pairing by day among the kept traces (CAN days 2-4, ECH days 1-3)
    pairs = [(d - 2, d - 1) for d in (2, 3) if cv[0, d - 2] and ev[0, d - 1]]
    assert pairs, "no pair left"
    i1, i2 = [a for a, _ in pairs], [b for _, b in pairs]
    n = len(pairs)
    # then each station's paired traces are whitened
    x1, _, _ = orc.preprocess(cp.reshape(3, N)[i1].reshape(1, -1), n, N, white=(0.005, 0.05), dt=4.0)
    x2, _, _ = orc.preprocess(ep.reshape(3, N)[i2].reshape(1, -1), n, N, white=(0.005, 0.05), dt=4.0)
    _, times, y = _read_bin(tmp_path / "CAN.00Z.ECH.00Z_pcc2.bin")
    assert y.shape == (n, 6001)
    assert list(times) == [fio.read_sac(CAN[a])[0].t for a in i1]
    err = np.abs(y - orc.run("pcc2", x1.reshape(n, N), x2.reshape(n, N), -3000, 3000)).max()
    print(f"preprocessed pcc2 ({n} pairs): max|cli-oracle| = {err:.2e}")
    assert err <= TOL
