"""Host logic of the request server that needs no GPU: the device-arena block
allocator (fastpcc_amd/csrc/arena.hpp), compiled with g++ and checked against a
byte-map model on randomised alloc/release sequences."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_allocator(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "arena_check")
    src = os.path.join(ROOT, "tests", "host", "arena_check.cpp")
    inc = os.path.join(ROOT, "fastpcc_amd", "csrc")
    subprocess.run([gxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-I", inc, src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr


def test_arena_allocator_asan(tmp_path):
    """The same check under AddressSanitizer and UBSan (host code only)."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "arena_check_asan")
    src = os.path.join(ROOT, "tests", "host", "arena_check.cpp")
    inc = os.path.join(ROOT, "fastpcc_amd", "csrc")
    r = subprocess.run([gxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover",
                        "-I", inc, src, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer runtime not available: " + r.stderr[-200:])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout + out.stderr
