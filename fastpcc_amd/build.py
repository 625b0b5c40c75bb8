"""Builds libfastpcc.so (HIP, gfx950) in-tree with hipcc, and the C drivers.

Every HIP source under csrc/ is compiled for gfx950 with packed-FP32 instruction
selection disabled, every C++ host source under csrc/ with g++, and all are linked
into fastpcc_amd/libfastpcc.so.  The C drivers under cli/ (the correlation
driver and the SAC-list-to-MSACS1 converter) are built with gcc into
fastpcc_amd/bin/, linked against the library.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfastpcc.so")
CLI = os.path.join(HERE, "cli")
BIN = os.path.join(HERE, "bin")
CXX = os.environ.get("CXX", "g++")
CC = os.environ.get("CC", "gcc")
CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-I", os.path.join(HERE, "..", "include")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FASTPCC_ARCH", "gfx950")

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(HERE, "..", "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    base, ext = os.path.splitext(os.path.basename(src))
    obj = os.path.join(BUILD, base + (".o" if ext == ".hip" else ".host.o"))
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj] if ext == ".hip" else [CXX, *CXXFLAGS, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{p.stdout}\n{p.stderr}")
    return obj, p.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, LIB)
    build_cli()
    return LIB


def build_cli():
    """gcc builds of the C drivers, linked against libfastpcc.so (rpath: ../)."""
    os.makedirs(BIN, exist_ok=True)
    for src in sorted(f for f in os.listdir(CLI) if f.endswith(".c")):
        exe = os.path.join(BIN, src[:-2])
        cmd = [CC, "-O2", "-std=gnu11", "-Wall", "-Wextra", "-I", os.path.join(HERE, "..", "include"),
               os.path.join(CLI, src), "-o", exe + ".tmp", "-L", HERE, "-lfastpcc", "-Wl,-rpath,$ORIGIN/..", "-lm"]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(f"gcc failed on {src}:\n{p.stdout}\n{p.stderr}")
        os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    print(build(verbose=True))
