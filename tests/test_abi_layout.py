"""The ctypes mirrors of the public structs have the C layout of include/fastpcc.h:
a small C program compiled with gcc against the header prints every field's
offset and each struct's size, which must equal ctypes' (a mismatched field
order once passed the output pointer where the station keys belonged)."""
import ctypes
import os
import subprocess

import pytest

from fastpcc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "fpcc_job": _lib.FpccJob,
    "fpcc_preproc": _lib.Preproc,
    "fpcc_server_config": _lib.ServerConfig,
    "fpcc_request": _lib.Request,
    "fpcc_request_status": _lib.RequestStatus,
    "fpcc_serving_stats": _lib.ServingStats,
}


def _c_layout(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fastpcc.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {tuple(line.split()[:2]): int(line.split()[2]) for line in out.splitlines()}


@pytest.mark.parametrize("cname", sorted(STRUCTS))
def test_struct_layout_matches_header(tmp_path, cname):
    c = _c_layout(tmp_path)
    cls = STRUCTS[cname]
    assert c[(cname, "size")] == ctypes.sizeof(cls)
    for fname, _ in cls._fields_:
        assert c[(cname, fname)] == getattr(cls, fname).offset, fname
