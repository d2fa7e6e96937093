"""The ctypes mirrors of the C-ABI structs (gpuowl_amd/_lib.py) have the layout of the
structs in include/owl.h: a small C program compiled here against the header prints each
struct's size and every field's offset, which must equal what ctypes computes."""
import os
import shutil
import subprocess

import pytest

from gpuowl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {"owl_gpu_info_t": _lib.GpuInfo, "owl_prp_result_t": _lib.PrpResult, "owl_ll_result_t": _lib.LLResult,
           "owl_serve_result_t": _lib.ServeResult, "owl_serve_info_t": _lib.ServeInfo,
           "owl_tune_entry_t": _lib.TuneEntry, "owl_result_t": _lib.ResultRec}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_ctypes_structs_match_the_header(tmp_path):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "owl.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        name, what, value = line.split()
        got[(name, what)] = int(value)
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == _lib.C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, (cname, field)
