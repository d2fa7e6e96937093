"""The fused row pass's chunked block order, checked on the CPU from the kernel header itself:
tests/row_order_check.cpp includes gpuowl_amd/csrc/kfixed.h and runs chunkedItem (the mapping
the device code uses) for every block count from 8 to 4200. It must store and publish every
block exactly once and make every group wait only on an earlier group of its own XCD."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_chunked_row_order_covers_blocks_and_waits_within_an_xcd(tmp_path):
    exe = tmp_path / "row_order_check"
    csrc = os.path.join(ROOT, "gpuowl_amd", "csrc")
    subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O1", "-I", csrc,
                    "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "row_order_check.cpp"),
                    "-o", str(exe)], check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], check=False, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout + out.stderr
