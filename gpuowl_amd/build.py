"""Builds gpuowl_amd/libgpuowl_amd.so (HIP kernels for gfx950 + C-ABI + host engine)
with hipcc, in-tree. Usage: python gpuowl_amd/build.py [-j N]"""
import argparse, os, subprocess, sys, concurrent.futures as cf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpuowl_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
SOURCES = ["kernels.hip", "kfix_row_a.hip", "kfix_row_b.hip", "kfix_row_c.hip", "kfix_colx_a.hip", "kfix_colx_b.hip", "kfix_colx_c.hip", "kfix_coly_a.hip", "kfix_coly_b.hip", "hipwrap.cpp", "FFTConfig.cpp", "Trig.cpp", "state.cpp", "Saver.cpp", "Worktodo.cpp", "Proof.cpp", "TrigBufCache.cpp", "Gpu.cpp", "GpuBatch.cpp", "Server.cpp", "Tune.cpp", "capi.cpp"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
# Packed FP32 instructions stay off (no FP32 math on this path anyway).
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-ffp-contract=fast"]


def compile_one(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + (DEVICE if src.endswith(".hip") else []) + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, r.returncode, r.stdout + r.stderr, out


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, log, out in ex.map(compile_one, SOURCES):
            if verbose or rc:
                sys.stderr.write(log)
            if rc:
                raise RuntimeError(f"hipcc failed on {src}")
            objs.append(out)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
