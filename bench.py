"""Benchmark: PRP-3 iterations per second at the 7.5M-word FFT on one MI355X.

Workload: p = 139,000,003 (smallest prime >= 139,000,000) at the reference's FFT
choice 1K:15:256 (N = 7,864,320 words). A step is one PRP iteration (one row pass
with the fused carry + one column pass with the fused squaring) over the
device-resident residue. W warm-up iterations, then exactly K timed iterations
bracketed by device synchronisation.

Also reported: the per-kernel roofline (HIP events on the engine's queue), the
iteration with Gerbicz updates (timePRP), graph-replayed iterations, the
round-off, a res64 parity check against the committed GMP golden, and the GMP
oracle's single-core rate on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "iter/s at 7.5M-word FFT (bit-exact res64); batched squarings·words/s vs plain loop"
HBM_PEAK_GBS = 8000.0


def golden_res(name):
    path = os.path.join(ROOT, "tests", "golden", f"res_{name}.txt")
    if not os.path.exists(path):
        return {}
    out = {}
    for line in open(path):
        if line.strip() and not line.startswith("#"):
            k, r = line.split()
            out[int(k)] = int(r, 16)
    return out


def cpu_baseline(E, iters):
    """The GMP oracle (one core) on a bounded sample of the same workload."""
    import oracle
    cli = oracle.gmp_cli()
    if not os.path.exists(cli):
        return None
    r = subprocess.run([cli, str(E), str(iters), "0", "0", "1"], capture_output=True, text=True, timeout=600)
    rec = json.loads(r.stdout.strip().splitlines()[-1])
    return {"value": iters / rec["secs"], "unit": "iter/s", "cores": 1, "kind": "port",
            "sample": f"{iters} PRP-3 squarings of a full-width residue mod 2^{E}-1 (after {rec['k0']} untimed "
                      f"squarings from 3) with the GMP oracle (oracle/gmp_loop.c), 1 thread",
            "secs": rec["secs"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=20)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (exponents shard across GPUs "
              "without communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import gpuowl_amd as owl
    from gpuowl_amd import workloads

    E = workloads.config_b()
    g = owl.Gpu(E)
    N = g.N

    # parity: res64 against the GMP golden, before timing
    parity = None
    gold = golden_res("B_prp_short")
    if gold:
        out, res = g.prp(200, block_size=100, check_step=100, res_step=100)
        parity = {"k": [100, 200], "ok": res == [gold[100], gold[200]], "max_roe": out.maxRoe,
                  "gerbicz_checks": out.nChecks, "gerbicz_errors": out.nErrors}

    # timed region: K steady-state iterations, synchronised on both sides
    us = g.time_squares(a.steps, a.warmup, use_graph=False)
    ms_per_step = us / 1000.0
    value = 1e6 / us

    us_graph = g.time_squares(a.steps, a.warmup, use_graph=True)
    us_prp = g.time_prp(iters=max(1000, a.steps // 1000 * 1000), warmup=a.warmup, block_size=1000)
    us_col, us_row = g.time_kernels(200)
    roe = g.read_roe()

    alg_bytes = 16 * N  # per launch: the FP64 state (8 B per word) read once and written once
    kernels = {"col": us_col, "row": us_row}
    dom = max(kernels, key=kernels.get)
    achieved = alg_bytes / (kernels[dom] * 1e-6) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_col (column pass: pairing + square)" if dom == "col" else "k_row (row pass: fused carry)",
                "us_per_launch": {"col": round(us_col, 2), "row": round(us_row, 2)},
                "alg_bytes_per_launch": alg_bytes}
    g.close()

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(E, a.cpu_iters)

    line = {"metric": METRIC, "value": round(value, 2), "unit": "iter/s", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"PRP-3 p={E} at 1K:15:256 (N={N} words)", "exponent": E, "fft_words": N,
                       "parallelism": "single exponent, one GPU"},
            "roofline": roofline, "cpu_baseline": cpu,
            "extra": {"us_per_iter": round(us, 2), "us_per_iter_graph": round(us_graph, 2),
                      "us_per_iter_prp_with_gerbicz": round(us_prp, 2), "roe": roe, "parity": parity}}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
