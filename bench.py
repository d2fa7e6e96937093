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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_batched_baseline(owl, exps, iters, procs):
    """The GMP oracle as a batched CPU baseline: `procs` single-threaded processes at once, each
    squaring a full-width residue of one of the exponents `iters` times; aggregate
    squarings*words/s with each exponent counted at its FFT length N (the GPU's unit)."""
    import oracle
    cli = oracle.gmp_cli()
    if not os.path.exists(cli):
        return None
    chosen = [exps[i % len(exps)] for i in range(procs)]
    t0 = time.time()
    ps = [subprocess.Popen([cli, str(E), str(iters), "0", "0", "1"], stdout=subprocess.PIPE, text=True)
          for E in chosen]
    secs = []
    for p in ps:
        out, _ = p.communicate(timeout=600)
        secs.append(json.loads(out.strip().splitlines()[-1])["secs"])
    wall = time.time() - t0
    words = sum(owl.best_fit(E, "1K:3:256")[1] for E in chosen)
    return {"value": round(words * iters / max(secs)), "unit": "squarings*words/s", "cores": procs,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"{procs} GMP processes at once (oracle/gmp_loop.c, one thread each), {iters} squarings of a "
                      f"full-width residue mod 2^p-1 each, p = the first {len(set(chosen))} primes >= 20,000,000, "
                      f"each counted at N = 1,572,864 words",
            "wall_s": round(wall, 2)}


def engine_key(info):
    return (info["spec"], info["R"], info["long_carry"], info["row"], info["col"])


class PlainLoop:
    """Steady squaring time of one exponent alone (the plain one-exponent-at-a-time loop),
    measured once per engine geometry."""

    def __init__(self, owl, iters=200, warmup=20):
        self.owl, self.iters, self.warmup, self.us = owl, iters, warmup, {}

    def us_per_iter(self, E, spec=""):
        g = self.owl.Gpu(E, spec)
        key = (g.spec, g.R, g.long_carry, g.row_kernel, g.col_kernel)
        if key not in self.us:
            self.us[key] = g.time_squares(self.iters, self.warmup, use_graph=False)
        g.close()
        return self.us[key]


def batch_leg(owl, exps, spec, iters, warmup, tune=True):
    """Batched squarings (graph-replayed) vs the plain loop: the same exponents one at a time.
    With tune, the batch tuner first picks each FFT length's W x H split for the batch (set-up,
    outside the timed region, like a tune file prepared beforehand); the plain loop keeps the
    single-job choices."""
    tuned = None
    if tune:
        t0 = time.time()
        chosen = {}
        for r in owl.tune_batch(exps, spec, iters=20):
            chosen.setdefault(2 * r["W"] * r["H"], r)   # the fastest entry of each length comes first
        tuned = {"seconds": round(time.time() - t0, 1),
                 "splits": {str(n): f'{r["W"]}x{r["H"]} {r["row"]}+{r["col"]}' for n, r in sorted(chosen.items())}}
    b = owl.Batch(exps, spec)
    us_batch = b.time_squares(iters, warmup, use_graph=True)
    words = b.total_words
    shapes = {}
    for i in range(b.n):
        shapes.setdefault(engine_key(b.job_info(i)), []).append(exps[i])
    groups = b.n_groups
    b.close()
    plain = PlainLoop(owl, iters, warmup)
    plain_us = 0.0
    for key, es in shapes.items():
        plain_us += plain.us_per_iter(es[0], key[0]) * len(es)
    rate = words / (us_batch * 1e-6)
    plain = words / (plain_us * 1e-6)
    return {"jobs": len(exps), "groups": groups, "total_words": words, "us_per_batch_iter": round(us_batch, 2),
            "squarings_words_per_s": round(rate), "plain_loop_squarings_words_per_s": round(plain),
            "speedup_vs_plain": round(plain_us / us_batch, 3),
            # the reference's byte model (64 bytes per word and squaring) over the batch, and the
            # 32 bytes per word and squaring this engine's two passes move
            "roofline_frac_64N_model": round(64 * rate / (HBM_PEAK_GBS * 1e9), 4),
            "hbm_frac_32N_model": round(32 * rate / (HBM_PEAK_GBS * 1e9), 4),
            "batch_tuning": tuned}


def pmc_pass(counter, steps=20, warmup=5):
    """One rocprofv3 counter pass over `bench.py --profile` in a child process.
    Returns {(kernel name, grid size): [values per dispatch]}."""
    import csv
    import glob
    import shutil
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    out = tempfile.mkdtemp(prefix="owl_pmc_", dir="/tmp")
    cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "run", "--",
           sys.executable, os.path.join(ROOT, "bench.py"), "--profile", "--steps", str(steps), "--warmup", str(warmup)]
    try:
        subprocess.run(cmd, capture_output=True, text=True, timeout=90, cwd="/tmp",
                       env=dict(os.environ, TMPDIR="/tmp"))
    except subprocess.TimeoutExpired:
        return None
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        return None
    rows = {}
    for r in csv.DictReader(open(files[0])):
        if r["Counter_Name"] == counter:
            rows.setdefault((r["Kernel_Name"], int(r["Grid_Size"])), []).append(float(r["Counter_Value"]))
    shutil.rmtree(out, ignore_errors=True)
    return rows


def pmc_traffic(kernel_prefix, copy_bytes):
    """HBM bytes per launch of the kernel whose name starts with kernel_prefix (its most
    frequent grid), from FETCH_SIZE and WRITE_SIZE, each scaled by what the counter reads
    for the state copy of copy_bytes read + copy_bytes written in the same run."""
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        rows = pmc_pass(counter)
        if not rows:
            return None
        copy = [v for (k, g), vs in rows.items() if "copyBuffer" in k for v in vs if v * 1024 > copy_bytes / 4]
        mine = {(k, g): vs for (k, g), vs in rows.items() if k.split("(")[0].split("<")[0].endswith(kernel_prefix)}
        if not copy or not mine:
            return None
        key = max(mine, key=lambda kg: len(mine[kg]))
        factor = copy_bytes / (sum(copy) / len(copy) * 1024)
        got[counter] = {"kib": round(sum(mine[key]) / len(mine[key]), 1), "factor": round(factor, 4),
                        "kernel": key[0].split("(")[0], "grid": key[1]}
    total = sum(v["kib"] * 1024 * v["factor"] for v in got.values())
    return {"bytes": round(total), "fetch": got["FETCH_SIZE"], "write": got["WRITE_SIZE"]}


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    i = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[i]


def serve_leg(owl, workloads, count, lam, slice_iters, full_below, block=1000, check=50_000):
    """Poisson arrivals (splitmix64 seed 7) of PRP-3 / LL requests, exponents log-uniform primes
    in [2e6, 1.4e8]: below full_below a request runs to completion (E or E - 2 iterations),
    otherwise one slice of slice_iters. Requests are submitted at their arrival time and the
    server steps rounds of `block` iterations; latency = completion - arrival."""
    reqs = workloads.poisson_requests(count, lam)
    length = [(E - 2 if kind == "ll" else E) if E < full_below else slice_iters for _, E, kind in reqs]
    plain = PlainLoop(owl)  # measured before the served run starts
    plain_s = sum(n * plain.us_per_iter(E) * 1e-6 for n, (_, E, _) in zip(length, reqs))
    s = owl.Server(block_size=block, check_step=check)
    t0 = s.now()
    arrive, res, i = {}, [], 0
    while len(res) < count:
        now = s.now() - t0
        while i < count and reqs[i][0] <= now:
            t, E, kind = reqs[i]
            arrive[s.submit(E, kind, length[i])] = t
            i += 1
        inf = s.info()
        if not inf["active"] and not inf["waiting"] and i < count:
            time.sleep(max(0.0, reqs[i][0] - (s.now() - t0)))
            continue
        res += s.step()
    wall = s.now() - t0
    inf = s.info()
    s.close()
    lat = [r["tDone"] - t0 - arrive[r["id"]] for r in res]
    bad = [r for r in res if r["status"] != 0 or r["nErrors"] or r["maxRoe"] >= 0.4]
    return {"requests": count, "lambda_per_s": lam, "slice_iters": slice_iters, "full_below": full_below,
            "block": block, "check_step": check, "wall_s": round(wall, 2), "rounds": inf["rounds"],
            "latency_p50_s": round(percentile(lat, 50), 3), "latency_p99_s": round(percentile(lat, 99), 3),
            "iter_per_s": round(inf["iterations"] / wall), "squarings_words_per_s": round(inf["wordIterations"] / wall),
            "plain_loop_equiv_s": round(plain_s, 2), "speedup_vs_plain": round(plain_s / wall, 3),
            "max_roe": round(max(r["maxRoe"] for r in res), 4), "failed_or_unclean": len(bad)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pmc", action="store_true", help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes")
    ap.add_argument("--cpu-iters", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched legs")
    ap.add_argument("--batch-iters", type=int, default=100)
    ap.add_argument("--no-batch-tune", action="store_true", help="batched legs with the single-job splits")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving sample")
    ap.add_argument("--serve-requests", type=int, default=8,
                    help="the first requests of the serving workload (the default 8 hold one run to completion)")
    ap.add_argument("--serve-lambda", type=float, default=1.0)
    ap.add_argument("--serve-slice", type=int, default=50_000)
    ap.add_argument("--serve-full-below", type=int, default=10_000_000,
                    help="requests below this exponent run to completion, the rest one slice")
    ap.add_argument("--cpu-batch-procs", type=int, default=0,
                    help="GMP processes of the batched CPU baseline (0 = the host's cores, at most 16)")
    ap.add_argument("--profile", action="store_true",
                    help="only the profiled workload: K steady iterations and one state copy, no JSON line")
    ap.add_argument("--profile-batch", action="store_true",
                    help="only the batched workload of 64 primes >= 20M at 1K:3:256: K steady batch iterations, no JSON line")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (exponents shard across GPUs "
              "without communication); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import gpuowl_amd as owl
    from gpuowl_amd import workloads

    def note(msg):
        # progress on stderr as each leg ends; stdout carries only the JSON line
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    if a.profile_batch:
        if not a.no_batch_tune:   # the split the batched leg runs (its candidates appear in a trace too)
            owl.tune_batch(workloads.uniform64_20m(), "1K:3:256", iters=20)
        b = owl.Batch(workloads.uniform64_20m(), "1K:3:256")
        us = b.time_squares(a.steps, a.warmup, use_graph=False)
        print(f"profile batch: {b.n} jobs, {a.warmup}+{a.steps} batch iterations, {us:.1f} us each")
        b.close()
        return

    E = workloads.p139m()
    g = owl.Gpu(E)
    N = g.N

    if a.profile:
        us = g.time_squares(a.steps, a.warmup, use_graph=False)
        us_copy = g.time_copy(1)
        g.close()
        print(f"profile workload: {a.warmup}+{a.steps} iterations, {us:.1f} us/iter; copy {us_copy:.1f} us")
        return

    # parity: res64 against the GMP golden, before timing
    parity = None
    gold = golden_res("p139000003_prp_short")
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
    us_copy = g.time_copy(50)
    roe = g.read_roe()
    note(f"{a.steps} timed iterations: {us:.2f} us/iter (row {us_row:.2f} us, column {us_col:.2f} us)")

    # Algorithmic bytes: 64*N per iteration in the reference's model, 16*N (state read once,
    # written once) for each of its four passes. Per launch, each of the two passes here must
    # read and write the state once: 16*N. (The column pass does the work of three reference
    # passes, middle-in, tail square and middle-out, but crediting it with their 48*N would
    # count HBM bytes it never moves and put it above the HBM peak.)
    alg = {"row": 16 * N, "col": 16 * N}
    kernels = {"col": us_col, "row": us_row}
    names = {"row": g.row_kernel, "col": g.col_kernel}
    dom = max(kernels, key=kernels.get)
    achieved = alg[dom] / (kernels[dom] * 1e-6) / 1e9
    copy_bytes = 8 * N  # the state copy reads 8*N bytes and writes 8*N bytes
    traffic = None
    if not a.no_pmc:
        prefix = "k_rowx" if names[dom].startswith("rowx") else "k_colx" if names[dom].startswith("colx") else \
            "k_row" if dom == "row" else "k_col"
        traffic = pmc_traffic(prefix, copy_bytes)
        note(f"counter passes done: {traffic['bytes'] if traffic else None} HBM bytes per launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic["bytes"] if traffic else None,
                "kernel": f"{names[dom]} ({'row pass: fused carry' if dom == 'row' else 'column pass: middle + square'})",
                "alg_bytes_per_launch": alg[dom],
                "us_per_launch": {"col": round(us_col, 2), "row": round(us_row, 2)},
                "traffic_detail": traffic,
                "iteration": {"alg_bytes": 64 * N, "us": round(us_col + us_row, 2),
                              "frac": round(64 * N / ((us_col + us_row) * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)},
                "copy_gbs": round(2 * copy_bytes / (us_copy * 1e-6) / 1e9, 1)}
    g.close()

    def leg(fn, *args):
        # an extra leg that fails is reported, never allowed to drop the bench line
        try:
            return fn(*args)
        except Exception as e:  # noqa: BLE001
            return {"error": f"{type(e).__name__}: {e}"}

    batched = None
    if not a.no_batch:
        batched = {}
        batched["uniform_64x20M_at_1K:3:256"] = leg(batch_leg, owl, workloads.uniform64_20m(), "1K:3:256",
                                                    a.batch_iters, 10, not a.no_batch_tune)
        note("batch of 64 exponents at 1K:3:256 done")
        batched["uneven_200_exponents_1M_to_60M"] = leg(batch_leg, owl, workloads.uneven200(), "", a.batch_iters, 10,
                                                        not a.no_batch_tune)
        note("batch of 200 uneven exponents done")

    served = None
    if not a.no_serve:
        served = leg(serve_leg, owl, workloads, a.serve_requests, a.serve_lambda, a.serve_slice,
                     a.serve_full_below)
        note("serving sample done")

    cpu = cpu_batch = None
    if not a.no_cpu:
        cpu = cpu_baseline(E, a.cpu_iters)
        if cpu:
            cpu["cpu_model"] = cpu_model()
        note("CPU baseline done")
        procs = a.cpu_batch_procs or min(os.cpu_count() or 1, 16)
        cpu_batch = leg(cpu_batched_baseline, owl, workloads.uniform64_20m(), 64, procs)
        note("batched CPU baseline done")

    line = {"metric": METRIC, "value": round(value, 2), "unit": "iter/s", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": round(ms_per_step, 5), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"PRP-3 p={E} at 1K:15:256 (N={N} words)", "exponent": E, "fft_words": N,
                       "parallelism": "single exponent, one GPU"},
            "roofline": roofline, "cpu_baseline": cpu,
            "extra": {"us_per_iter": round(us, 2), "us_per_iter_graph": round(us_graph, 2),
                      "us_per_iter_prp_with_gerbicz": round(us_prp, 2), "roe": roe, "parity": parity,
                      "batched": batched, "served": served, "cpu_baseline_batched": cpu_batch,
                      "reference_opencl": "not measured: no OpenCL CPU runtime (PoCL) on the box and the "
                                          "reference is not shipped there"}}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
