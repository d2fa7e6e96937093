"""Batches: every job's res64 stream equals GMP's and equals the same exponent run
alone; uneven batches group jobs by FFT geometry."""
import os
import sys

import pytest

import gpuowl_amd as owl
import oracle
from gpuowl_amd import workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def announce(*a):
    print("CASE", *a, file=sys.stderr, flush=True)


def test_uniform_batch_matches_gmp_and_single_runs():
    spec = "256:2:256"
    exps = workloads.primes_from(3_900_000, 4)
    announce("uniform batch", spec, exps)
    iters, step = 300, 100
    with owl.Batch(exps, spec) as b:
        assert b.n_groups == 1
        out, rows = b.prp(iters, block_size=100, check_step=100, res_step=step)
    for E, r, o in zip(exps, rows, out):
        assert r == oracle.gmp_res64(E, iters, step)
        assert o.nErrors == 0 and o.nChecks == 3 and o.maxRoe < 0.4
    with owl.Gpu(exps[2], spec) as g:
        _, alone = g.prp(iters, block_size=100, check_step=100, res_step=step)
    assert alone == rows[2]


def test_uneven_batch_groups_by_shape():
    # reference FFT choices 256K, 384K and 512K words; one exponent at low bits/word
    exps = [1_029_689, 2_287_861, 3_503_053, 5_300_023, 1_500_007]
    announce("uneven batch", exps)
    with owl.Batch(exps) as b:
        Ns = [b.job_info(i)["N"] for i in range(b.n)]
        assert Ns == [owl.best_fit(E)[1] for E in exps]
        # a group shares kernels and launches: same geometry, line block and carry mode
        keys = set()
        for i in range(b.n):
            j = b.job_info(i)
            keys.add((j["spec"], j["R"], j["long_carry"], j["row"], j["col"]))
        assert b.n_groups == len(keys)
        out, rows = b.prp(200, block_size=100, check_step=100, res_step=50)
    for E, r, o in zip(exps, rows, out):
        assert r == oracle.gmp_res64(E, 200, 50)
        assert o.nErrors == 0


def test_batch_recovers_from_fault_in_one_job():
    exps = workloads.primes_from(3_700_000, 3)
    announce("batch gerbicz recovery", exps)
    with owl.Batch(exps, "256:2:256") as b:
        b.arm_fault(1, 230)
        out, rows = b.prp(400, block_size=50, check_step=100, res_step=100)
    for E, r, o in zip(exps, rows, out):
        assert r == oracle.gmp_res64(E, 400, 100)
    assert out[1].nErrors == 1 and out[0].nErrors == 0


def test_batch_ll():
    exps = workloads.primes_from(2_000_000, 3)
    with owl.Batch(exps, "256:2:256") as b:
        out, rows = b.ll(120, res_step=40)
    for E, r in zip(exps, rows):
        assert r == oracle.gmp_res64(E, 120, 40, ll=True)


def load_golden(name):
    path = os.path.join(ROOT, "tests", "golden", f"res_{name}.txt")
    if not os.path.exists(path):
        return None
    res = {}
    for line in open(path):
        if line.strip() and not line.startswith("#"):
            k, r = line.split()
            res[int(k)] = int(r, 16)
    return res


def test_uniform_batch_64x20m_vs_goldens():
    # the first 64 primes >= 20,000,000, all at 1K:3:256, advanced together
    exps = workloads.uniform64_20m()
    gold = [load_golden(f"p{p}_prp") for p in exps]
    have = [i for i, g in enumerate(gold) if g]
    if not have:
        pytest.skip("goldens of the 64 primes >= 20M not generated")
    announce("64-prime batch at 1K:3:256", len(exps))
    with owl.Batch(exps, "1K:3:256") as b:
        assert b.n_groups == 1 and b.job_info(0)["N"] == 1_572_864
        out, rows = b.prp(1000, block_size=500, check_step=500, res_step=500)
    for i in have:
        assert rows[i] == [gold[i][500], gold[i][1000]], f"job {i} (p = {exps[i]})"
        assert out[i].nErrors == 0 and out[i].maxRoe < 0.4
