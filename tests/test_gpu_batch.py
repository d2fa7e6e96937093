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


@pytest.mark.slow
def test_uneven_200_batch_vs_goldens():
    # the 200 uneven exponents (splitmix64 seed 1234, log-uniform primes in [1e6, 6e7]) in one batch
    # at the batch tuner's splits: two Gerbicz blocks of 500 and a verification per job, res64 at
    # 500 and 1,000 against GMP
    exps = workloads.uneven200()
    gold = [load_golden(f"p{p}_prp") for p in exps]
    missing = [p for p, g in zip(exps, gold) if not g]
    if missing:
        pytest.skip(f"{len(missing)} goldens of the uneven exponents not generated")
    announce("uneven 200-exponent batch", len(exps))
    owl.tune_batch(exps, "", iters=20)
    try:
        with owl.Batch(exps, "") as b:
            groups = b.n_groups
            out, rows = b.prp(1000, block_size=500, check_step=1000, res_step=500)
    finally:
        owl.tune_batch_clear()
    assert groups >= 10
    for i, p in enumerate(exps):
        assert rows[i] == [gold[i][500], gold[i][1000]], f"job {i} (p = {p})"
        assert out[i].nErrors == 0 and out[i].nChecks >= 1 and out[i].maxRoe < 0.4, (p, out[i].maxRoe)


@pytest.mark.slow
def test_batch_full_test_checks_past_the_end():
    # a full PRP test in a batch (iters == E = 859,433, not a multiple of the check step): the
    # batch runs on to the next block boundary past E and is Gerbicz-checked there before x_E
    # counts; M859433 is prime, so x_E = 9 (and a second job of a composite exponent beside it
    # reports its own x_E)
    E, other = 859_433, 859_447
    announce("batch full test", E, other)
    with owl.Batch([E, other], "256:2:256") as b:
        out, _ = b.prp(E, block_size=1000, check_step=10_000)
    assert out[0].k == E and out[0].isPrime and out[0].res64 == 9 and out[0].nErrors == 0
    assert out[1].k == E and not out[1].isPrime and out[1].nErrors == 0
    # the last check came after E: checks at 10,000 ... 850,000 and at 860,000
    assert out[0].nChecks == E // 10_000 + 1
