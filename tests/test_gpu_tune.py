"""Autotuner: every configuration it times for a 256K-word FFT gives GMP's residues, and
a recorded choice is what a later engine of that FFT length runs."""
import sys

import pytest

import gpuowl_amd as owl
import oracle
from gpuowl_amd import workloads

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, file=sys.stderr, flush=True)


def test_tuned_configurations_are_exact():
    E = workloads.primes_from(3_900_000, 1)[0]
    announce("tune", E)
    res = owl.tune(E, "256:2:256", iters=100)
    assert len(res) >= 3
    assert all(a["us"] <= b["us"] for a, b in zip(res, res[1:]))
    want = oracle.gmp_res64(E, 200, 100)
    N = owl.best_fit(E, "256:2:256")[1]
    try:
        for c in res:
            announce("tuned config", c)
            owl.tune_set(N, c["width"], c["line_block"], c["flags"], c["us"])
            with owl.Gpu(E, "256:2:256") as g:
                assert (g.W, g.R, g.row_kernel, g.col_kernel) == (c["W"], c["R"], c["row"], c["col"])
                out, got = g.prp(200, block_size=100, check_step=100, res_step=100)
            assert got == want, c
    finally:
        owl.tune_clear()


def test_batch_tuning_picks_a_split_for_batches_only():
    exps = workloads.primes_from(3_900_000, 8)
    N = owl.best_fit(exps[0], "256:2:256")[1]
    announce("tune batch", exps[0], len(exps))
    try:
        res = owl.tune_batch(exps, "256:2:256", iters=10)
        assert len(res) >= 2 and {2 * r["W"] * r["H"] for r in res} == {N}
        assert all(a["us"] <= b["us"] for a, b in zip(res, res[1:]))
        best = res[0]
        with owl.Batch(exps, "256:2:256") as b:
            info = b.job_info(0)
            assert (info["row"], info["col"]) == (best["row"], best["col"])
            outs, _ = b.prp(200, block_size=100, check_step=100)
        for E, o in zip(exps[:3], outs[:3]):
            assert o.res64 == oracle.gmp_res64(E, 200, 200)[0], E
        with owl.Gpu(exps[0], "256:2:256") as g:   # single jobs keep the reference's split
            assert g.W == 256
    finally:
        owl.tune_batch_clear()
