"""Serving: requests admitted while others run (continuous batching) end on the same
res64 as the same request run alone and as GMP; admission respects the memory budget."""
import json
import sys

import pytest

import gpuowl_amd as owl
import oracle
from gpuowl_amd import workloads

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, file=sys.stderr, flush=True)


def expected(E, kind, iters):
    return oracle.gmp_res64(E, iters, iters, ll=(kind == "ll"))[0]


def alone(E, kind, iters):
    with owl.Gpu(E) as g:
        if kind == "ll":
            out, _ = g.ll(iters)
        else:
            out, _ = g.prp(iters, block_size=200, check_step=400)
        return out.res64


def test_mix_matches_alone_and_gmp():
    # two FFT geometries (256K and 384K words), PRP and LL, lengths off the round size,
    # requests arriving while the first ones run
    first = [(2_000_003, "prp", 1234), (2_000_029, "ll", 1000), (5_300_023, "prp", 801)]
    later = [(2_000_081, "prp", 1399), (5_300_039, "ll", 617), (3_900_067, "prp", 400)]
    announce("serve mix", first + later)
    got = {}
    with owl.Server(block_size=200, check_step=400) as s:
        ids = {s.submit(E, k, n): (E, k, n) for E, k, n in first}
        for r in s.step() + s.step():
            got[r["id"]] = r
        assert s.info()["active"] == 3
        ids.update({s.submit(E, k, n): (E, k, n) for E, k, n in later})
        for r in s.drain():
            got[r["id"]] = r
        info = s.info()
    assert set(got) == set(ids)
    assert info["iterations"] == sum(n for _, _, n in first + later)
    for rid, (E, k, n) in ids.items():
        r = got[rid]
        assert r["status"] == 0 and r["nErrors"] == 0 and r["maxRoe"] < 0.4, r
        assert r["res64"] == expected(E, k, n), (E, k, n)
        assert r["res64"] == alone(E, k, n), (E, k, n)
        if k == "prp":
            assert r["nChecks"] >= 1
        assert r["tSubmit"] <= r["tAdmit"] <= r["tDone"]


def test_rounds_balance_small_and_large_ffts():
    # a 1.5M-word job beside 256K-word ones: the small group runs six blocks per round,
    # with its requested iterations and check points falling inside the round
    big = workloads.next_prime(12 * 1_572_864)
    reqs = [(big, "prp", 400), (3_900_067, "prp", 1234), (3_900_091, "ll", 777)]
    announce("serve balance", reqs)
    with owl.Server(block_size=100, check_step=200) as s:
        ids = {s.submit(E, k, n): (E, k, n) for E, k, n in reqs}
        res = {r["id"]: r for r in s.drain()}
    assert set(res) == set(ids)
    for rid, (E, k, n) in ids.items():
        r = res[rid]
        assert r["status"] == 0 and r["nErrors"] == 0, r
        assert r["res64"] == expected(E, k, n), (E, k, n)


def test_admission_by_memory_budget():
    exps = workloads.primes_from(3_000_000, 4)
    spec = "256:2:256"
    N = owl.best_fit(exps[0], spec)[1]
    one = 60 * N  # a job's footprint bound is about 52 bytes per word
    announce("serve budget", exps)
    with owl.Server(mem_budget=2 * one, block_size=100, check_step=200, spec=spec) as s:
        for E in exps:
            s.submit(E, "prp", 300)
        s.step()
        i = s.info()
        assert i["active"] == 2 and i["waiting"] == 2 and i["bytesInUse"] <= i["budget"]
        res = s.drain()
    assert sorted(r["E"] for r in res) == exps
    for r in res:
        assert r["res64"] == expected(r["E"], "prp", 300)


@pytest.mark.slow
def test_full_mersenne_prime_runs_to_completion():
    # M859433 is prime: PRP ends on 9 at k = E and LL on 0 at k = E - 2 (long-carry geometry).
    # The requests come from worktodo lines and the results go out as results lines: the PRP
    # one reports the type-1 residue x_E / 9 = 1.
    E = 859_433
    aid = "0123456789ABCDEF0123456789ABCDEF"
    lines = [f"PRP={aid},1,2,{E},-1,75,0", f"DoubleCheck={E},75,1"]
    announce("serve full", E)
    with owl.Server(block_size=1000, check_step=50_000) as s:
        for line in lines:
            t = owl.worktodo_parse(line)
            s.submit(t["exponent"], t["kind"])
        res = s.drain()
    assert len(res) == 2
    for r in res:
        assert r["status"] == 0 and r["isPrime"], r
    prp = next(r for r in res if not r["ll"])
    assert prp["res64Type1"] == 1 and prp["res2048"] == "0" * 511 + "1"
    n = owl.best_fit(E)[1]
    line = json.loads(owl.result_json("prp", E, prp["isPrime"], prp["res64Type1"], n, res2048=prp["res2048"],
                                      n_errors=prp["nErrors"], aid=aid))
    assert (line["status"], line["exponent"], line["worktype"], line["res64"], line["aid"]) == \
        ("P", E, "PRP-3", "0000000000000001", aid)
    ll = next(r for r in res if r["ll"])
    assert ll["res64Type1"] == 0 and ll["res2048"] == ""
    assert json.loads(owl.result_json("ll", E, ll["isPrime"], ll["res64"], n))["status"] == "P"


def test_fault_in_one_request_is_caught_and_recovered():
    # one data bit of the first PRP request flips after the round that passes iteration 500;
    # its Gerbicz check at 800 fails, it reloads the state verified at 400 and ends on the
    # right residue; the other member of its group and the LL request are unaffected
    reqs = [(2_000_003, "prp", 1600), (2_000_029, "prp", 1600), (2_000_081, "ll", 900)]
    announce("serve fault", reqs)
    with owl.Server(block_size=200, check_step=400) as s:
        ids = {s.submit(E, k, n): (E, k, n) for E, k, n in reqs}
        faulty = min(ids)
        s.arm_fault(faulty, 500)
        res = {r["id"]: r for r in s.drain()}
    assert set(res) == set(ids)
    for rid, (E, k, n) in ids.items():
        r = res[rid]
        assert r["status"] == 0, r
        assert r["nErrors"] == (1 if rid == faulty else 0), r
        assert r["res64"] == expected(E, k, n), (E, k, n)


def load_golden(name):
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"res_{name}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip() and not line.startswith("#"):
            k, r = line.split()
            last = (int(k), int(r, 16))
    return last


@pytest.mark.slow
def test_serving_policy_prefix_matches_alone():
    # The first 8 requests of the serving workload (Poisson arrivals, splitmix64 seed 7, exponents
    # log-uniform primes in [2e6, 1.4e8]) under its policy: below 10,000,000 a request runs to
    # completion (E or E - 2 iterations), otherwise one 50,000-iteration slice; Gerbicz blocks of
    # 1,000 checked every 50,000 and past the end. Requests are submitted as the server's clock
    # passes their arrival. Every slice equals the same request run alone, and GMP where a golden
    # covers it (the slices below 6e7). The full run (3,146,987 PRP, 3.1M squarings) is verified by
    # its own Gerbicz checks, every 50,000 iterations and past the end; running it alone as well
    # would double the test's time (the server's full-run parity with the plain loop is shown by
    # test_full_mersenne_prime_runs_to_completion's known answer).
    reqs = workloads.poisson_requests(8, 1.0)
    length = [(E - 2 if kind == "ll" else E) if E < 10_000_000 else 50_000 for _, E, kind in reqs]
    assert any(n > 50_000 for n in length)  # the prefix holds a full run
    announce("serving policy prefix", [(E, k, n) for (_, E, k), n in zip(reqs, length)])
    got, ids = {}, {}
    with owl.Server(block_size=1000, check_step=50_000) as s:
        t0, i = s.now(), 0
        while i < len(reqs) or s.info()["active"] or s.info()["waiting"]:
            while i < len(reqs) and reqs[i][0] <= s.now() - t0:
                _, E, kind = reqs[i]
                ids[s.submit(E, kind, length[i])] = (E, kind, length[i])
                i += 1
            if not s.info()["active"] and not s.info()["waiting"]:
                # nothing to run before the next arrival: jump the clock by submitting it now
                _, E, kind = reqs[i]
                ids[s.submit(E, kind, length[i])] = (E, kind, length[i])
                i += 1
            for r in s.step():
                got[r["id"]] = r
    assert set(got) == set(ids)
    for rid, (E, kind, n) in ids.items():
        r = got[rid]
        assert r["status"] == 0 and r["nErrors"] == 0 and r["maxRoe"] < 0.4, r
        if kind == "prp":
            assert r["nChecks"] >= 1
        if n > 50_000:
            assert kind == "prp" and r["nChecks"] >= n // 50_000, r
            continue
        announce("alone", E, kind, n)
        assert r["res64"] == alone_policy(E, kind, n), (E, kind, n)
        gold = load_golden(f"p{E}_{kind}_50000") if n == 50_000 else None
        if gold:
            assert gold == (50_000, r["res64"]), (E, kind)


def alone_policy(E, kind, iters):
    with owl.Gpu(E) as g:
        if kind == "ll":
            out, _ = g.ll(iters)
        else:
            out, _ = g.prp(iters, block_size=1000, check_step=50_000)
        return out.res64
