"""GPU parity: the HIP path (through the C-ABI) against the oracle on the same inputs.
res64 and full residues must be bit-exact; round-off is bounded."""
import random
import sys

import pytest

import gpuowl_amd as owl
import oracle
from gpuowl_amd import workloads

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, file=sys.stderr, flush=True)


def exponent_for(spec, bpw):
    N = owl.best_fit(1000, spec)[1]
    return workloads.next_prime(int(bpw * N)), N


# (spec, bits per word): shapes covering radices 2..16 in both passes
SHAPES = [
    ("64:2:64", 12.0),
    ("256:2:256", 15.3),
    ("256:3:256", 8.0),
    ("256:5:256", 9.0),
    ("256:7:256", 9.0),
    ("256:11:256", 8.0),
    ("256:13:256", 8.0),
    ("512:9:256", 8.0),
    ("1K:3:256", 8.0),
    ("512:6:512", 5.0),
    ("256:9:512", 12.0),    # H = 4608: run as 1024 x 1152
    ("256:15:512", 17.0),   # H = 7680: run as 1024 x 1920
]


@pytest.mark.parametrize("spec,bpw", SHAPES)
def test_prp_res64_vs_gmp(spec, bpw):
    E, N = exponent_for(spec, bpw)
    announce("prp", spec, E)
    iters, step = 120, 40
    with owl.Gpu(E, spec) as g:
        out, res = g.prp(iters, block_size=40, check_step=40, res_step=step)
        assert res == oracle.gmp_res64(E, iters, step)
        assert out.nErrors == 0 and out.nChecks == iters // 40
        assert out.maxRoe < 0.4


@pytest.mark.parametrize("spec,bpw", [("64:2:64", 12.0), ("256:3:256", 8.0), ("1K:3:256", 8.0)])
def test_ll_res64_vs_gmp(spec, bpw):
    E, N = exponent_for(spec, bpw)
    announce("ll", spec, E)
    with owl.Gpu(E, spec) as g:
        out, res = g.ll(90, res_step=30)
        assert res == oracle.gmp_res64(E, 90, 30, ll=True)
        assert out.res64 == res[-1]


VARIANTS = {"default": {}, "generic": {"generic": True}, "coly": {"col_one_region": True},
            "onepair": {"col_one_pair": True}}


@pytest.mark.parametrize("spec,bpw,variant", [("64:2:64", 12.0, "default"), ("256:2:256", 15.0, "default"),
                                              ("256:2:256", 15.0, "generic"), ("256:2:256", 15.0, "coly"),
                                              ("256:13:256", 8.0, "default"), ("1K:3:256", 12.0, "default"),
                                              ("1K:3:256", 12.0, "generic"), ("1K:3:256", 12.0, "coly")])
def test_full_residue_vs_gmp(spec, bpw, variant):
    E, N = exponent_for(spec, bpw)
    announce("residue", spec, E, variant)
    want = oracle.gmp_residue(E, 50)
    with owl.Gpu(E, spec, **VARIANTS[variant]) as g:
        g.write_int(owl.DATA, 3)
        g.square_loop(owl.DATA, owl.DATA, 0, 50)
        same = g.read_words(owl.DATA) == want   # (not inside assert: pytest would diff megabit lists)
        assert same, "full residue differs from GMP"
        assert g.res64(owl.DATA) == want[0] | (want[1] << 32)


@pytest.mark.parametrize("spec,bpw", [("64:2:64", 12.0), ("256:5:256", 9.0)])
def test_mod_mul_vs_gmp(spec, bpw):
    E, N = exponent_for(spec, bpw)
    announce("modmul", spec, E)
    rnd = random.Random(11)
    M = (1 << E) - 1
    with owl.Gpu(E, spec) as g:
        for mul3 in (False, True):
            a, b = rnd.randrange(M), rnd.randrange(M)
            g.write_int(owl.DATA, a)
            g.write_int(owl.AUX, b)
            g.mod_mul(owl.DATA, owl.AUX, mul3)
            wa, wb = owl.int_to_words(a, E), owl.int_to_words(b, E)
            same = g.read_words(owl.DATA) == oracle.gmp_mulmod(E, wa, wb, mul3)
            assert same, "product differs from GMP"
            same = g.read_words(owl.AUX) == wb
            assert same, "second operand was modified"


def test_gerbicz_check_and_fault():
    spec = "256:2:256"
    E, N = exponent_for(spec, 15.0)
    announce("gerbicz", spec, E)
    B = 25
    rnd = random.Random(3)
    nw = owl.n_words(E)
    c = [rnd.getrandbits(32) for _ in range(nw)]
    c[-1] &= (1 << (E % 32)) - 1
    # the Gerbicz relation: data = check^(2^B - 1) * 3, i.e. check * data = check^(2^B) * 3
    # data is found as y with y * c == c^(2^B) * 3 by building it as c^(2^B - 1) = prod c^(2^i), i < B
    d = [1] + [0] * (nw - 1)
    p = list(c)
    for i in range(B):
        d = oracle.gmp_mulmod(E, d, p)
        p = oracle.gmp_square_n(E, p, 1)
    d = oracle.gmp_mulmod(E, d, [3] + [0] * (nw - 1))
    with owl.Gpu(E, spec) as g:
        g.write_words(owl.CHECK, c)
        g.write_words(owl.DATA, d)
        assert g.do_check(B)
        same = g.read_words(owl.CHECK) == oracle.gmp_mulmod(E, c, d)
        assert same, "check update differs from GMP"
        g.write_words(owl.CHECK, c)
        d2 = list(d)
        d2[5] ^= 1 << 7                           # a fault in the data
        g.write_words(owl.DATA, d2)
        assert not g.do_check(B)


def test_prp_recovers_from_fault_after_a_verified_check():
    # one flipped data bit at k = 230: the check at 300 fails, the run reloads the state
    # verified at 200 (whose check already holds x_200) and ends on GMP's residues
    spec = "256:2:256"
    E = workloads.primes_from(3_800_000, 1)[0]
    announce("gerbicz recovery", spec, E)
    with owl.Gpu(E, spec) as g:
        g.arm_fault(230)
        out, rows = g.prp(400, block_size=50, check_step=100, res_step=100)
    assert out.nErrors == 1 and out.nChecks == 4
    assert rows == oracle.gmp_res64(E, 400, 100)


def test_is_equal_and_copy():
    spec = "64:2:64"
    E, N = exponent_for(spec, 12.0)
    with owl.Gpu(E, spec) as g:
        g.write_int(owl.DATA, 12345)
        g.copy(owl.AUX, owl.DATA)
        assert g.is_equal(owl.DATA, owl.AUX)
        g.write_int(owl.AUX, 12346)
        assert not g.is_equal(owl.DATA, owl.AUX)


def test_p4000037_prp_golden(golden):
    head, res = golden("p4000037_prp")
    E = head["E"]
    announce("p = 4,000,037 golden", E)
    with owl.Gpu(E) as g:
        assert g.spec == "256:2:256:101"
        out, got = g.prp(100_000, block_size=1000, check_step=10_000, res_step=1000)
    assert got == [res[k] for k in range(1000, 100_001, 1000)]
    assert out.nErrors == 0 and out.nChecks == 10 and out.maxRoe < 0.4


@pytest.mark.parametrize("variant", ["default", "coly", "generic", "onepair"])
def test_p139000003_prp_golden_short(golden, variant):
    head, res = golden("p139000003_prp_short")
    E = head["E"]
    announce("p = 139,000,003 prp golden", E, variant)
    with owl.Gpu(E, **VARIANTS[variant]) as g:
        assert g.spec == "1K:15:256:101" and g.N == 7_864_320
        out, got = g.prp(200, block_size=100, check_step=100, res_step=100)
        assert got == [res[100], res[200]]
        assert out.nErrors == 0 and out.maxRoe < 0.4


def test_p139000003_ll_golden_short(golden):
    head, res = golden("p139000003_ll_short")
    E = head["E"]
    announce("p = 139,000,003 ll golden", E)
    with owl.Gpu(E) as g:
        out, got = g.ll(200, res_step=100)
        assert got == [res[100], res[200]]


def test_p139000003_prp_golden_20000(golden):
    # the full stream: 20,000 PRP-3 iterations at 7.5M words, res64 every 1,000, Gerbicz
    # checks every 5,000, against GMP
    head, res = golden("p139000003_prp")
    E = head["E"]
    announce("p = 139,000,003 prp golden, 20,000 iterations", E)
    with owl.Gpu(E) as g:
        out, got = g.prp(20_000, block_size=1000, check_step=5000, res_step=1000)
    assert got == [res[k] for k in range(1000, 20_001, 1000)]
    assert out.nErrors == 0 and out.nChecks == 4 and out.maxRoe < 0.4


def test_p139000003_ll_golden_20000(golden):
    head, res = golden("p139000003_ll")
    E = head["E"]
    announce("p = 139,000,003 ll golden, 20,000 iterations", E)
    with owl.Gpu(E) as g:
        out, got = g.ll(20_000, res_step=1000)
    assert got == [res[k] for k in range(1000, 20_001, 1000)]


@pytest.mark.parametrize("spec,bpw", [("64:2:64", 3.0), ("256:3:256", 3.2), ("256:3:256", 8.0), ("1K:3:256", 5.0)])
def test_low_bits_per_word_stays_exact(spec, bpw):
    # the fused carry must stay stable (no growth of the hand-off word) at low bits per word
    E, N = exponent_for(spec, bpw)
    announce("low bpw", spec, E)
    with owl.Gpu(E, spec) as g:
        out, res = g.prp(200, block_size=100, check_step=100, res_step=50)
        assert res == oracle.gmp_res64(E, 200, 50)
        assert out.nErrors == 0 and out.maxRoe < 0.1


def test_long_carry_mode_matches():
    spec = "256:3:256"
    E, N = exponent_for(spec, 8.0)
    with owl.Gpu(E, spec, long_carry=True) as g:
        out, res = g.prp(60, block_size=20, check_step=20, res_step=20)
        assert res == oracle.gmp_res64(E, 60, 20)


@pytest.mark.parametrize("spec,bpw", [("64:2:64", 3.0), ("256:3:256", 3.1)])
def test_carries_past_whole_blocks_match(spec, bpw):
    # one line per block and a full carry every iteration at ~3 bits per word: carries often
    # cross whole 2-word blocks, so the carry fix (the marked ranks, in word order) runs on most
    # iterations
    E, N = exponent_for(spec, bpw)
    announce("carries past blocks", spec, E)
    with owl.Gpu(E, spec, long_carry=True, line_block=1) as g:
        out, res = g.prp(40, block_size=20, check_step=20, res_step=20)
        assert res == oracle.gmp_res64(E, 40, 20)
        assert out.nErrors == 0


@pytest.mark.slow
def test_mersenne_prime_1257787():
    # M1257787 is prime: the type-1 PRP residue after E squarings is 9
    E = 1_257_787
    announce("M1257787 full PRP")
    with owl.Gpu(E) as g:
        out, _ = g.prp(E, block_size=1000, check_step=100_000)
        assert out.isPrime and out.res64 == 9 and out.nErrors == 0
