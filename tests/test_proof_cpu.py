"""Proof building blocks on the CPU: SHA3-256 against hashlib (FIPS 202), including the
136-byte block boundary, and the proof points of src/Proof.cpp:130-164 through properties
that fix them: 2^power sorted points ending at E, power 1 = {ceil(E/2), E}, each power's
points contain the previous power's, and every point below E is a sum of distinct level
spans ceil(E/2), ceil(ceil(E/2)/2), ..."""
import hashlib
import itertools
import random

import pytest

import gpuowl_amd as owl


def test_sha3_256_matches_hashlib():
    rng = random.Random(3)
    for n in (0, 1, 31, 32, 135, 136, 137, 271, 272, 273, 1000, 125_001):
        data = bytes(rng.getrandbits(8) for _ in range(n))
        assert owl.sha3_256(data) == hashlib.sha3_256(data).digest(), n
    assert owl.sha3_256(b"abc").hex() == "3a985da74fe225b2045c172d6bd390bd855f086e3e9d525b46bfe24511431532"


def spans(E, power):
    out, s = [], (E + 1) // 2
    for _ in range(power):
        out.append(s)
        s = (s + 1) // 2
    return out


@pytest.mark.parametrize("E", [1_000_003, 859_433, 124_647_911])
def test_proof_points(E):
    assert owl.proof_points(E, 1) == [(E + 1) // 2, E]
    prev = None
    for power in range(1, 11):
        pts = owl.proof_points(E, power)
        assert len(pts) == 2 ** power and pts == sorted(pts) and pts[-1] == E and len(set(pts)) == len(pts)
        if prev is not None:
            assert set(prev) <= set(pts)
        prev = pts
        if power <= 6:
            sums = {sum(c) for r in range(1, power + 1) for c in itertools.combinations(spans(E, power), r)}
            assert set(pts[:-1]) <= sums


def test_bad_power_is_an_error():
    for power in (0, 13):
        with pytest.raises(owl.OwlError):
            owl.proof_points(1_000_003, power)
    with pytest.raises(owl.OwlError):
        owl.proof_points(1_000_004, 4)   # E must be odd


def test_reference_proof_algorithm_round_trip():
    # tests/proofref.py restates the reference's proof construction and verification; a proof it
    # builds for a Mersenne prime and for a composite verifies, and one altered bit of B or of any
    # middle makes it invalid (each enters the hash chain and the final identity)
    from tests import proofref
    for E, power, prime in ((4423, 3, True), (4099, 4, False)):
        raw = proofref.make(E, power)
        assert proofref.verify(raw) == (True, prime)
        assert proofref.points(E, power) == owl.proof_points(E, power)
        _, _, b, mids = proofref.parse(raw)
        head = len(raw) - (power + 1) * proofref.nbytes(E)
        for part in range(power + 1):
            bad = bytearray(raw)
            bad[head + part * proofref.nbytes(E) + 3] ^= 0x10
            assert proofref.verify(bytes(bad))[0] is False, (E, part)


def test_reference_proof_gmp_arithmetic_agrees():
    # the GMP-backed arithmetic the GPU proof test uses for E near 10^6 equals the Python one
    import oracle
    from tests import proofref
    rnd = random.Random(9)
    E = 4423
    py, gmp = proofref.Arith(E), proofref.GmpArith(E, oracle)
    for _ in range(5):
        a, b, h = rnd.randrange(py.M), rnd.randrange(py.M), rnd.getrandbits(64)
        assert gmp.mul(a, b) == py.mul(a, b) == a * b % py.M
        assert gmp.pow(a, h) == py.pow(a, h) == pow(a, h, py.M)
        assert gmp.square_n(a, 13) == py.square_n(a, 13) == pow(a, 2**13, py.M)
