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
