"""PRP proofs on the GPU (src/Proof.cpp:84-126, 245-297): A^h * B against Python integers,
and whole proofs for a Mersenne prime and a composite: made from a full PRP-3 run,
verified, and rejected once a middle or B is altered."""
import random
import sys

import pytest

import gpuowl_amd as owl

pytestmark = pytest.mark.gpu


def announce(*a):
    print("CASE", *a, file=sys.stderr, flush=True)


def words(x, E):
    return owl.int_to_words(x, E)


def test_exp_mul_matches_python_integers():
    E = 1_000_003
    M = (1 << E) - 1
    rng = random.Random(21)
    A, B = rng.randrange(2, M), rng.randrange(2, M)
    announce("exp_mul", E)
    with owl.Gpu(E) as g:
        for h, sq in [(0, False), (1, False), (0b1011, True), (0b1100101001, False)]:
            got = owl.words_to_int(g.exp_mul(words(A, E), h, words(B, E), sq))
            want = pow(A, h, M) * pow(B, 2 if sq else 1, M) % M
            assert got == want, (h, sq)


def proof_case(tmp_path, E, power, prime):
    path = tmp_path / f"{E}-{power}.proof"
    with owl.Gpu(E) as g:
        out = g.prp_proof(tmp_path / "residues", path, power, block_size=1000)
        assert out.k == E and out.isPrime == prime and out.nErrors == 0
        nbytes = (E - 1) // 8 + 1
        raw = path.read_bytes()
        header = b"PRP PROOF\nVERSION=2\nHASHSIZE=64\nPOWER=%d\nNUMBER=M%d\n" % (power, E)
        assert raw.startswith(header) and len(raw) == len(header) + (power + 1) * nbytes
        assert g.proof_verify(path) == (True, prime)
        # altered proofs fail: one bit of the last middle, one bit of B
        for offset in (len(raw) - nbytes // 2, len(header) + 7):
            bad = bytearray(raw)
            bad[offset] ^= 0x04
            tampered = tmp_path / "tampered.proof"
            tampered.write_bytes(bytes(bad))
            assert g.proof_verify(tampered)[0] is False, offset


@pytest.mark.slow
def test_proof_of_a_mersenne_prime(tmp_path):
    announce("proof M859433 power 4")
    proof_case(tmp_path, 859_433, 4, True)


@pytest.mark.slow
def test_proof_of_a_composite(tmp_path):
    announce("proof M1000003 power 5")
    proof_case(tmp_path, 1_000_003, 5, False)


@pytest.mark.slow
def test_engine_proof_passes_reference_algorithm(tmp_path):
    # the engine's proof file, checked by the reference's algorithm restated in Python
    # (tests/proofref.py, GMP arithmetic) rather than by the engine's own verifier
    import oracle
    from tests import proofref
    E, power = 1_000_003, 9
    announce("proof M1000003 power 9, reference-algorithm check")
    path = tmp_path / f"{E}-{power}.proof"
    with owl.Gpu(E) as g:
        out = g.prp_proof(tmp_path / "residues", path, power, block_size=1000)
    assert out.k == E and not out.isPrime and out.nErrors == 0
    raw = path.read_bytes()
    assert proofref.parse(raw)[:2] == (E, power)
    assert proofref.points(E, power) == owl.proof_points(E, power)
    # (altered files are rejected by the same verifier in tests/test_proof_cpu.py)
    assert proofref.verify(raw, proofref.GmpArith(E, oracle)) == (True, False)
