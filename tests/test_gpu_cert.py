"""The CERT task (Gpu::isCERT, src/Gpu.cpp:1665-1725) on the GPU: `squarings` plain squarings of
the start value read from M<E>.cert, hashed with SHA3-256 over (E-1)/8+1 bytes and reported as the
results line's sha3-hash (the digest's little-endian 64-bit words from the last to the first,
src/Task.cpp:182-186) -- against hashlib over GMP's B^(2^squarings) mod 2^E - 1."""
import hashlib
import random
import sys

import pytest

import gpuowl_amd as owl
import oracle

pytestmark = pytest.mark.gpu


def reference_hex(digest):
    return "".join(format(int.from_bytes(digest[8 * i: 8 * i + 8], "little"), "016x") for i in (3, 2, 1, 0))


@pytest.mark.parametrize("E,squarings,seed", [(1_000_003, 500, 4), (4_000_037, 257, 5)])
def test_cert_matches_gmp(tmp_path, E, squarings, seed):
    print("CASE cert", E, squarings, file=sys.stderr, flush=True)
    M = (1 << E) - 1
    nbytes = (E - 1) // 8 + 1
    B = random.Random(seed).getrandbits(E) % M
    path = tmp_path / f"M{E}.cert"
    path.write_bytes(B.to_bytes(nbytes, "little"))
    with owl.Gpu(E) as g:
        got, r64 = g.cert(path, squarings, remove_file=True)
    assert not path.exists()
    X = owl.words_to_int(oracle.gmp_square_n(E, owl.int_to_words(B, E), squarings)) % M
    assert r64 == X & (2**64 - 1)
    assert got == reference_hex(hashlib.sha3_256(X.to_bytes(nbytes, "little")).digest())


def test_cert_short_file_is_an_error(tmp_path):
    E = 1_000_003
    path = tmp_path / f"M{E}.cert"
    path.write_bytes(b"\x01" * 100)
    with owl.Gpu(E) as g:
        with pytest.raises(owl.OwlError):
            g.cert(path, 200)
