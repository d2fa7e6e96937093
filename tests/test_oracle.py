"""Pins the oracle (CPU): GMP loop vs Python integers and Mersenne-prime known answers;
the C IBDWT restatement vs GMP and vs the reference's own state.cpp (oracle/_ref)."""
import random

import pytest

import oracle

MERSENNE_PRIME_EXPONENTS = [31, 61, 89, 107, 127, 521, 607, 1279, 2203, 2281, 3217, 4253, 4423]


def py_res64(E, iters, step, ll=False):
    M = (1 << E) - 1
    x = 4 if ll else 3
    out = []
    for k in range(1, iters + 1):
        x = (x * x - (2 if ll else 0)) % M
        if k % step == 0:
            out.append(x & (2**64 - 1))
    return out


@pytest.mark.parametrize("E", [61, 127, 1279, 4423, 9941])
def test_gmp_known_answers(E):
    # a Mersenne prime: type-1 PRP residue 9 after E squarings, LL residue 0 after E-2
    assert oracle.gmp_res64(E, E, E) == [9]
    assert oracle.gmp_residue(E, E - 2, ll=True) == [0] * oracle.n_words(E)


@pytest.mark.parametrize("E", [67, 257, 1277, 4099])  # composite Mersenne numbers
def test_gmp_composite(E):
    assert oracle.gmp_res64(E, E, E) != [9]
    assert oracle.gmp_residue(E, E - 2, ll=True) != [0] * oracle.n_words(E)


@pytest.mark.parametrize("E,ll", [(89, False), (521, True), (3217, False), (9689, True)])
def test_gmp_vs_python(E, ll):
    assert oracle.gmp_res64(E, 300, 7, ll) == py_res64(E, 300, 7, ll)


def test_gmp_mulmod():
    rnd = random.Random(5)
    for E in (127, 1279, 86243):
        M = (1 << E) - 1
        a, b = rnd.randrange(M), rnd.randrange(M)
        wa = [(a >> (32 * i)) & 0xFFFFFFFF for i in range(oracle.n_words(E))]
        wb = [(b >> (32 * i)) & 0xFFFFFFFF for i in range(oracle.n_words(E))]
        for mul3 in (False, True):
            got = sum(w << (32 * i) for i, w in enumerate(oracle.gmp_mulmod(E, wa, wb, mul3)))
            assert got == a * b * (3 if mul3 else 1) % M


def test_gmp_square_n():
    # a^(2^n) mod 2^E - 1 against Python's pow, including n = 0, the residue M - 1 (= -1),
    # and the non-canonical input M (= 0); a dropped or extra squaring changes the result
    rnd = random.Random(11)
    for E in (61, 127, 1279, 86243):
        M = (1 << E) - 1
        nw = oracle.n_words(E)
        for a in (rnd.randrange(M), M - 1, 3, M):
            wa = [(a >> (32 * i)) & 0xFFFFFFFF for i in range(nw)]
            for n in (0, 1, 2, 7, 33):
                got = sum(w << (32 * i) for i, w in enumerate(oracle.gmp_square_n(E, wa, n)))
                assert got % M == pow(a, 1 << n, M), (E, a, n)
                assert got < (1 << E)


def test_bitlen_closed_form():
    # bits_k = ceil((k+1)E/N) - ceil(kE/N), extra(k)/N = ceil(kE/N) - kE/N; sum of bits = E
    for E, N in ((127, 16), (4423, 256), (100003, 4096), (4000037, 262144)):
        tot = 0
        for k in range(0, N, max(1, N // 512)):
            ck = -(-k * E // N)
            assert oracle.bitlen(N, E, k) == -(-(k + 1) * E // N) - ck
            assert oracle.extra(N, E, k) == ck * N - k * E
        tot = sum(oracle.bitlen(N, E, k) for k in range(N))
        assert tot == E


def test_expand_compact_roundtrip():
    rnd = random.Random(1)
    for E, N in ((127, 16), (1279, 64), (86243, 4096), (1257787, 65536)):
        M = (1 << E) - 1
        for x in (0, 1, 3, M - 1, rnd.randrange(M), rnd.randrange(M)):
            words = [(x >> (32 * i)) & 0xFFFFFFFF for i in range(oracle.n_words(E))]
            d = oracle.expand(words, N, E)
            for k in range(N):
                b = oracle.bitlen(N, E, k)
                assert -(1 << (b - 1)) <= d[k] < (1 << (b - 1)) or k == 0
            assert oracle.compact(d, N, E) == words
            # the digits represent x: sum d_k 2^(sum of earlier bits) == x mod M
            pos, v = 0, 0
            for k in range(N):
                v += d[k] << pos
                pos += oracle.bitlen(N, E, k)
            assert v % M == x % M


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_expand_compact_vs_reference_state_cpp():
    rnd = random.Random(2)
    for E, N in ((1279, 64), (86243, 4096), (4000037, 262144)):
        M = (1 << E) - 1
        x = rnd.randrange(M)
        words = [(x >> (32 * i)) & 0xFFFFFFFF for i in range(oracle.n_words(E))]
        d = oracle.expand(words, N, E)
        assert d == oracle.ref_expand(words, N, E)
        # a non-normalized (negative-top) digit vector exercises the final borrow
        d2 = list(d)
        d2[-1] = -(1 << (oracle.bitlen(N, E, N - 1) - 1))
        assert oracle.compact(d2, N, E) == oracle.ref_compact(d2, N, E)


def test_residue_from_ends_matches_compact():
    rnd = random.Random(3)
    for E, N in ((86243, 4096), (1257787, 65536)):
        M = (1 << E) - 1
        for _ in range(4):
            x = rnd.randrange(M)
            words = [(x >> (32 * i)) & 0xFFFFFFFF for i in range(oracle.n_words(E))]
            d = oracle.expand(words, N, E)
            ends = d[N - 32:] + d[:32]
            assert oracle.residue_from_ends(ends, N, E) == x & (2**64 - 1)


@pytest.mark.parametrize("E,N,ll", [(1279, 128, False), (4423, 512, True), (86243, 8192, False),
                                    (100003, 3 * 4096, False)])
def test_ibdwt_vs_gmp(E, N, ll):
    got, roe = oracle.ibdwt_run(E, N, 60, 10, ll)
    assert got == oracle.gmp_res64(E, 60, 10, ll)
    assert roe < 0.1


def test_ibdwt_mersenne_prime_ll():
    E, N = 2203, 256
    d = oracle.expand([4] + [0] * (oracle.n_words(E) - 1), N, E)
    for _ in range(E - 2):
        d, roe = oracle.ibdwt_square(d, N, E, ll=True)
    assert oracle.compact(d, N, E) == [0] * oracle.n_words(E) or all(x == 0 for x in d)


def test_ibdwt_detects_planted_error():
    # a planted mistake in one digit must change the res64 stream
    E, N = 4423, 512
    d = oracle.expand([3] + [0] * (oracle.n_words(E) - 1), N, E)
    d[7] += 1
    x = d
    for _ in range(20):
        x, _ = oracle.ibdwt_square(x, N, E)
    ends = x[N - 32:] + x[:32]
    assert oracle.residue_from_ends(ends, N, E) != oracle.gmp_res64(E, 20, 20)[0]
