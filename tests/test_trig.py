"""Twiddle tables against mpmath (CPU).

root1(n, k) (TrigBufCache's root1, src/TrigBufCache.cpp:22-52) is the entry every table of the
engine holds: exp(-2 pi i k / n) rounded once to double. It must be within 1 ulp of the exact
value and on the unit circle to |c^2 + s^2 - 1| <= 2^-52. The kernels build most twiddles as
products of two (the column pass's first stage: three) such entries,
w^m = TA[m mod 64] * TB[m div 64]; those products are emulated here bit for bit (the kernels'
complex multiply is fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)) and checked against
the exact value: each factor is off by at most half an ulp (2^-54 for entries in [1/2, 1)) and
each product adds a rounding, so two factors stay within 2^-52 and three within 1.5 * 2^-52."""
import math
import random
from fractions import Fraction

import mpmath
import pytest

import gpuowl_amd as owl

mpmath.mp.prec = 200
TWO = mpmath.mpf(2)


def exact(n, k):
    a = 2 * mpmath.pi * (k % n) / n
    return mpmath.cos(a), -mpmath.sin(a)


def ulp(x):
    """The ulp of the double nearest x; floored at 1e-50, far below every double of interest and
    above mpmath's own error at 200 bits (cos(pi/2) comes out as ~1e-61, not 0)."""
    x = abs(float(x))
    if x < 1e-40:
        return 1e-50
    m, e = math.frexp(x)  # x = m * 2^e, m in [0.5, 1)
    return 2.0 ** (e - 53)


def fma(a, b, c):
    """a * b + c rounded once (Python's int / int division rounds correctly)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def cmul(a, b):
    return (fma(a[0], b[0], -(a[1] * b[1])), fma(a[0], b[1], a[1] * b[0]))


def err(z, n, k):
    c, s = exact(n, k)
    return max(abs(mpmath.mpf(z[0]) - c), abs(mpmath.mpf(z[1]) - s))


def sample_ks(n, count, seed):
    rnd = random.Random(seed)
    ks = set(range(0, min(n, 70)))
    for j in range(9):  # octant boundaries and their neighbours
        base = n * j // 8
        ks.update(x % n for x in (base - 1, base, base + 1))
    if n <= 4096:
        ks.update(range(n))
    else:
        ks.update(rnd.randrange(n) for _ in range(count))
    return sorted(ks)


@pytest.mark.parametrize("n", [256, 1024, 3840, 4096, 262144 // 2, 7864320 // 2])
def test_root1_within_one_ulp_and_on_unit_circle(n):
    for k in sample_ks(n, 1500, n):
        c, s = owl.root1(n, k)
        ec, es = exact(n, k)
        assert abs(mpmath.mpf(c) - ec) <= ulp(ec), (n, k, "cos")
        assert abs(mpmath.mpf(s) - es) <= ulp(es), (n, k, "sin")
        norm = Fraction(c) ** 2 + Fraction(s) ** 2 - 1
        assert abs(norm) <= Fraction(1, 2 ** 52), (n, k)
    # exact at multiples of pi/2
    for j, want in enumerate([(1.0, 0.0), (0.0, -1.0), (-1.0, 0.0), (0.0, 1.0)]):
        assert owl.root1(n, n * j // 4) == want


@pytest.mark.parametrize("L", [512, 1024, 3840, 4096])
def test_two_factor_table_twiddles(L):
    # w_L^m = TA[m & 63] * TB[m >> 6] with TA[i] = root1(L, i), TB[j] = root1(L, 64 j):
    # the row pass's LDS table (W >= 512) and the column pass's w_H^m
    worst = 0.0
    for m in range(L):
        z = cmul(owl.root1(L, m & 63), owl.root1(L, 64 * (m >> 6)))
        worst = max(worst, float(err(z, L, m)))
    assert worst <= 2.0 ** -52, worst


@pytest.mark.parametrize("W,H", [(1024, 3840), (512, 1536), (256, 512)])
def test_mid_twiddles(W, H):
    # the column pass's mid twiddle w_ND^(a l) = MA[a][l & 63] * MB[a][l >> 6] and, in the first
    # stage, base * S[a][k] with S[a][k] = w_ND^(a k H/16) (three factors)
    ND = W * H
    rnd = random.Random(W + H)
    worst2 = worst3 = 0.0
    for _ in range(300):
        a, j, k = rnd.randrange(W), rnd.randrange(H // 16), rnd.randrange(1, 16)
        base = cmul(owl.root1(ND, a * (j & 63)), owl.root1(ND, a * 64 * (j >> 6) % ND))
        worst2 = max(worst2, float(err(base, ND, a * j)))
        z = cmul(base, owl.root1(ND, a * k * (H // 16) % ND))
        worst3 = max(worst3, float(err(z, ND, a * (j + k * (H // 16)))))
    assert worst2 <= 2.0 ** -52, worst2
    assert worst3 <= 1.5 * 2.0 ** -52, worst3
