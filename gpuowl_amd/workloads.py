"""Synthetic benchmark workloads (exponents only; every residue starts from 3 for PRP or
4 for LL).

  p4m()           p = smallest prime >= 4,000,000 at the reference's FFT choice (256:2:256)
  p139m()         p = smallest prime >= 139,000,000 at 1K:15:256 (N = 7,864,320)
  uniform64_20m() the first 64 primes >= 20,000,000, all at 1K:3:256 (N = 1,572,864)
  uneven200()     200 primes log-uniform in [1e6, 6e7]: splitmix64(seed 1234), u = (x >> 11) * 2^-53,
                  e = floor(exp(ln 1e6 + u ln 60)), p = nextprime(e), redrawn on a duplicate
  poisson_requests(count, lam)
                  serving arrivals: splitmix64(seed 7); draws alternate exponent u (log-uniform
                  prime in [2e6, 1.4e8]), kind u (PRP if < 0.5 else LL), inter-arrival -ln(1-u)/lambda
"""
import math

MASK = (1 << 64) - 1


def splitmix64(seed):
    x = seed & MASK
    while True:
        x = (x + 0x9E3779B97F4A7C15) & MASK
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        yield z ^ (z >> 31)


def is_prime(n):
    if n < 2:
        return False
    small = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)
    for p in small:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def next_prime(n):
    while not is_prime(n):
        n += 1
    return n


def primes_from(n, count):
    out, p = [], n
    while len(out) < count:
        p = next_prime(p)
        out.append(p)
        p += 1
    return out


def p4m():
    return next_prime(4_000_000)


def p139m():
    return next_prime(139_000_000)


def uniform64_20m():
    return primes_from(20_000_000, 64)


def uneven200(count=200, seed=1234, lo=1e6, span=60.0):
    g = splitmix64(seed)
    out, seen = [], set()
    while len(out) < count:
        u = (next(g) >> 11) * 2.0**-53
        p = next_prime(int(math.exp(math.log(lo) + u * math.log(span))))
        if p in seen:
            continue
        seen.add(p)
        out.append(p)
    return out


def poisson_requests(count, lam, seed=7, lo=2e6, hi=1.4e8):
    """[(arrival_time_s, exponent, kind)] for count Poisson arrivals at rate lam per second."""
    g = splitmix64(seed)
    t, out = 0.0, []
    for _ in range(count):
        u = (next(g) >> 11) * 2.0**-53
        p = next_prime(int(math.exp(math.log(lo) + u * math.log(hi / lo))))
        kind = "prp" if (next(g) >> 11) * 2.0**-53 < 0.5 else "ll"
        t += -math.log(1.0 - (next(g) >> 11) * 2.0**-53) / lam
        out.append((t, p, kind))
    return out
