"""Writes the res64 golden fixtures of tests/golden/ with the GMP oracle
(oracle/_build/gmp_loop), one process per job, in parallel.

Each fixture res_<name>.txt holds a header line and "k res64" lines:
    # E=<E> kind=<prp|ll> iters=<n> step=<s>
    1000 0123456789abcdef
Fixtures are named by exponent and kind:
  p4000037_prp            100,000 PRP-3 iterations of p = 4,000,037 (smallest prime >= 4,000,000)
  p139000003_{prp,ll}     20,000 PRP-3 / LL iterations of p = 139,000,003, and *_short: 200
  p<E>_prp                1,000 PRP-3 iterations of each of the first 64 primes >= 20,000,000
                          and of each of the 200 uneven exponents (splitmix64 seed 1234,
                          log-uniform primes in [1e6, 6e7]; gpuowl_amd.workloads.uneven200)
  p<E>_<prp|ll>_50000     the 50,000-iteration serving slices of the first 8 Poisson requests
                          (seed 7) with exponents in [1e7, 6e7)
Usage: python tests/golden/make_goldens.py [--jobs N] [--set all|uneven200] [names...]"""
import argparse, os, subprocess, sys, concurrent.futures as cf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GMP = os.path.join(ROOT, "oracle", "_build", "gmp_loop")


def is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
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
    out = []
    p = n
    while len(out) < count:
        p = next_prime(p)
        out.append(p)
        p += 1
    return out


def jobs():
    J = {}
    A = next_prime(4_000_000)
    B = next_prime(139_000_000)
    J[f"p{A}_prp"] = (A, 0, 100_000, 1000)
    J[f"p{B}_prp"] = (B, 0, 20_000, 1000)
    J[f"p{B}_ll"] = (B, 1, 20_000, 1000)
    J[f"p{B}_prp_short"] = (B, 0, 200, 100)
    J[f"p{B}_ll_short"] = (B, 1, 200, 100)
    for p in primes_from(20_000_000, 64):
        J[f"p{p}_prp"] = (p, 0, 1000, 500)
    for p in uneven200():
        J[f"p{p}_prp"] = (p, 0, 1000, 500)
    for E, kind in serving_slices():
        J[f"p{E}_{kind}_50000"] = (E, int(kind == "ll"), 50_000, 50_000)
    return J


def serving_slices(count=8, full_below=10_000_000, gmp_below=60_000_000):
    """The 50,000-iteration slices among the first `count` serving requests (splitmix64 seed 7)
    whose GMP run stays within a few core-hours (exponents below gmp_below)."""
    sys.path.insert(0, ROOT)
    from gpuowl_amd import workloads
    return [(E, kind) for _, E, kind in workloads.poisson_requests(count, 1.0)
            if full_below <= E < gmp_below]


def uneven200():
    sys.path.insert(0, ROOT)
    from gpuowl_amd import workloads
    return workloads.uneven200()


def run(name, E, ll, iters, step):
    path = os.path.join(HERE, f"res_{name}.txt")
    tmp = path + ".part"
    with open(tmp, "w") as f:
        f.write(f"# E={E} kind={'ll' if ll else 'prp'} iters={iters} step={step}\n")
        f.flush()
        p = subprocess.Popen([GMP, str(E), str(iters), str(ll), str(step)], stdout=subprocess.PIPE, text=True)
        for line in p.stdout:
            if line.startswith("{"):
                continue
            f.write(line)
            f.flush()
        rc = p.wait()
    if rc:
        raise RuntimeError(f"{name}: gmp_loop exit {rc}")
    os.replace(tmp, path)
    return name


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--set", default="all", choices=["all", "uneven200", "serving"])
    ap.add_argument("--missing", action="store_true", help="only fixtures not written yet")
    ap.add_argument("names", nargs="*")
    a = ap.parse_args()
    all_jobs = jobs()
    if a.names:
        names = a.names
    elif a.set == "uneven200":
        names = [f"p{p}_prp" for p in uneven200()]
    elif a.set == "serving":
        names = [f"p{E}_{kind}_50000" for E, kind in serving_slices()]
    else:
        names = list(all_jobs)
    if a.missing:
        names = [n for n in names if not os.path.exists(os.path.join(HERE, f"res_{n}.txt"))]
    # largest exponents first, so the longest jobs do not start last
    names.sort(key=lambda n: -all_jobs[n][0] * all_jobs[n][2])
    with cf.ThreadPoolExecutor(max_workers=a.jobs) as ex:
        futs = [ex.submit(run, n, *all_jobs[n]) for n in names]
        for f in cf.as_completed(futs):
            print("done", f.result(), flush=True)
