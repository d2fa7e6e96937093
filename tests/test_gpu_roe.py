"""Round-off at the top of each FFT shape's range: the largest prime that the reference's
bits-per-word table gives a shape (its maxExp, fitted at Gumbel Z = 28 on the reference's own
kernels, src/fftbpw.h, src/tune.cpp:106-172) run on this engine's kernels for 2,000 measured
squarings (Gpu::measureROE, src/Gpu.cpp:1268-1335). This engine takes twiddles as products of
two correctly rounded table entries and carries in FP64, so its round-off differs from the
reference's; the Z it reaches at each maxExp is what bounds the exponents it can take there."""
import sys

import pytest

import gpuowl_amd as owl
from gpuowl_amd import workloads

pytestmark = pytest.mark.gpu

SHAPES = ["256:2:256", "1K:3:256", "1K:15:256", "512:15:512", "256:9:256", "1K:5:256"]


def prev_prime(n):
    while not workloads.is_prime(n):
        n -= 1
    return n


@pytest.mark.parametrize("spec", SHAPES)
def test_round_off_at_max_exponent(spec):
    full, N, max_exp = owl.best_fit(1000, spec)
    E = prev_prime(max_exp)
    with owl.Gpu(E, spec) as g:
        r = g.measure_roe(2000, 50)
    print(f"ROE {spec} E={E} bpw={E / N:.3f} max={r['max']:.4f} mean={r['mean']:.4f} sd={r['sd']:.4f} "
          f"Z={r['z']:.2f}", file=sys.stderr, flush=True)
    assert r["n"] == 2000
    assert r["max"] < 0.4, r
    assert r["z"] >= 28, r


def test_round_off_statistics_track_the_bits_per_word():
    # the same shape at three exponents: round-off grows with the bits per word
    spec = "256:2:256"
    _, N, max_exp = owl.best_fit(1000, spec)
    got = []
    for bpw in (14.0, 16.0, max_exp / N):
        with owl.Gpu(prev_prime(int(bpw * N)), spec) as g:
            got.append(g.measure_roe(500, 20))
    assert got[0]["mean"] < got[1]["mean"] < got[2]["mean"]
    assert got[0]["z"] > got[1]["z"] > got[2]["z"]
