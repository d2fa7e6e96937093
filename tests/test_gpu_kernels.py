"""Kernel-level numerics: each HIP pass against the numpy FP64 model of the same op
(tests/fftmodel.py) on dense random digits. FFT outputs must agree to a relative
1e-12 (FP64 with correctly rounded twiddles); rounded digits must agree exactly."""
import re

import numpy as np
import pytest

import gpuowl_amd as owl
from gpuowl_amd import workloads
from tests.fftmodel import Model

pytestmark = pytest.mark.gpu

# (spec, bits per word, kernel variant): "default" = shape-specialized passes where compiled
# (two-region column pass, two column pairs per group at 1K:15:256), "coly" = the one-region
# column pass, "onepair" = the two-region pass with one pair per group, "generic" = runtime-plan
# kernels
SHAPES = [("64:2:64", 12.0, "default"), ("256:2:256", 15.0, "default"), ("256:2:256", 15.0, "generic"),
          ("256:2:256", 15.0, "coly"), ("256:3:256", 8.0, "default"), ("256:3:256", 8.0, "coly"),
          ("1K:3:256", 8.0, "default"), ("1K:3:256", 8.0, "generic"), ("1K:3:256", 8.0, "coly"),
          ("256:13:256", 8.0, "default"), ("512:9:256", 8.0, "default"), ("1K:15:256", 17.6, "default"),
          ("1K:15:256", 17.6, "generic"), ("1K:15:256", 17.6, "coly"), ("1K:15:256", 17.6, "onepair")]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def setup(spec, bpw, variant, seed=1):
    N = owl.best_fit(1000, spec)[1]
    E = workloads.next_prime(int(bpw * N))
    g = owl.Gpu(E, spec, generic=variant == "generic", col_one_region=variant == "coly",
                col_one_pair=variant == "onepair")
    m = Model(E, g.W, g.H, g.R)
    return g, m, np.random.default_rng(seed)


@pytest.mark.parametrize("spec,bpw,variant", SHAPES)
def test_lead_in_and_square(spec, bpw, variant):
    g, m, rng = setup(spec, bpw, variant)
    with g:
        d = m.random_digits(rng)
        g.write_digits(owl.DATA, d.tolist())
        g.run_pass(g.PASS_LEADIN, owl.DATA)
        f = g.read_freq()
        assert rel(f, m.lead_in(d)) < 1e-12
        g.run_pass(g.PASS_SQUARE)
        sq = g.read_freq()
        assert rel(sq, m.col_square(f)) < 1e-12


@pytest.mark.parametrize("spec,bpw,variant", SHAPES)
def test_fused_row_pass(spec, bpw, variant):
    g, m, rng = setup(spec, bpw, variant, seed=2)
    with g:
        d = m.random_digits(rng)
        g.write_digits(owl.DATA, d.tolist())
        g.run_pass(g.PASS_LEADIN, owl.DATA)
        g.run_pass(g.PASS_SQUARE)
        sq = g.read_freq()
        g.run_pass(g.PASS_FUSED)
        got = g.read_freq()
        digits, roe = m.fused_digits(m.row_inverse(sq))
        want = m.lead_in(digits)
        err = np.abs(got - want) / np.max(np.abs(want))
        bad = np.nonzero(err > 1e-11)[0]
        assert len(bad) == 0, f"{len(bad)} elements differ; first at HBM index {bad[:8]}"


@pytest.mark.parametrize("spec,bpw,variant", SHAPES)
def test_carry_out_digits(spec, bpw, variant):
    g, m, rng = setup(spec, bpw, variant, seed=3)
    with g:
        d = m.random_digits(rng)
        g.write_digits(owl.DATA, d.tolist())
        g.run_pass(g.PASS_LEADIN, owl.DATA)
        g.run_pass(g.PASS_SQUARE)
        sq = g.read_freq()
        g.run_pass(g.PASS_CARRYOUT, owl.AUX)
        got = np.array(g.read_digits(owl.AUX))
        want, roe = m.full_digits(m.row_inverse(sq))
        bad = np.nonzero(got != want)[0]
        assert len(bad) == 0, f"{len(bad)} digits differ; first at k = {bad[:8]}"


def test_kernel_selection():
    # the 7.5M-word shape runs the shape-specialized passes (two column pairs per group: W/4 = 256
    # groups, one per CU); generic=True forces the runtime-plan ones
    E = workloads.next_prime(139_000_000)
    with owl.Gpu(E) as g:
        assert (g.row_kernel, g.col_kernel) == ("rowx1024r4", "colx3840m")
    with owl.Gpu(E, col_one_pair=True) as g:
        assert (g.row_kernel, g.col_kernel) == ("rowx1024r4", "colx3840")
    with owl.Gpu(E, col_one_region=True) as g:
        assert (g.row_kernel, g.col_kernel) == ("rowx1024r4", "coly3840")
    with owl.Gpu(E, generic=True) as g:
        assert re.fullmatch(r"row\d+", g.row_kernel) and re.fullmatch(r"col\d+", g.col_kernel)


@pytest.mark.parametrize("spec,bpw,variant", [("1K:3:256", 12.0, "default"), ("1K:3:256", 12.0, "coly"),
                                              ("1K:15:256", 17.6, "default"), ("1K:15:256", 17.6, "coly"),
                                              ("1K:15:256", 17.6, "generic"), ("1K:15:256", 17.6, "onepair")])
def test_mul_pass(spec, bpw, variant):
    # COL_FWD on a second operand, then COL_MUL, against the model's spectrum product
    g, m, rng = setup(spec, bpw, variant, seed=4)
    with g:
        a, b = m.random_digits(rng), m.random_digits(rng)
        g.write_digits(owl.AUX, b.tolist())
        g.write_digits(owl.DATA, a.tolist())
        g.run_pass(g.PASS_LEADIN, owl.AUX)
        fb = g.read_freq()
        g.write_freq(fb, 2)
        g.run_pass(g.PASS_SPECTRUM)
        spec_gpu = g.read_freq(2)
        assert rel(spec_gpu, m.spectrum_layout(fb)) < 1e-12
        g.run_pass(g.PASS_LEADIN, owl.DATA)
        fa = g.read_freq()
        g.run_pass(g.PASS_MUL)
        assert rel(g.read_freq(), m.col_square(fa, mul_spec=m.spectrum(fb))) < 1e-12


# Round-off pin: the state is set to H times the forward row transform of digits d + delta*e_k
# (small digits, so the FFTs are accurate to ~1e-12); the row pass then sees word k at d_k + delta
# and every other word at an integer, so the round-off it reports must be |delta| -- whatever
# block, column and word k is in (the redo of block 0, the last word of a block, ...).
@pytest.mark.parametrize("spec,variant", [("256:2:256", "default"), ("1K:3:256", "default"),
                                          ("1K:15:256", "default"), ("1K:15:256", "generic")])
def test_round_off_is_reported_exactly(spec, variant):
    g, m, rng = setup(spec, 8.0, variant, seed=7)
    with g:
        d = m.random_digits(rng).astype(np.float64)
        N = m.N
        for k, delta in [(0, 0.3125), (N - 1, -0.4375), (2 * g.R * 7 + 1, 0.28125), (N // 2 + 5, -0.15625)]:
            for pas in (g.PASS_FUSED, g.PASS_CARRYOUT):
                y = d.copy()
                y[k] += delta
                g.write_freq(m.H * m.lead_in(y))
                g.read_roe()  # clears the slots
                g.run_pass(pas, owl.AUX if pas == g.PASS_CARRYOUT else 0)
                roe = g.read_roe()
                assert roe["n"] == 1 and abs(roe["max"] - abs(delta)) < 1e-6, (k, delta, pas, roe)
        # an integer state reports (almost) no round-off
        g.write_freq(m.H * m.lead_in(d))
        g.read_roe()
        g.run_pass(g.PASS_FUSED)
        assert g.read_roe()["max"] < 1e-6
