"""Kernel-level numerics: each HIP pass against the numpy FP64 model of the same op
(tests/fftmodel.py) on dense random digits. FFT outputs must agree to a relative
1e-12 (FP64 with correctly rounded twiddles); rounded digits must agree exactly."""
import numpy as np
import pytest

import gpuowl_amd as owl
from gpuowl_amd import workloads
from tests.fftmodel import Model

pytestmark = pytest.mark.gpu

SHAPES = [("64:2:64", 12.0), ("256:2:256", 15.0), ("256:3:256", 8.0), ("1K:3:256", 8.0), ("256:13:256", 8.0),
          ("512:9:256", 8.0)]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def setup(spec, bpw, seed=1):
    N = owl.best_fit(1000, spec)[1]
    E = workloads.next_prime(int(bpw * N))
    g = owl.Gpu(E, spec)
    m = Model(E, g.W, g.H, g.R)
    return g, m, np.random.default_rng(seed)


@pytest.mark.parametrize("spec,bpw", SHAPES)
def test_lead_in_and_square(spec, bpw):
    g, m, rng = setup(spec, bpw)
    with g:
        d = m.random_digits(rng)
        g.write_digits(owl.DATA, d.tolist())
        g.run_pass(g.PASS_LEADIN, owl.DATA)
        f = g.read_freq()
        assert rel(f, m.lead_in(d)) < 1e-12
        g.run_pass(g.PASS_SQUARE)
        sq = g.read_freq()
        assert rel(sq, m.col_square(f)) < 1e-12


@pytest.mark.parametrize("spec,bpw", SHAPES)
def test_fused_row_pass(spec, bpw):
    g, m, rng = setup(spec, bpw, seed=2)
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


@pytest.mark.parametrize("spec,bpw", SHAPES)
def test_carry_out_digits(spec, bpw):
    g, m, rng = setup(spec, bpw, seed=3)
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
