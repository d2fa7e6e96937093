"""owl_gpu_read_digits_noalloc (staging through the library's static device array) against
owl_gpu_read_digits (a staging buffer allocated per call), through the C-ABI: two calls in a
row, the second on a smaller FFT, and the rejection of an FFT larger than the static array."""
import pytest

import gpuowl_amd as owl
from gpuowl_amd import _lib

pytestmark = pytest.mark.gpu

MAX_WORDS = 1 << 23  # OWL_NOALLOC_MAX_WORDS (job.h)


def _run(E, spec, iters):
    g = owl.Gpu(E, spec)
    try:
        g.write_int(owl.DATA, 3)
        g.square_loop(owl.DATA, owl.DATA, 0, iters)
        want = g.read_digits()
        got = g.read_digits_noalloc()
        return g.N, want, got, g.res64()
    finally:
        g.close()


def test_noalloc_readout_matches_allocating_readout():
    n1, want1, got1, _ = _run(139_000_003, "1K:15:256", 3)
    assert n1 == 7_864_320 and n1 <= MAX_WORDS
    assert got1 == want1
    assert any(want1)
    # the second call is smaller: only its N words are staged and read back
    n2, want2, got2, _ = _run(4_000_037, "256:2:256", 5)
    assert n2 == 262_144
    assert got2 == want2
    assert any(want2)


def test_noalloc_readout_rejects_ffts_above_the_static_array():
    g = owl.Gpu(200_000_033, "1K:16:512")
    try:
        assert g.N > MAX_WORDS
        with pytest.raises(_lib.OwlError, match="static staging"):
            g.read_digits_noalloc()
    finally:
        g.close()
