"""The reference's tune.txt on the CPU (src/TuneEntry.cpp:7-84, src/FFTConfig.cpp:235-263):
bestFit takes the cheapest listed FFT that covers the exponent, the list stays the
cost/maxExp Pareto frontier of what was added (checked against a brute-force frontier),
and the file is written in the reference's "%6.1f %14s # %u" lines."""
import random

import pytest

import gpuowl_amd as owl

SPECS = ["256:2:256:101", "256:3:256:101", "256:4:256:101", "512:3:256:101", "1K:3:256:101", "1K:5:256:101",
         "1K:9:256:101", "1K:15:256:101"]


def max_exp(spec):
    return owl.best_fit(1000, spec)[2]


@pytest.fixture(autouse=True)
def clean():
    owl.fft_tune_clear()
    yield
    owl.fft_tune_clear()


def test_best_fit_takes_the_cheapest_covering_entry(tmp_path):
    path = tmp_path / "tune.txt"
    path.write_text("  22.8  256:2:256:101 # 5124128\n"
                    "# a comment line\n"
                    "  30.1  256:4:256:101\n"
                    "  41.0  1K:3:256:101 # ignored maxExp comment\n")
    owl.fft_tune_load(path)
    m2, m4, m3k = max_exp("256:2:256:101"), max_exp("256:4:256:101"), max_exp("1K:3:256:101")
    assert [s for _, s, _ in owl.fft_tune_list()] == ["256:2:256:101", "256:4:256:101", "1K:3:256:101"]
    assert owl.best_fit(m2)[0] == "256:2:256:101"
    assert owl.best_fit(m2 + 1)[0] == "256:4:256:101"      # 256:3:256 is not listed
    assert owl.best_fit(m4 + 1)[0] == "1K:3:256:101"
    spec, _, mx = owl.best_fit(m3k + 1)                       # past every entry: the shape table
    assert spec not in ("256:2:256:101", "256:4:256:101", "1K:3:256:101") and mx > m3k
    owl.fft_tune_clear()
    # the shape table: variants 101, 202 per shape (the reference's limits); this engine caps both
    # variants of 256:2:256 at the same bits per word, so it moves on to the next shape
    r2 = owl.best_fit_ref(1000, "256:2:256:101")[2]
    assert owl.best_fit_ref(r2 + 1)[0] == "256:2:256:202"
    assert owl.best_fit(m2 + 1)[0] == "256:3:256:101"


def test_unordered_file_is_an_error(tmp_path):
    path = tmp_path / "tune.txt"
    path.write_text("30.0 256:4:256:101\n20.0 256:2:256:101\n")
    with pytest.raises(owl.OwlError):
        owl.fft_tune_load(path)
    path.write_text("20.0 256:4:256:101\n30.0 256:2:256:101\n")   # costlier but covers less
    with pytest.raises(owl.OwlError):
        owl.fft_tune_load(path)


def test_update_rules_by_hand():
    assert owl.fft_tune_add(10.0, "256:3:256:101")
    assert owl.fft_tune_add(20.0, "1K:3:256:101")
    assert not owl.fft_tune_add(15.0, "256:2:256:101")    # covers less than the 10.0 entry
    assert owl.fft_tune_add(5.0, "256:4:256:101")          # cheaper and covers more: replaces 10.0
    assert [(c, s) for c, s, _ in owl.fft_tune_list()] == [(5.0, "256:4:256:101"), (20.0, "1K:3:256:101")]


def frontier(entries):
    """Entries not dominated by an earlier-or-equal-cost entry covering at least as much."""
    keep = [e for e in entries if not any(f is not e and f[0] <= e[0] and f[1] >= e[1] for f in entries)]
    return sorted(keep)


def test_update_keeps_the_pareto_frontier():
    rng = random.Random(11)
    for trial in range(20):
        owl.fft_tune_clear()
        costs = rng.sample(range(10, 500), len(SPECS))
        added = []
        for spec, cost in zip(rng.sample(SPECS, len(SPECS)), costs):
            owl.fft_tune_add(float(cost), spec)
            added.append((float(cost), max_exp(spec)))
        got = [(c, m) for c, _, m in owl.fft_tune_list()]
        assert got == frontier(added), trial


def test_saved_in_the_reference_layout(tmp_path):
    for cost, spec in [(22.84, "256:2:256:101"), (41.0, "1K:3:256:101"), (143.55, "1K:15:256:101")]:
        owl.fft_tune_add(cost, spec)
    path = tmp_path / "tune.txt"
    owl.fft_tune_save(path)
    want = "".join("%6.1f %14s # %u\n" % (c, s, max_exp(s))
                   for c, s in [(22.84, "256:2:256:101"), (41.0, "1K:3:256:101"), (143.55, "1K:15:256:101")])
    assert path.read_text() == want
    owl.fft_tune_clear()
    owl.fft_tune_load(path)
    assert [s for _, s, _ in owl.fft_tune_list()] == ["256:2:256:101", "1K:3:256:101", "1K:15:256:101"]
