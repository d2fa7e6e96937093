"""Host-side checks on the CPU: the C-ABI library loads and exports every function
declared in include/*.h (no GPU call is made); FFT selection matches the
reference's documented choices; the workload generators match their definition."""
import os
import re

import pytest

import gpuowl_amd
from gpuowl_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(owl_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = gpuowl_amd.load()
    names = declared_functions()
    assert len(names) >= 40
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_no_fallback_when_library_missing(monkeypatch, tmp_path):
    from gpuowl_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(_lib.OwlError):
        _lib.load()


@pytest.mark.parametrize("E,spec,N", [
    (4_000_037, "256:2:256:101", 262144),        # p = 4,000,037: the reference picks 256:2:256
    (139_000_003, "1K:15:256:101", 7864320),      # p = 139,000,003: 1K:15:256, maxExp ~142.56M
    (1_029_689, "256:2:256:101", 262144),
    (56_430_277, "1K:6:256:101", 3145728),
])
def test_best_fit_matches_reference_choice(E, spec, N):
    for fit in (gpuowl_amd.best_fit_ref, gpuowl_amd.best_fit):  # the engine's limits agree here
        s, n, mx = fit(E)
        assert (s, n) == (spec, N)
        assert mx >= E


def test_best_fit_p139000003_max_exp():
    for fit in (gpuowl_amd.best_fit_ref, gpuowl_amd.best_fit):
        s, n, mx = fit(139_000_003)
        assert abs(mx - 142_560_000) < 10_000


def test_best_fit_uneven200_distribution():
    # the reference's choices for config D (its own limits)
    ps = workloads.uneven200()
    assert ps[:5] == [19_917_551, 11_330_491, 2_287_861, 3_503_053, 19_297_693]
    assert min(ps) == 1_029_689 and max(ps) == 56_430_277
    Ns = [gpuowl_amd.best_fit_ref(p)[1] for p in ps]
    assert Ns.count(262144) == 83 and Ns.count(393216) == 22
    assert sum(Ns) == 155_713_536
    assert sum(1 for p, n in zip(ps, Ns) if p / n < 12) == 58


def test_engine_limits_cap_only_the_measured_shapes():
    # The engine's limit is the reference's, lowered only for the shapes in engine_bpw.inc (their
    # top exponents measured above the Z = 28 round-off criterion on this engine's kernels); of
    # config D's 200 exponents, the five in those top ranges move to another FFT (of the same or the
    # next size).
    ps = workloads.uneven200()
    moved = {p: (gpuowl_amd.best_fit_ref(p), gpuowl_amd.best_fit(p)) for p in ps}
    moved = {p: (r, e) for p, (r, e) in moved.items() if r[:2] != e[:2]}
    assert sorted(moved) == [5_124_299, 19_917_551, 24_741_121, 48_995_393, 49_039_027]
    for p, (ref, eng) in moved.items():
        assert eng[1] >= ref[1] and eng[0] != ref[0] and eng[2] >= p and ref[2] >= p
        # the reference's FFT is still the choice with its limits, and p is above this engine's
        assert gpuowl_amd.best_fit(p, ref[0])[2] < p
    # the 256:2:256 cap: 19.442 bits per word (the reference's 101/202 limits are 19.547/19.636)
    s, n, mx = gpuowl_amd.best_fit(5_000_000)
    assert s == "256:2:256:101" and mx == int(19.442 * 262144)
    assert gpuowl_amd.best_fit(5_110_000)[0] == "256:3:256:101"
    assert gpuowl_amd.best_fit_ref(5_110_000)[0] == "256:2:256:101"


def test_spec_parsing_and_errors():
    assert gpuowl_amd.best_fit(1000, "1K:3:256")[1] == 1572864
    assert gpuowl_amd.best_fit(1000, "1.5M")[1] == 1572864
    assert gpuowl_amd.best_fit(1000, "256:2:256:202:1")[0] == "256:2:256:202:1"
    for bad in ("1K:3", "1K:3:256:999", "1K:3:256:101:7", "0:1:1"):
        with pytest.raises(gpuowl_amd.OwlError):
            gpuowl_amd.best_fit(1000, bad)


def test_workloads():
    assert workloads.p4m() == 4_000_037
    assert workloads.p139m() == 139_000_003
    c = workloads.uniform64_20m()
    assert len(c) == 64 and c[0] == 20_000_003 and c[-1] == 20_000_879


def test_server_rejects_bad_requests_before_running():
    # no device work happens before the first step, so this runs without a GPU
    with gpuowl_amd.Server(mem_budget=1 << 30) as s:
        with pytest.raises(gpuowl_amd.OwlError):
            s.submit(4_000_037, "prp", 0)
        with pytest.raises(gpuowl_amd.OwlError):
            s.submit(4_000_037, "prp", 4_000_038)
        with pytest.raises(gpuowl_amd.OwlError):
            s.submit(4_000_037, "ll", 4_000_036)
        with pytest.raises(gpuowl_amd.OwlError):
            s.submit(756_839, "prp", 10)   # below 3 bits per word at the smallest FFT
        rid = s.submit(4_000_037, "ll", 4_000_035)
        assert rid == 1 and s.info()["waiting"] == 1
    with pytest.raises(gpuowl_amd.OwlError):
        gpuowl_amd.Server(mem_budget=1 << 30, block_size=100, check_step=150)


def test_tune_table_round_trip(tmp_path):
    owl = gpuowl_amd
    owl.tune_clear()
    assert owl.tune_get(7_864_320) is None
    owl.tune_set(7_864_320, 1024, 4, 0, 141.5)
    owl.tune_set(262_144, 256, 2, 4, 22.8)
    path = tmp_path / "tune.txt"
    owl.tune_save(path)
    owl.tune_clear()
    owl.tune_load(path)
    assert owl.tune_get(7_864_320) == {"width": 1024, "line_block": 4, "flags": 0, "us": 141.5}
    assert owl.tune_get(262_144)["flags"] == 4
    bad = tmp_path / "bad.txt"
    bad.write_text("262144 256 two 0 1.0\n")
    with pytest.raises(owl.OwlError):
        owl.tune_load(bad)
    owl.tune_clear()


def test_kernel_names_cover_the_specialized_passes():
    names = gpuowl_amd.kernel_names()
    assert len(names) == len(set(names))
    for n in ("rowx1024r4", "rowx256r2", "colx3840", "colx3840m", "coly3840", "coly512", "carryB", "readResidue"):
        assert n in names


def test_batch_tuning_rejects_bad_arguments():
    # validated before any device work
    from gpuowl_amd import _lib
    lib = gpuowl_amd.load()
    n = _lib.C.c_uint32()
    assert lib.owl_tune_batch(None, 0, b"", 10, None, 0, _lib.C.byref(n)) != 0
    gpuowl_amd.tune_batch_clear()
