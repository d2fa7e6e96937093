"""Every compiled shape-specialized pass, forced through the tune table onto a geometry
it serves, gives GMP's PRP residues (with Gerbicz checks) over a short run; the multi-pair
column passes run in a batch large enough to choose them."""
import re
import sys

import pytest

import gpuowl_amd as owl
import oracle
from gpuowl_amd import workloads

pytestmark = pytest.mark.gpu

NAMES = [n for n in owl.kernel_names() if re.match(r"(rowx|colx|coly)\d", n)]


def announce(*a):
    print("CASE", *a, file=sys.stderr, flush=True)


def reference_spec(nd):
    """A reference FFT spec (width:middle:height) with width*middle*height == nd."""
    fmt = {256: "256", 512: "512", 1024: "1K", 4096: "4K"}
    for a in (256, 512, 1024, 4096):
        for c in (256, 512, 1024):
            if nd % (a * c) == 0 and 1 <= nd // (a * c) <= 16:
                return f"{fmt[a]}:{nd // (a * c)}:{fmt[c]}"
    return None


def config_for(name):
    """(spec, W, R, flags, row kernel, col kernel prefix) that selects kernel `name`."""
    m = re.match(r"rowx(\d+)r(\d+)", name)
    if m:
        W, R = int(m.group(1)), int(m.group(2))
        H = {2048: 256, 4096: 128}.get(W, 512)
        return reference_spec(W * H), W, R, 0, name, None
    kind, H, multi = re.match(r"(col[xy])(\d+)(m)?$", name).groups()
    H = int(H)
    if multi:  # W = 1024 has the most column pairs per job to share out
        return reference_spec(1024 * H), 1024, 2, 0, None, name
    for W in (256, 1024, 512):
        spec = reference_spec(W * H)
        if spec:
            return spec, W, 2, 4 if kind == "coly" else 0, None, name
    return None


@pytest.mark.parametrize("name", NAMES)
def test_specialized_kernel_is_exact(name):
    cfg = config_for(name)
    assert cfg and cfg[0], f"no reference shape serves {name}"
    spec, W, R, flags, row, col = cfg
    if name.endswith("m"):
        # the multi-pair column pass (compiled for H >= 2304, one group per CU) runs where a
        # launch has groups to spare: one job already gives W/4 = 256 groups of two pairs
        N = owl.best_fit(1000, spec)[1]
        exps = [workloads.next_prime(15 * N)]
        announce(name, spec, W, R, exps[0], "x", len(exps))
        owl.tune_set(N, W, R, flags)
        try:
            with owl.Batch(exps, spec) as b:
                assert b.job_info(0)["col"] == name
                outs, _ = b.prp(60, block_size=30, check_step=30)
            for E, o in zip(exps, outs):
                assert o.res64 == oracle.gmp_res64(E, 60, 60)[0], E
                assert o.nErrors == 0 and o.maxRoe < 0.4
        finally:
            owl.tune_clear()
        return
    N = owl.best_fit(1000, spec)[1]
    E = workloads.next_prime(15 * N)
    announce(name, spec, W, R, flags, E)
    owl.tune_set(N, W, R, flags)
    try:
        with owl.Gpu(E, spec) as g:
            assert (g.W, g.R) == (W, R)
            if row:
                assert g.row_kernel == row
            if col:
                assert g.col_kernel == col
            out, got = g.prp(60, block_size=30, check_step=30, res_step=30)
        assert got == oracle.gmp_res64(E, 60, 30)
        assert out.nErrors == 0 and out.maxRoe < 0.4
    finally:
        owl.tune_clear()
