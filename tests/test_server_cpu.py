"""Server request checks that need no GPU: a request whose footprint alone exceeds the memory
budget is rejected at submit (it could never be admitted, and would otherwise wait forever),
as are requests no FFT or engine covers; fault arming checks its arguments."""
import pytest

import gpuowl_amd as owl


def test_request_larger_than_the_budget_is_rejected():
    with owl.Server(mem_budget=1 << 20, block_size=100, check_step=200) as s:
        with pytest.raises(owl.OwlError, match="budget"):
            s.submit(4_000_037, "prp", 1000)
        assert s.info()["waiting"] == 0


def test_invalid_requests_are_rejected():
    with owl.Server(mem_budget=1 << 40, block_size=100, check_step=200) as s:
        with pytest.raises(owl.OwlError):
            s.submit(4_000_037, "prp", 4_000_038)  # past E
        with pytest.raises(owl.OwlError):
            s.submit(4_000_037, "ll", 4_000_036)   # past E - 2
        with pytest.raises(owl.OwlError):
            s.submit(4_000_037, "prp", 0)
        rid = s.submit(4_000_037, "prp", 1000)
        lid = s.submit(4_000_037, "ll", 1000)
        with pytest.raises(owl.OwlError):
            s.arm_fault(lid, 10)      # LL has no error check
        with pytest.raises(owl.OwlError):
            s.arm_fault(rid, 1000)    # not inside the request
        with pytest.raises(owl.OwlError):
            s.arm_fault(rid + 100, 10)
        s.arm_fault(rid, 10)
        assert s.info()["waiting"] == 2
