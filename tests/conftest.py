import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: long-running GPU test")


def next_prime(n):
    from tests.golden.make_goldens import next_prime as np_
    return np_(n)


@pytest.fixture(scope="session")
def golden():
    """Loads tests/golden/res_<name>.txt -> (header dict, {k: res64})."""
    def load(name):
        path = os.path.join(ROOT, "tests", "golden", f"res_{name}.txt")
        if not os.path.exists(path):
            pytest.skip(f"golden fixture {name} not generated")
        head, res = {}, {}
        for line in open(path):
            if line.startswith("#"):
                for kv in line[1:].split():
                    k, v = kv.split("=")
                    head[k] = int(v) if v.isdigit() else v
            elif line.strip():
                k, r = line.split()
                res[int(k)] = int(r, 16)
        return head, res
    return load
