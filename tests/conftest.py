import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtwopcf_hip)")
    config.addinivalue_line("markers", "slow: long CPU test")


def parse_params(text):
    p = {}
    for line in str(text).splitlines():
        if "=" in line and not line.startswith("#"):
            k, v = line.split("#")[0].split("=", 1)
            p[k.strip()] = v.strip()
    return p


def load_golden(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    oracle.lib()
    return oracle
