import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def release():
    import json
    import numpy as np
    here = os.path.join(ROOT, "tests", "golden")
    z = np.load(os.path.join(here, "couette_release.npz"), allow_pickle=False)
    with open(os.path.join(here, "couette_release.json")) as f:
        g = json.load(f)
    return dict(u=z["u"], v=z["v"], w=z["w"], meta=g)
