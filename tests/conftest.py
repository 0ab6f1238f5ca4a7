import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p in (HERE, os.path.join(REPO, "arbor-cuda_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle, ORACLE_SO

    if not os.path.exists(ORACLE_SO):
        pytest.fail("oracle/liboracle.so is not built: run `make -C oracle` (or __graft_entry__.build())")
    return Oracle()


@pytest.fixture(scope="session")
def golden():
    import json

    import numpy as np

    with open(os.path.join(HERE, "golden", "golden.json")) as f:
        g = json.load(f)
    crops = dict(np.load(os.path.join(HERE, "golden", "crops.npz")))
    return g, crops
