import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return load


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        oracle.build(ref=False)
    return oracle.Oracle()


@pytest.fixture(scope="session")
def gpu():
    """The HIP library with a visible device.  A GPU test on a box without a
    visible device fails rather than skipping: there is no CPU fallback."""
    import erlbm_amd
    n = erlbm_amd.device_count()
    if n < 1:
        pytest.fail("no HIP device visible to liberlbm.so")
    return erlbm_amd


def published_maps():
    d = np.load(os.path.join(GOLDEN, "published_maps.npz"), allow_pickle=False)
    return {k: np.unpackbits(d[k])[: 400 * 400].reshape(400, 400) for k in d.files}


def read_csv(name):
    import csv
    with open(os.path.join(GOLDEN, name)) as fh:
        return list(csv.DictReader(fh))
