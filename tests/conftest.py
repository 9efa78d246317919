import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def net27():
    import mumfim_amd as mf
    return mf.load_network(os.path.join(GOLDEN, "network_2_7", "Network_2_7.txt"))


def voigt(S):
    return [S[0][0], S[1][1], S[2][2], S[1][2], S[0][2], S[0][1]]
