import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def anchors():
    with open(os.path.join(GOLDEN, "reference_anchors.json"), encoding="utf-8") as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def closed_forms():
    with open(os.path.join(GOLDEN, "closed_forms.json"), encoding="utf-8") as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    oracle.lib()
    return oracle
