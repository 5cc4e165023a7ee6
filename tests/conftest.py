import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(REPO, "tests", "golden", "cases.json")) as f:
        return json.load(f)["cases"]


@pytest.fixture(scope="session")
def ref_cones():
    """Light cones and per-edge F_p texts written by the reference's own ExtraData and
    emission functions (src/maxcut.h:31-233), recorded by tests/golden/make_golden.py."""
    import json
    with open(os.path.join(REPO, "tests", "golden", "cones.json")) as f:
        return json.load(f)["graphs"]
