import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_kats.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: larger CPU-side oracle cases")


def pytest_runtest_logstart(nodeid, location):
    # Name the case before any GPU work starts, unbuffered, so a run that ends
    # the process still says which test was running.
    sys.__stderr__.write(f"[case] {nodeid}\n")
    sys.__stderr__.flush()


@pytest.fixture(scope="session")
def kats():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu_ctx():
    import cuyashe_amd as ya
    ctx = ya.Context(0)
    yield ctx
    try:
        ctx.close()
    except ya.YashError as e:
        # a failed test left rings or plans alive: keep the context (the
        # process ends next) rather than free what they still reference
        sys.__stderr__.write(f"[gpu_ctx] context kept alive: {e}\n")
