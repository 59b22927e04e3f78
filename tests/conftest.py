import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def pytest_runtest_logstart(nodeid, location):
    # name the case before any GPU work, so a lost box still says which one ran
    print(f"\n[case] {nodeid}", flush=True)


@pytest.fixture(scope="session")
def cb():
    import cosmicblup_amd

    return cosmicblup_amd


@pytest.fixture(scope="session")
def ctx(cb):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return cb._lib.default_context()
