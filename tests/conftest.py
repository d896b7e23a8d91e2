import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libsav_hip.so")


@pytest.fixture(scope="session")
def jv():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope="session")
def so():
    from oracle import sav_oracle
    return sav_oracle
