import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

import __graft_entry__  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def mpm():
    """The product package (loads libmpm_hip.so; raises if it is not built)."""
    m = __graft_entry__.load_package()
    m.load_library()
    return m


@pytest.fixture(scope="session")
def orc():
    """The CPU oracle (checker)."""
    from oracle import oracle
    oracle.lib()
    return oracle
