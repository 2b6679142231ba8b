import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
DTE_LIB = os.path.join(ROOT, "datamininglibs_amd", "libdte.so")
HOST_TEST = os.path.join(ROOT, "tests", "cpp", "test_host_api")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: larger CPU oracle runs")


def _make(path, target):
    if not os.path.exists(target):
        subprocess.run(["make", "-C", path, "-s", "-j8"], check=True)


@pytest.fixture(scope="session", autouse=True)
def _built_libraries():
    """The oracle (gcc, seconds) is built when missing; libdte.so is built by
    __graft_entry__.build() and only rebuilt here when absent."""
    _make(os.path.join(ROOT, "oracle"), ORACLE_LIB)
    _make(os.path.join(ROOT, "datamininglibs_amd"), DTE_LIB)
    _make(os.path.join(ROOT, "datamininglibs_amd"), HOST_TEST)
    yield
