"""Shared fixtures.  `gpu` marks tests that need an MI355X (run with -m gpu)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def pkg():
    return ge.load_package()


@pytest.fixture(scope="session")
def po():
    o = ge.load_oracle()
    o.lib()
    return o


@pytest.fixture(scope="session")
def golden_log():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "vanadis_log_default.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu(pkg):
    lib = pkg.load()
    if pkg.device_count() < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    return lib
