import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long-running case")


@pytest.fixture(scope="session")
def workloads(tmp_path_factory):
    """Synthetic FCIDUMPs generated once per session: name -> info dict."""
    from mcdfci_amd import synth

    base = tmp_path_factory.mktemp("fcidump")
    cache = {}

    def get(name):
        if name not in cache:
            cache[name] = synth.generate(name, str(base / name))
        return cache[name]

    return get


@pytest.fixture(scope="session")
def h2o_fcidump():
    return os.path.join(GOLDEN, "h2o_sto3g.FCIDUMP")
