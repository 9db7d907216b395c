import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libdummest_amd.so")


@pytest.fixture(scope="session")
def engine():
    import dummest_amd as D
    e = D.Engine(0)
    yield e
    e.close()


@pytest.fixture(autouse=True)
def _calibration_cache_dir(tmp_path_factory, monkeypatch):
    # The engine executable reads and writes a calibration cache when it ends;
    # every test gets a fresh cache directory of its own.
    monkeypatch.setenv("DUMMEST_CACHE_DIR", str(tmp_path_factory.mktemp("cache")))
