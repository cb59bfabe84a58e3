import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP libraries")
    config.addinivalue_line("markers", "slow: longer CPU checks")


_ENGINES = {}


@pytest.fixture(scope="session")
def engine_factory():
    from cumayo_amd import Engine

    def get(version):
        if version not in _ENGINES:
            _ENGINES[version] = Engine(version)
        return _ENGINES[version]

    return get
