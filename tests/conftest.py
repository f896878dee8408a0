import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library on a GPU)")
    config.addinivalue_line(
        "markers", "slow: long-running (minutes) parity case")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pp_oracle
    pp_oracle.lib()
    return pp_oracle
