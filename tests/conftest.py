import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librtg.so on cuda:0 / HIP device 0)")


@pytest.fixture(scope="session")
def ctx():
    from _pkg import rtg
    c = rtg.Context(0)
    print("device:", c.device_info(), flush=True)
    yield c
    c.close()
