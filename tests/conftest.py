import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle_built():
    import oracle_lib
    if not os.path.exists(oracle_lib.ORACLE_SO):
        pytest.fail("oracle/liboracle.so is missing: run `make -C oracle`")
    return oracle_lib
