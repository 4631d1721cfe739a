import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def lib():
    import bounce_amd
    return bounce_amd.load_library()


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    return oracle_lib.oracle()


@pytest.fixture(scope="session")
def reference():
    import oracle_lib
    r = oracle_lib.reference()
    if r is None:
        pytest.skip("oracle/_ref (the compiled reference) is not built here")
    return r


@pytest.fixture(scope="session")
def ctx():
    import bounce_amd
    c = bounce_amd.Context(0)
    yield c
    c.close()
