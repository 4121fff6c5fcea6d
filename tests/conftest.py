import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "mojo-petsc4py_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libmpk on the device)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    O.set_mode("ref")
    return O


@pytest.fixture(scope="session")
def native():
    """libmpk with the device opened; only for tests marked gpu."""
    from mpk import _native
    return _native.lib()
