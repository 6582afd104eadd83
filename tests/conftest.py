import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import ORACLE_SO, Oracle
    if not os.path.exists(ORACLE_SO):
        pytest.fail("oracle/liboracle.so is not built (run __graft_entry__.build())")
    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle_lib import REF_SO, Reference
    if not os.path.exists(REF_SO):
        pytest.skip("the reference build (oracle/_ref) is not present")
    return Reference()
