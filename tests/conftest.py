import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PKG = "protein_encoding_v2_nsga-iii_amd"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu on a machine that has one)")


@pytest.fixture(scope="session")
def cdp():
    return importlib.import_module(PKG)


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ref_tables(cdp):
    return cdp.tables.reference_tables()
