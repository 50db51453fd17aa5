import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library's C-ABI)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib("strict")
    return O


@pytest.fixture(scope="session")
def ctx64():
    from fdps_sph_amd import Context
    c = Context(0, "fp64")
    from fdps_sph_amd import eos_tables as E
    c.load_table(0, E.PATHS["granite"])
    c.load_table(1, E.PATHS["iron"])
    c.load_table(2, E.PATHS["dunite"])
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctxmix():
    from fdps_sph_amd import Context
    c = Context(0, "mixed")
    from fdps_sph_amd import eos_tables as E
    c.load_table(0, E.PATHS["granite"])
    c.load_table(1, E.PATHS["iron"])
    c.load_table(2, E.PATHS["dunite"])
    yield c
    c.close()
