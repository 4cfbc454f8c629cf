import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "italy-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    import ib_oracle

    return ib_oracle


@pytest.fixture(scope="session")
def meshgen():
    from italy_bake import meshgen

    return meshgen


@pytest.fixture(scope="session")
def ctx():
    from italy_bake import bake

    c = bake.Context()
    yield c
    c.close()
