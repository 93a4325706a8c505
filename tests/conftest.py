import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libegt.so")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def egt_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from multigpuegtsmoothing_amd import egt as E

    return E
