import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@pytest.fixture(scope="session")
def params_fs2fs3():
    from ccpol23_plus_2025_mbmd_amd.params import load_variant
    return load_variant("fs2fs3")


@pytest.fixture(scope="session")
def params_fs3():
    from ccpol23_plus_2025_mbmd_amd.params import load_variant
    return load_variant("fs3")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
