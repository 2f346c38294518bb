import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
ORACLE_LIB = os.path.join(ROOT, "oracle", "_build", "libfs_oracle.so")
HIP_LIB = os.path.join(ROOT, "fluidprjct_amd", "libfluidsim_hip.so")
REF_INPUT = os.path.join(ROOT, "tests", "golden", "inputdata")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from fluidprjct_amd._abi import FsLib
    if not os.path.exists(ORACLE_LIB):
        pytest.fail("oracle not built: make -C oracle")
    return FsLib(ORACLE_LIB)


@pytest.fixture(scope="session")
def hip():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected without a GPU")
    from fluidprjct_amd import load_library
    return load_library()
