import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def hip():
    """The HIP library and a visible GPU; the product path has no fallback."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("gpu test collected without a visible HIP device")
    import swspec_amd
    from swspec_amd import _lib
    _lib.load()
    return swspec_amd
