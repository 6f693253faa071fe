import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _init_torch_gpu():
    """torch ships its own HIP runtime beside the one libnckde links: it must
    initialise first in a process that uses both (tests that hand torch device
    buffers to the library).  On a machine without a GPU this fails quietly."""
    try:
        import torch
        torch.cuda.init()
    except Exception:
        pass


def pytest_configure(config):
    _init_torch_gpu()
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
