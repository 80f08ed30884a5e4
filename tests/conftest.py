import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs libkronmult_hip.so on an MI355X (needs a GPU)")


@pytest.fixture(scope="session")
def gpu():
    """The GPU tests fail (never skip) without a device: they are the parity proof."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from kronmult993_amd import _lib
    _lib.lib()  # raises if libkronmult_hip.so is not built
    dev = torch.device("cuda:0")
    print(f"[gpu] {torch.cuda.get_device_name(dev)}", flush=True)
    return dev
