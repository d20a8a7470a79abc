import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """Initialises the HIP backend on device 0 (fails the test when absent).

    torch is imported first when available: it carries its own HIP runtime,
    and the backend library then binds to that same runtime, so device
    pointers of torch tensors (the autograd harness) and of the backend are
    interchangeable.  Loaded the other way round, two runtimes would coexist
    and torch would see no device."""
    try:
        import torch

        torch.cuda.is_available()
    except ImportError:
        pass
    from simall_amd import capi

    lib = capi.load()
    capi.check(lib.mi_init(0))
    return lib
