import os
import sys

import pytest

# torch, where a test uses it beside the library, has to be loaded first: it carries its own HIP
# runtime under the same soname as the one libjkg.so links, and a process can hold only one
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; parity tests of the HIP path")


def golden(name):
    import numpy as np
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o
    o.lib()
    return o


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
