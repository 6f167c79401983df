import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a visible GPU and the built lib21cm_hip.so")


def pkg(name=""):
    """Import a module of the 21cmpy_amd package (its name starts with a digit)."""
    return importlib.import_module("21cmpy_amd" + ("." + name if name else ""))


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    lib = pkg("_lib")
    lib.load()
    return lib
