import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); run with -m gpu")


def _ensure_built():
    import importlib.util
    spec = importlib.util.spec_from_file_location("gpu_mm_amd_build", os.path.join(REPO, "gpu_mm_amd", "_build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    if not (os.path.exists(b.lib_path()) and os.path.exists(b.ext_path())):
        b.build()
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()


_ensure_built()


def gpu_available():
    import torch
    return torch.cuda.is_available()


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def gmm():
    import gpu_mm_amd
    return gpu_mm_amd


@pytest.fixture(scope="session")
def dev():
    import torch
    return torch.device("cuda", 0)
