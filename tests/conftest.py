import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for path in (ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(os.path.abspath(__file__))):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")


def pytest_collection_modifyitems(config, items):
    if os.path.exists("/dev/kfd"):
        return
    skip = pytest.mark.skip(reason="no AMD GPU (/dev/kfd) on this machine")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session", autouse=True)
def _torch_gpu_context_first(request):
    """torch ships its own HIP runtime; when the library's runtime initialises the GPU
    first, torch's later initialisation finds no GPU. Tests that use torch device
    tensors therefore get torch's GPU context set up before any library call."""
    if os.path.exists("/dev/kfd") and any("gpu" in item.keywords for item in request.session.items):
        import torch

        if torch.cuda.is_available():
            torch.cuda.init()
    yield
