import os
import sys

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU test")


def pytest_collection_finish(session):
    """GPU sessions: bring up PyTorch's device before the library's first
    HIP call.  The library and PyTorch share the process's HIP runtime, and
    PyTorch finds no device when the library initialised it first; the
    learned-form tests hold their tensors in PyTorch."""
    if not any(item.get_closest_marker("gpu") for item in session.items):
        return
    try:
        import torch
        torch.cuda.init()
    except Exception as e:  # no GPU here: the GPU tests fail on their own
        print(f"[conftest] torch.cuda.init() failed: {e}", file=sys.stderr)
