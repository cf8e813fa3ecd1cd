"""Shared fixtures: the ``gpu`` marker, the package loader and the CUDA device.

``gpu`` tests run only on a box with a GPU (``pytest -m gpu``); everything else runs on the CPU and
never calls a kernel.
"""

from __future__ import annotations

import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: tens of seconds of CPU oracle work")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__

    return __graft_entry__.load_package()


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("a gpu-marked test ran without a GPU")
    return torch.device("cuda:0")
