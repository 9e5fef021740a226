"""Shared test setup: the `gpu` marker and import paths.

GPU tests run only with ``-m gpu`` on a machine with an MI355X; they fail
(never skip) when the GPU or libssim is missing, so a CPU fallback can never
pass them.
"""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X and the built libssim.so")


@pytest.fixture(scope="session")
def dev():
    """The torch device of GPU 0 (raises if there is none)."""
    from solarsystemsimulator_amd.device import require_gpu
    return require_gpu(0)
