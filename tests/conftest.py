"""Test configuration: the `gpu` marker, repo-root imports, built libraries.

Tests marked `gpu` need a HIP device and run the product path (the C ABI of
libblobs_gpu.so).  The rest run on the CPU: the oracle against closed forms,
invariants, brute force and the reference's one golden vector on this path,
the host-side construction logic, and the ABI's exported symbols.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the HIP path through the C ABI")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session", autouse=True)
def _built_libraries():
    """Build the oracle (test infrastructure) if missing; the product
    libraries must already be built (`python -m blobs_amd.build`)."""
    from oracle import oracle
    if not oracle.LIB.exists():
        oracle.build()
    yield


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (device enumeration only)
        return torch.cuda.device_count() > 0
    except Exception:
        return False
