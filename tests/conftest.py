"""pytest configuration: the `gpu` marker and device fixtures.

Tests marked `gpu` need an MI355X and the built libreikna_amd.so; everything
else runs on the CPU (oracle, planner rules, API semantics, library symbols).
"""

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def device():
    from reikna_amd.runtime import Device, device_count

    if device_count() == 0:
        pytest.fail("no GPU visible: gpu tests must run on the GPU box")
    return Device(0)


@pytest.fixture(scope="session")
def queue(device):
    from reikna_amd.runtime import Queue

    return Queue(device)


@pytest.fixture
def some_queue(queue):
    return queue
