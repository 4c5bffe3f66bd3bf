"""Shared test set-up: import paths, the `gpu` marker and session-wide handles."""
from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "photon-engine_amd" / "python"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and calls libphoton_hip.so kernels")


@pytest.fixture(scope="session")
def dev():
    from photon_hip import Device, device_count

    assert device_count() > 0, "no HIP device visible"
    # Scenes and films keep a reference to their device; it is released when the last of them is.
    yield Device(0)


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    g = ROOT / "tests" / "golden"
    return {"sobol": dict(np.load(g / "sobol_ref.npz")), "hits": dict(np.load(g / "hits_ref.npz"))}
