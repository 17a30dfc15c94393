"""Shared pytest configuration: markers, import paths and oracle build."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "cat-surface-gpu_amd"
for path in (str(ROOT), str(PKG)):
    if path not in sys.path:
        sys.path.insert(0, path)

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libcsg_hip.so")


def _ensure_oracle() -> None:
    lib = ROOT / "oracle" / "liboracle.so"
    if not lib.is_file():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


_ensure_oracle()


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    cache = {}

    def load(name: str):
        if name not in cache:
            cache[name] = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
        return cache[name]

    return load


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")
    return torch.device("cuda", 0)
