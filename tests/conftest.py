import os
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
for p in (REPO, REPO / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def orc():
    import pyoracle

    return pyoracle


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library; GPU tests must fail loudly (not skip) when it is missing."""
    from miluphpc_amd import lib

    return lib


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor)
    den = np.where(den == 0, 1.0, den)
    return np.abs(a - b) / den


def vec_rel_err(ax, ay, az, bx, by, bz, floor=0.0):
    num = np.sqrt((ax - bx) ** 2 + (ay - by) ** 2 + (az - bz) ** 2)
    den = np.sqrt(bx ** 2 + by ** 2 + bz ** 2)
    den = np.maximum(den, floor)
    return num / np.where(den == 0, 1.0, den)


def print_case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)
