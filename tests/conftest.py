"""Shared test setup.

Markers: `gpu` tests need a gfx950 device and call the HIP path through the
C ABI; everything else runs on the CPU. The CPU checkers under oracle/ are
test infrastructure and are imported here only as the checker.
"""

from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "ultrasound-engine_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (runs the HIP path)")
    config.addinivalue_line("markers", "slow: larger sizes, still minutes at most")


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


@pytest.fixture(scope="session")
def usx_mod():
    import usx
    return usx


@pytest.fixture(scope="session")
def port():
    from oracle import oracle as O
    return O.load("port")


@pytest.fixture(scope="session")
def O():
    from oracle import oracle as O
    return O
