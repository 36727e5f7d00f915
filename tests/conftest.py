import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "fluid-2d_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls libfs2d_hip.so")
    config.addinivalue_line("markers", "slow: longer CPU-side reference runs")


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    den = float(np.linalg.norm(b.astype(np.float64)))
    num = float(np.linalg.norm(a.astype(np.float64) - b.astype(np.float64)))
    return num / den if den > 0 else num


def describe(a: np.ndarray, b: np.ndarray) -> str:
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    idx = np.unravel_index(int(np.argmax(d)), d.shape)
    return (f"max|diff|={d.max():.3e} at {idx}, rel-L2={rel_l2(a, b):.3e}, "
            f"differing cells={int((a.view(np.uint32) != b.view(np.uint32)).sum())}")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    return oracle.Oracle()


@pytest.fixture(scope="session")
def reference_lib():
    import oracle

    if not oracle.reference_available():
        pytest.skip("compiled reference (oracle/_ref) not built here")
    return oracle.Reference()


def golden(name: str):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)
