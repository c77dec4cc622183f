import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libpourlab3d.so kernels)")
    config.addinivalue_line("markers", "slow: long-running CPU oracle check")


@pytest.fixture(scope="session")
def pins():
    with open(os.path.join(GOLDEN, "pins.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def recording():
    with open(os.path.join(GOLDEN, "pour3d_5mm_header.json")) as f:
        header = json.load(f)
    z = np.load(os.path.join(GOLDEN, "pour3d_5mm_frames.npz"))
    return header, {"origin": z["origin"], "frames": list(z["frames"]),
                    "positions": z["positions"]}


def quantise(x, y, z, q):
    """Recording quantisation (src/view3d/format.ts:58): Float32 position / quantum, JS rounding."""
    pos = np.stack([x, y, z], 1).astype(np.float32).astype(np.float64)
    return np.clip(np.floor(pos / q + 0.5), 0, 65535).astype(np.int64)
