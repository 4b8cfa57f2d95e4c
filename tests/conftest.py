import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libbcm_hip.so on the device)")


def load_golden(name):
    """Load a fixture written by tests/golden/make_golden.py."""
    from bcmtools_amd._lib import BLOCK_DTYPE
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    out = {k: z[k] for k in z.files}
    out["blocks"] = out["blocks"].view(BLOCK_DTYPE).copy()
    for k in ("n", "vc", "method", "iters", "inner"):
        if k in out:
            out[k] = int(out[k])
    out.setdefault("inner", 1)  # nLoopInner of the sor1 fixtures written before it was stored
    return out


@pytest.fixture
def golden():
    return load_golden
