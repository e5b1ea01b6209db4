import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "optimized-gpu-accelerator_amd"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import pdhg_amd  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def load_netlib(name):
    """Netlib fixture (tests/golden/netlib_<name>.npz) as a pdhg_amd.LP and its known optimum."""
    d = np.load(os.path.join(ROOT, "tests", "golden", f"netlib_{name}.npz"))
    lp = pdhg_amd.LP(int(d["m"]), int(d["n"]), d["start"], d["index"], d["value"], d["c"], d["l"], d["u"], d["L"],
                     d["U"], float(d["offset"]))
    return lp, float(d["optimum"])


def load_plan(name):
    """Planning-model fixture (tests/golden/<name>.npz, minimisation form) and
    the reference's published optimum in that form."""
    d = np.load(os.path.join(ROOT, "tests", "golden", f"{name}.npz"))
    lp = pdhg_amd.LP(int(d["m"]), int(d["n"]), d["start"], d["index"], d["value"], d["c"], d["l"], d["u"], d["L"],
                     d["U"], float(d["offset"]))
    return lp, float(d["optimum"])


@pytest.fixture(scope="session")
def gpu():
    """The HIP device must be there when GPU tests run: no silent fallback."""
    info = pdhg_amd.gpu_info()
    assert info["available"], f"GPU tests need an MI355X: {info}"
    return info
